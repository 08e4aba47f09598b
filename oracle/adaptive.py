"""Adaptive sampling of the per-pixel entry, restated over per-sample values.

TEST INFRASTRUCTURE ONLY: the checker of the HIP renderer's adaptive mode,
loaded by tests/.  The product (vork_amd) never imports this package.

Follows crates/vork-render/shaders/path.slang:
  * 1796-1802  moments reset on the exposure's first sample; minimum and
               threshold defaults (minimum max(16, sqrt(samples)), 0.01),
  * 1804-1811  a stopped pixel outputs its mean,
  * 1812-1816  settled: every channel's sqrt(variance / max(i-1, 1) / i) at
               most threshold * max(|mean|, 0.01), tested from index `minimum`,
  * 1817-1836  the test is published in slot i & 1; the pixel stops when it
               and its clamped 3x3 neighbourhood passed slot (i & 1) ^ 1,
  * 1964-1966  Welford update of mean and variance with the sample value.
Every operation is a float32 operation in the shader's order, so the result is
bit-exact for the same per-sample values.
"""
import numpy as np

F = np.float32


def default_minimum(samples):
    """path.slang:1801 with PATH_ADAPTIVE_MIN_SAMPLES = 0 (constants.rs:266)."""
    return max(16, int(np.sqrt(np.float32(samples))))


def settled_test(mean, variance, index, threshold):
    """path.slang:1812-1816 for sample index >= 1; arrays (..., 4)."""
    a, b = F(max(index - 1, 1)), F(index)
    error = np.sqrt(np.maximum(variance / a / b, F(0)))
    return np.all(error <= F(threshold) * np.maximum(np.abs(mean), F(0.01)), axis=-1)


def render(values, minimum, threshold=0.01):
    """Adaptive exposure over `values`, float32 (S, H, W, 4): the RGBA value of
    sample index k of every pixel.  Returns (film, traced, mean, variance):
    the float32 film sum over the S samples, a bool (S, H, W) mask of the
    samples that traced a path, and the final moments."""
    values = np.asarray(values, dtype=F)
    S, H, W, _ = values.shape
    film = np.zeros((H, W, 4), F)
    mean = np.zeros((H, W, 4), F)
    variance = np.zeros((H, W, 4), F)
    settled = np.zeros((2, H, W), bool)
    stopped = np.zeros((H, W), bool)
    traced = np.zeros((S, H, W), bool)
    ys = np.clip(np.arange(H)[:, None] + np.array([-1, 0, 1])[None, :], 0, H - 1)
    xs = np.clip(np.arange(W)[:, None] + np.array([-1, 0, 1])[None, :], 0, W - 1)
    for k in range(S):
        if k == 0:
            mean[:] = 0
            variance[:] = 0
            stopped[:] = False
        run = ~stopped
        test = settled_test(mean, variance, k, threshold) if k >= minimum else np.zeros((H, W), bool)
        settled[k & 1][run] = test[run]
        prev = settled[(k & 1) ^ 1]
        neighbourhood = np.ones((H, W), bool)
        for dy in range(3):
            for dx in range(3):
                neighbourhood &= prev[ys[:, dy]][:, xs[:, dx]]
        stop = run & test & neighbourhood
        stopped |= stop
        run = ~stopped
        out = np.where(run[..., None], values[k], mean)
        film += out
        traced[k] = run
        v = values[k]
        delta = v - mean
        new_mean = mean + delta / F(k + 1)
        new_variance = variance + delta * (v - new_mean)
        mean = np.where(run[..., None], new_mean, mean)
        variance = np.where(run[..., None], new_variance, variance)
    return film, traced, mean, variance
