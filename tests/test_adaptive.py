"""Pins of the adaptive-sampling restatement (oracle/adaptive.py) by closed
forms: constant pixels stop right after `minimum`, one noisy pixel holds its
clamped 3x3 neighbourhood, the Welford moments equal the two-pass mean and
squared deviations, and the first stop index of a known sequence follows from
the direct standard-error formula."""
import numpy as np

from oracle import adaptive

F = np.float32


def constant_values(S, H, W, value=0.5):
    v = np.full((S, H, W, 4), value, F)
    v[..., 3] = 1.0
    return v


def test_default_minimum():
    assert adaptive.default_minimum(16) == 16
    assert adaptive.default_minimum(64) == 16
    assert adaptive.default_minimum(1024) == 32
    assert adaptive.default_minimum(4096) == 64


def test_constant_pixels_stop_after_minimum():
    S, H, W, m = 40, 6, 7, 16
    film, traced, mean, var = adaptive.render(constant_values(S, H, W), m)
    # zero variance settles at index m; with every neighbour settled at m the
    # pixel stops at m + 1, having traced samples 0..m
    assert traced[: m + 1].all() and not traced[m + 1:].any()
    assert np.array_equal(film[..., :3], np.full((H, W, 3), S * 0.5, F))
    assert np.array_equal(film[..., 3], np.full((H, W), float(S), F))
    assert np.array_equal(mean[..., 0], np.full((H, W), 0.5, F)) and not var.any()


def noisy_at(S, H, W, y, x):
    v = constant_values(S, H, W)
    v[:, y, x, :3] = np.where(np.arange(S) % 2 == 0, 0.0, 1.0)[:, None]
    return v


def test_noisy_pixel_holds_its_neighbourhood():
    S, H, W, m = 40, 8, 9, 16
    _, traced, _, _ = adaptive.render(noisy_at(S, H, W, 3, 4), m)
    runs_all = traced.all(axis=0)
    want = np.zeros((H, W), bool)
    want[2:5, 3:6] = True
    assert np.array_equal(runs_all, want)
    assert traced.sum() == 9 * S + (H * W - 9) * (m + 1)


def test_noisy_corner_clamps_the_neighbourhood():
    S, H, W, m = 30, 5, 5, 16
    _, traced, _, _ = adaptive.render(noisy_at(S, H, W, 0, 0), m)
    want = np.zeros((H, W), bool)
    want[:2, :2] = True
    assert np.array_equal(traced.all(axis=0), want)


def test_tiny_threshold_is_the_plain_film():
    rng = np.random.default_rng(3)
    S, H, W = 24, 4, 5
    v = rng.random((S, H, W, 4), dtype=F)
    film, traced, _, _ = adaptive.render(v, 4, threshold=1e-30)
    assert traced.all()
    plain = np.zeros((H, W, 4), F)
    for k in range(S):
        plain += v[k]
    assert np.array_equal(film, plain)


def test_welford_moments_equal_two_pass_sums():
    rng = np.random.default_rng(5)
    S = 200
    v = rng.random((S, 1, 1, 4)).astype(F) * 3
    _, _, mean, var = adaptive.render(v, S)  # never tests, so never stops
    x = v[:, 0, 0].astype(np.float64)
    np.testing.assert_allclose(mean[0, 0], x.mean(axis=0), rtol=1e-5)
    np.testing.assert_allclose(var[0, 0], ((x - x.mean(axis=0)) ** 2).sum(axis=0), rtol=1e-4)


def test_first_stop_follows_the_standard_error():
    # every pixel alternates 1, 3: after k samples the standard error is
    # sqrt(M2 / (k - 1) / k) with M2 the squared deviations of the first k
    S, t, m = 120, 0.095, 4
    seq = np.where(np.arange(S) % 2 == 0, 1.0, 3.0)

    def settled(k):
        x = seq[:k]
        err = np.sqrt(((x - x.mean()) ** 2).sum() / max(k - 1, 1) / k)
        assert abs(err - t * x.mean()) > 1e-4 * x.mean()  # no borderline case
        return k >= m and err <= t * abs(x.mean())

    first = next(k for k in range(m + 1, S) if settled(k) and settled(k - 1))
    v = np.zeros((S, 3, 3, 4), F)
    v[..., :3] = seq[:, None, None, None]
    v[..., 3] = 1
    _, traced, _, _ = adaptive.render(v, m, threshold=t)
    assert traced[:first].all() and not traced[first:].any(), first
