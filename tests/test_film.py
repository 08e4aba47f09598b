"""Film output: resolve, views, PNG / EXR writers and readers (CPU only)."""
import json
import os

import numpy as np
import pytest

from vork_amd import film as F

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_resolve_divides_by_the_sample_count():
    f = np.zeros((2, 3, 4), np.float32)
    f[..., :3] = 8.0
    f[..., 3] = 4.0
    f[1, 2] = 0
    r = F.resolve(f)
    assert np.allclose(r[0, 0], 2.0) and np.allclose(r[1, 2], 0.0)


def test_srgb_transfer_closed_form():
    # IEC 61966-2-1 values
    c = np.array([0.0, 0.0031308, 0.18, 0.5, 1.0])
    e = F.encode_srgb(c)
    assert e[0] == 0.0 and e[-1] == pytest.approx(1.0)
    assert e[1] == pytest.approx(12.92 * 0.0031308)
    assert e[2] == pytest.approx(1.055 * 0.18 ** (1 / 2.4) - 0.055)
    assert np.allclose(F.decode_srgb(e), c, atol=1e-12)
    # exposure +1 doubles the linear value; gamma acts on the codes
    assert np.allclose(F.view_standard(0.09, exposure=1.0), F.encode_srgb(0.18))
    assert np.allclose(F.view_standard(0.18, gamma=2.0), F.encode_srgb(0.18) ** 0.5)


def _synthetic_lut():
    n = F.AGX_N
    step = (F.AGX_MAX_LOG2 - F.AGX_MIN_LOG2) / (n - 1)
    axis = np.r_[0.0, 2.0 ** (F.AGX_MIN_LOG2 + step * np.arange(1, n))]
    b, g, r = np.meshgrid(axis, axis, axis, indexing="ij")
    return F.agx_analytic(np.stack([r, g, b], -1)), axis


def test_lut_view_returns_grid_samples_and_interpolates_between():
    lut, axis = _synthetic_lut()
    rng = np.random.default_rng(1)
    idx = rng.integers(0, F.AGX_N, size=(200, 3))
    x = axis[idx]
    got = F.agx_lut(x, lut)
    want = lut[idx[:, 2], idx[:, 1], idx[:, 0]]
    assert np.allclose(got, want, atol=1e-12)
    # off the grid: trilinear in the log index, close to the function sampled
    y = np.exp2(rng.uniform(-10, 10, size=(500, 3)))
    assert np.abs(F.agx_lut(y, lut) - F.agx_analytic(y)).max() < 0.02
    # exact black and the linear first interval
    assert np.allclose(F.agx_lut(np.zeros((1, 3)), lut), lut[0, 0, 0])


def test_exr_zip_round_trip_in_the_agx_library_layout(tmp_path):
    lut, _ = _synthetic_lut()
    n = F.AGX_N
    # library image: x = 65 b + r, y = 65 look + g; two looks
    band = lut.transpose(1, 0, 2, 3).reshape(n, n * n, 3)
    img = np.concatenate([band, band[::-1]], 0).astype(np.float16).astype(np.float32)
    path = os.path.join(tmp_path, "agx.exr")
    F.write_exr(path, np.concatenate([img, np.ones(img.shape[:2] + (1,), np.float32)], -1), half=True)
    back = F.load_agx_lut(path, look=0)
    assert np.array_equal(back, lut.astype(np.float16).astype(np.float64))


@pytest.mark.parametrize("half,compression", [(False, "none"), (False, "zip"), (True, "zip")])
def test_exr_round_trip(tmp_path, half, compression):
    rng = np.random.default_rng(3)
    rgba = rng.normal(size=(37, 29, 4)).astype(np.float32) * 10
    path = os.path.join(tmp_path, "x.exr")
    F.write_exr(path, rgba, half=half, compression=compression)
    ch = F.read_exr(path)
    got = np.stack([ch["R"], ch["G"], ch["B"], ch["A"]], -1)
    want = rgba.astype(np.float16).astype(np.float32) if half else rgba
    assert np.array_equal(got, want)


@pytest.mark.parametrize("depth", [8, 16])
def test_png_round_trip(tmp_path, depth):
    rng = np.random.default_rng(4)
    codes = rng.uniform(-0.1, 1.1, size=(21, 17, 3))
    path = os.path.join(tmp_path, "x.png")
    F.write_png(path, codes, depth)
    back = F.read_png(path)
    top = (1 << depth) - 1
    assert np.array_equal(back, np.floor(np.clip(codes, 0, 1) * top + 0.5).astype(back.dtype))


def test_analytic_agx_against_the_reference_lut_samples():
    """The analytic AgX is an approximation of the baked view: grey axis
    monotonic, black exact; deviation from the reference LUT at its grid
    nodes stated here (mean 0.018, max 0.21 code units on 481 samples)."""
    with open(os.path.join(GOLDEN, "agx_samples.json")) as f:
        d = json.load(f)
    x = np.array([s["linear"] for s in d["samples"]])
    y = np.array([s["display"] for s in d["samples"]])
    a = F.agx_analytic(x)
    err = np.abs(a - y)
    assert err.mean() < 0.025 and err.max() < 0.25
    grey = sorted((s["linear"][0], s["display"][0]) for s in d["samples"] if len(set(s["index"])) == 1)
    ref = np.array([g[1] for g in grey])
    assert (np.diff(ref) >= 0).all()
    mine = F.agx_analytic(np.array([[g[0]] * 3 for g in grey]))[:, 0]
    assert (np.diff(mine) >= -1e-9).all()
    assert mine[0] == 0.0 and ref[0] == 0.0


def test_write_outputs_and_meta(tmp_path):
    f = np.zeros((8, 10, 4), np.float32)
    f[..., :3] = np.linspace(0, 4, 80).reshape(8, 10, 1)
    f[..., 3] = 4
    png = os.path.join(tmp_path, "a.png")
    F.write_outputs(png, f, view="standard", meta={"version": 3, "warnings": []})
    codes = F.read_png(png)
    assert np.array_equal(codes, F.quantize(F.view_standard(F.resolve(f))))
    assert json.load(open(png + ".meta.json"))["version"] == 3
    exr = os.path.join(tmp_path, "a.exr")
    F.write_outputs(exr, f)
    ch = F.read_exr(exr)
    assert np.allclose(ch["R"], F.resolve(f)[..., 0])
