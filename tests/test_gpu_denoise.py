"""The HIP U-Net (bf16 MFMA convolutions) against the plain fp32 torch
reference of the same network with the same bf16 rounding points: random
weights on an odd image size (edges of the 8x16 tiles, odd pooling sizes),
and the shipped weights on a rendered 4-spp film with its guides."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from vork_amd import denoise as D, native, scene as S  # noqa: E402

pytestmark = pytest.mark.gpu


def reference(blob_model, film, guides, samples):
    x = D.features(torch.from_numpy(film)[None], torch.from_numpy(guides)[None], samples)
    with torch.no_grad():
        y = D.radiance(blob_model(x, bf16=True))
    return y[0].permute(1, 2, 0).numpy()


def compare(out, ref):
    assert np.isfinite(out).all() and (out[..., 3] == 1).all()
    err = np.abs(out[..., :3] - ref) / (np.abs(ref) + 0.05)
    assert np.mean(err) < 4e-3 and np.mean(err < 0.03) > 0.999, (np.mean(err), np.mean(err < 0.03))


def test_random_weights_on_an_odd_size(tmp_path):
    torch.manual_seed(3)
    blob = D.export(D.make_unet())
    p = tmp_path / "w.vhdn"
    p.write_bytes(blob)
    model = D.load_model(str(p)).eval()
    rng = np.random.default_rng(4)
    h, w, spp = 75, 133, 4
    film = (rng.random((h, w, 4)) * 3 * spp).astype(np.float32)
    g = np.zeros((h, w, 2, 4), np.float32)
    g[..., 0, :3] = rng.random((h, w, 3)) * spp
    g[..., 0, 3] = spp
    n = rng.normal(size=(h, w, 3))
    g[..., 1, :3] = n / np.linalg.norm(n, axis=-1, keepdims=True) * spp
    g[..., 1, 3] = rng.random((h, w)) * spp
    r = native.Renderer(0)
    den = D.Denoiser(r.device, blob)
    compare(den(film, g, spp), reference(model, film, g, spp))
    den.close()


def test_shipped_weights_on_a_rendered_film():
    sc = S.config0(width=96, height=64)
    r = native.Renderer(0)
    inst = r.install(sc)
    film, g = inst.render_guides(sc.render_params(samples=4, seed=9, adaptive=False))
    den = D.Denoiser(r.device)
    out = den(film, g, 4)
    compare(out, reference(D.load_model().eval(), film, g, 4))
    # device pointers give the same image
    tf, tg = torch.from_numpy(film).cuda(), torch.from_numpy(g).cuda()
    assert np.array_equal(den(tf, tg, 4).cpu().numpy(), out)
    den.close()


def test_bad_weight_blobs_are_refused():
    r = native.Renderer(0)
    blob = D.export(D.make_unet())
    for bad in (b"nope", blob[:-4], blob + b"\0", blob[:12] + b"\x20\0\0\0\x21\0\0\0" + blob[20:]):
        with pytest.raises(native.VorkError) as e:
            D.Denoiser(r.device, bad)
        assert e.value.status == 2


def test_shipped_weights_at_1080p():
    """The denoiser at the size it serves, 1920x1080: the coarse levels take
    the channel split the kernel picks for this size.  Input: a 4-spp render
    with guides of a 1.8M-triangle GGX scene."""
    sc = S.config1(seed=2, width=1920, height=1080, sphere_level=8, terrain=512, materials=16)
    r = native.Renderer(0)
    inst = r.install(sc)
    film, g = inst.render_guides(sc.render_params(samples=4, seed=9, adaptive=False))
    den = D.Denoiser(r.device)
    tf, tg = torch.from_numpy(film).cuda(), torch.from_numpy(g).cuda()
    out = den(tf, tg, 4).cpu().numpy()
    torch.backends.cudnn.allow_tf32 = False
    torch.backends.cuda.matmul.allow_tf32 = False
    model = D.load_model().eval().cuda()
    with torch.no_grad():
        x = D.features(tf[None], tg[None], 4)
        ref = D.radiance(model(x, bf16=True))[0].permute(1, 2, 0).cpu().numpy()
    compare(out, ref)
    den.close()
