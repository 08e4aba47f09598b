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


def _batch(rng, n, h, w, spp):
    film = (rng.random((n, h, w, 4)) * 2 * spp).astype(np.float32)
    g = np.zeros((n, h, w, 2, 4), np.float32)
    g[..., 0, :3] = rng.random((n, h, w, 3)) * spp
    g[..., 0, 3] = spp
    nrm = rng.normal(size=(n, h, w, 3))
    g[..., 1, :3] = nrm / np.linalg.norm(nrm, axis=-1, keepdims=True) * spp
    g[..., 1, 3] = rng.random((n, h, w)) * spp
    target = (rng.random((n, h, w, 3)) * 1.5).astype(np.float32)
    return film, g, target


def test_training_gradients_match_torch_autograd():
    """The hand-written backward pass (k_relu_grad, k_wgrad, the transposed
    k_conv3x3, k_route, k_loss) against torch fp32 autograd of the same network
    with the same bf16 rounding points, on a batch of two odd-sized images:
    every layer's weight and bias gradient agrees to bf16 precision."""
    torch.manual_seed(7)
    blob = D.export(D.make_unet())
    rng = np.random.default_rng(3)
    n, h, w, spp = 2, 37, 45, 4
    film, g, target = _batch(rng, n, h, w, spp)
    r = native.Renderer(0)
    tr = D.Trainer(r.device, blob)
    tf, tg, tt = (torch.from_numpy(a).cuda() for a in (film, g, target))
    loss = tr.step(tf, tg, tt, spp, 0.0, update=False)
    grads = D.split_gradients(*tr.gradients())
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "w.vhdn")
        open(p, "wb").write(blob)
        model = D.load_model(p)
    x = D.features(torch.from_numpy(film), torch.from_numpy(g), spp)
    pred = model(x, bf16=True)
    ref_loss = D.loss_fn(pred, torch.from_numpy(target).permute(0, 3, 1, 2))
    ref_loss.backward()
    assert abs(loss - ref_loss.item()) <= 2e-3 * abs(ref_loss.item()), (loss, ref_loss.item())
    for k, (m, (gw, gb)) in enumerate(zip(model.convs(), grads)):
        co, ci = m.weight.shape[:2]
        want_w, want_b = m.weight.grad.numpy(), m.bias.grad.numpy()
        got_w, got_b = gw[:co, :ci], gb[:co]
        ew = np.linalg.norm(got_w - want_w) / np.linalg.norm(want_w)
        eb = np.linalg.norm(got_b - want_b) / (np.linalg.norm(want_b) + 1e-12)
        assert ew < 3e-2 and eb < 3e-2, (k, ew, eb)
        assert np.abs(gw[co:]).max(initial=0) == 0 and np.abs(gw[:, ci:]).max(initial=0) == 0, k
    tr.close()


def test_training_follows_torch_adam_and_exports_a_blob_the_denoiser_reads():
    """Adam on the device (k_adam over fp32 master weights, bf16 copies for the
    kernels) follows torch's Adam from the same starting weights on the same
    batch over its first steps, to bf16 precision; the loss falls; the exported blob
    loads into vh_denoiser_create and denoises like the torch reference of the
    same weights."""
    rng = np.random.default_rng(5)
    film, g, _ = _batch(rng, 2, 32, 48, 4)
    target = np.ascontiguousarray(film[..., :3] / 4)
    torch.manual_seed(3)
    start = D.make_unet()
    blob = D.export(start)
    import tempfile, os
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "w.vhdn")
        open(p, "wb").write(blob)
        model = D.load_model(p)
    opt = torch.optim.Adam(model.parameters(), lr=1e-3, eps=1e-8)
    x = D.features(torch.from_numpy(film), torch.from_numpy(g), 4)
    t = torch.from_numpy(target).permute(0, 3, 1, 2)
    r = native.Renderer(0)
    tr = D.Trainer(r.device, blob)
    tf, tg, tt = (torch.from_numpy(a).cuda() for a in (film, g, target))
    ours, theirs = [], []
    for _ in range(30):
        ours.append(tr.step(tf, tg, tt, 4, 1e-3))
        loss = D.loss_fn(model(x, bf16=True), t)
        opt.zero_grad()
        loss.backward()
        opt.step()
        theirs.append(loss.item())
    print("losses (hip, torch):", [(round(a, 4), round(b, 4)) for a, b in zip(ours[::5], theirs[::5])])
    assert ours[-1] < 0.85 * ours[0], ours[::5]
    # the same trajectory while rounding differences stay small (later steps
    # of either can amplify them)
    for a, b in zip(ours[:10], theirs[:10]):
        assert abs(a - b) <= 0.02 * b, (ours, theirs)
    blob = tr.export()
    den = D.Denoiser(r.device, blob)
    out = den(film[0], g[0], 4)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "w.vhdn")
        open(p, "wb").write(blob)
        compare(out, reference(D.load_model(p).eval(), film[0], g[0], 4))
    den.close()
    tr.close()
