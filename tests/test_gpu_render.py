"""GPU parity of the wavefront integrator against the CPU oracle.

Both restate the reference estimator with the same counter-based sampler, so
the same (pixel, sample) follows the same path up to float rounding (FMA
contraction and the GPU's sin/cos/log2 differ in the last bits).  The bars:
  * per-image relMSE of GPU vs oracle at equal spp far below the noise level,
  * relMSE against a 4096-spp oracle reference equal within 2 %,
  * bit-identical re-renders for a fixed seed, and a film that is linear in
    the sample range (two half ranges sum to the full range).
"""
import numpy as np
import pytest

import oracle
from vork_amd import scene as S
from tests.helpers import rel_mse

pytestmark = pytest.mark.gpu


def _render_pair(renderer, sc, spp, seed=1):
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    p = sc.render_params(samples=spp, seed=seed)
    g = inst.render(p)
    o32, o64 = ora.render(p)
    return inst, ora, g, o32, o64, p


def test_config0_matches_oracle_at_equal_spp(renderer):
    sc = S.config0(width=64, height=64)
    inst, ora, g, o32, o64, p = _render_pair(renderer, sc, 16)
    assert np.all(np.isfinite(g))
    assert np.array_equal(g[..., 3], np.full(g.shape[:2], 16.0, np.float32))
    gm, om = g[..., :3] / 16, o64[..., :3] / 16
    noise = rel_mse(om, ora.render(sc.render_params(samples=16, seed=77))[1][..., :3] / 16)
    diff = rel_mse(gm, om)
    assert diff < 0.02 * noise, (diff, noise)
    close = np.abs(gm - om) <= 1e-3 * np.maximum(om, 1e-2)
    assert close.mean() > 0.97, close.mean()
    st = inst.stats()
    assert st["paths"] == 64 * 64 * 16
    assert st["dropped_samples"] == 0


def test_config0_relmse_vs_reference_within_two_percent(renderer):
    sc = S.config0(width=48, height=48)
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    # the 4096-spp FP64-accumulated oracle reference
    ref_p = sc.render_params(samples=4096, seed=1000, adaptive=False)
    ref = ora.render(ref_p, film32=False)[1][..., :3] / 4096
    errs_g, errs_o = [], []
    for seed in range(1, 5):
        p = sc.render_params(samples=16, seed=seed)
        errs_g.append(rel_mse(inst.render(p)[..., :3] / 16, ref))
        errs_o.append(rel_mse(ora.render(p, film32=False)[1][..., :3] / 16, ref))
    eg, eo = np.mean(errs_g), np.mean(errs_o)
    assert abs(eg - eo) <= 0.02 * eo, (eg, eo)


def test_ggx_scene_matches_oracle(renderer):
    sc = S.small_config1(width=64, height=36, spp=8)
    inst, ora, g, o32, o64, p = _render_pair(renderer, sc, 8)
    gm, om = g[..., :3] / 8, o64[..., :3] / 8
    noise = rel_mse(om, ora.render(sc.render_params(samples=8, seed=91))[1][..., :3] / 8)
    diff = rel_mse(gm, om)
    assert diff < 0.05 * noise, (diff, noise)


def test_seed_determinism_and_sample_range_linearity(renderer):
    sc = S.config0(width=40, height=40)
    inst = renderer.install(sc)
    p = sc.render_params(samples=16, seed=3)
    a = inst.render(p)
    b = inst.render(p)
    assert np.array_equal(a, b)
    p1 = sc.render_params(samples=16, seed=3)
    p1.sample_end = 8
    p2 = sc.render_params(samples=16, seed=3)
    p2.sample_begin = 8
    f = inst.render(p1)
    f = inst.render(p2, film=f)
    np.testing.assert_allclose(f, a, rtol=2e-6, atol=1e-6)
    p3 = sc.render_params(samples=16, seed=3)
    p3.batch_samples = 3
    np.testing.assert_allclose(inst.render(p3), a, rtol=2e-6, atol=1e-6)


def test_batch_layouts_give_the_same_film(renderer):
    # a batch of B samples keeps the largest power of two dividing B of a
    # pixel's samples in consecutive path slots (1, 2 and 16 here; 16 also
    # takes the film kernel's staged path); every layout sums a pixel's samples
    # in sample order, so the films are bit-identical
    sc = S.config0(width=40, height=24)
    inst = renderer.install(sc)
    films = []
    for batch in (1, 2, 6, 16, 0):
        p = sc.render_params(samples=16, seed=11, adaptive=False)
        p.batch_samples = batch
        films.append(inst.render(p))
    for f in films[1:]:
        assert np.array_equal(f, films[0])
    assert films[0][..., 3].min() > 0


def test_device_film_accumulates(renderer):
    torch = pytest.importorskip("torch")
    sc = S.config0(width=32, height=32)
    inst = renderer.install(sc)
    p = sc.render_params(samples=4, seed=5)
    host = inst.render(p)
    dev = torch.zeros((32, 32, 4), dtype=torch.float32, device="cuda")
    inst.render(p, device_film=dev)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dev.cpu().numpy(), host)


def test_instanced_flock_matches_oracle(renderer):
    """Config 2 at test size: thousands of instances of one BLAS, TLAS rebuilt
    for a later animation frame through the refit entry point."""
    sc = S.config2(width=64, height=36, spp=4, birds=3000)
    sc.camera = dict(eye=(0.0, 24.0, 60.0), target=(0.0, 21.0, 0.0))
    inst, ora, g, o32, o64, p = _render_pair(renderer, sc, 4)
    gm, om = g[..., :3] / 4, o64[..., :3] / 4
    noise = rel_mse(om, ora.render(sc.render_params(samples=4, seed=57))[1][..., :3] / 4)
    assert rel_mse(gm, om) < 0.05 * noise
    # a later frame, through VH_TLAS_REFIT on the same handle
    frames = S.flock_frames(sc, 3)
    sc2 = S.Scene(sc.meshes, frames[2], sc.materials, sc.lights, sc.camera, sc.sky, sc.width, sc.height, sc.spp,
                  sc.max_depth, sc.seed, "config2")
    inst.tlas.build(sc2.instance_array(), refit=True)
    g2 = inst.render(p)
    o2 = oracle.OracleScene(sc2).render(p)[1]
    assert rel_mse(g2[..., :3] / 4, o2[..., :3] / 4) < 0.05 * noise
    assert not np.array_equal(g2, g)


def test_scene_json_scenes_render_like_the_oracle(renderer, tmp_path):
    """Scenes read by the scene.json loader: the reference's analytic fixture
    (perspective camera, rectangle key light) and config 0 written and read
    back."""
    import os
    from vork_amd import scene_json as J
    from tests.test_scene_json import GOLDEN, analytic_perspective
    sc, _ = J.load(analytic_perspective(), J.parse_set(["samples=8", "seed=3"]), base_dir=GOLDEN)
    inst, ora, g, o32, o64, p = _render_pair(renderer, sc, 8, seed=3)
    noise = rel_mse(o64[..., :3] / 8, ora.render(sc.render_params(samples=8, seed=44))[1][..., :3] / 8)
    assert rel_mse(g[..., :3] / 8, o64[..., :3] / 8) < 0.05 * noise
    assert g[..., :3].mean() > 0.01
    path = os.path.join(tmp_path, "c0.json")
    J.dump(S.config0(width=48, height=48), path)
    sc0, _ = J.load(path, {"samples": 8, "seed": 1})
    inst, ora, g, o32, o64, p = _render_pair(renderer, sc0, 8)
    noise = rel_mse(o64[..., :3] / 8, ora.render(sc0.render_params(samples=8, seed=45))[1][..., :3] / 8)
    assert rel_mse(g[..., :3] / 8, o64[..., :3] / 8) < 0.05 * noise


def test_render_call_writes_png_exr_and_meta(tmp_path):
    """`python -m vork_amd.render SCENE --out FILE --set ...` end to end: the
    EXR holds the film mean, the PNG its standard-view codes."""
    import json
    import os
    from vork_amd import film as F, render as R, scene_json as J
    scene = os.path.join(tmp_path, "c0.json")
    J.dump(S.config0(width=32, height=32), scene)
    exr = os.path.join(tmp_path, "c0.exr")
    png = os.path.join(tmp_path, "c0.png")
    assert R.main([scene, "--out", exr, "--set", "samples=4", "--set", "seed=2"]) == 0
    assert R.main([scene, "--out", png, "--view", "standard", "--set", "samples=4", "--set", "seed=2"]) == 0
    ch = F.read_exr(exr)
    rgb = np.stack([ch["R"], ch["G"], ch["B"]], -1)
    assert rgb.shape == (32, 32, 3) and np.isfinite(rgb).all() and rgb.mean() > 0
    codes = F.read_png(png).astype(np.int32)
    want = F.quantize(F.view_standard(rgb.astype(np.float64))).astype(np.int32)
    # the EXR holds the mean rounded to float32: a code may sit one step away
    assert np.abs(codes - want).max() <= 1 and (codes == want).mean() > 0.99
    meta = json.load(open(png + ".meta.json"))
    assert meta["version"] == 3 and meta["render_settings"]["samples"] == 4
    assert meta["render_settings"]["resolution"] == [32, 32]
    assert R.main([scene, "--out", png, "--set", "bogus=1"]) == 2


def _per_sample_values(inst, sc, samples, seed):
    """Each sample index rendered alone into a zero film: its RGBA value."""
    out = []
    for k in range(samples):
        p = sc.render_params(samples=samples, seed=seed, adaptive=False)
        p.sample_begin, p.sample_end = k, k + 1
        out.append(inst.render(p))
    return np.stack(out)


def test_adaptive_film_equals_the_restatement_over_the_gpu_samples(renderer):
    from oracle import adaptive
    sc = S.config0(width=48, height=40)
    inst = renderer.install(sc)
    spp, seed = 40, 4
    values = _per_sample_values(inst, sc, spp, seed)
    for threshold in (0.05, 0.2):
        p = sc.render_params(samples=spp, seed=seed)
        p.adaptive_threshold = threshold
        g = inst.render(p)
        film, traced, _, _ = adaptive.render(values, adaptive.default_minimum(spp), threshold)
        st = inst.stats()
        assert 0 < traced[-1].sum() < traced[0].sum(), traced.sum(axis=(1, 2))
        assert st["paths"] == traced.sum()
        assert st["stopped_samples"] == traced.size - traced.sum()
        assert np.array_equal(g, film)
    # ranges continue the exposure: [0, 25) then [25, 40) == [0, 40)
    p1 = sc.render_params(samples=spp, seed=seed)
    p1.adaptive_threshold, p1.sample_end = 0.05, 25
    p2 = sc.render_params(samples=spp, seed=seed)
    p2.adaptive_threshold, p2.sample_begin = 0.05, 25
    whole = sc.render_params(samples=spp, seed=seed)
    whole.adaptive_threshold = 0.05
    a = inst.render(whole)
    f = inst.render(p2, film=inst.render(p1))
    assert np.array_equal(f, a)


def test_adaptive_at_a_tiny_threshold_is_the_plain_render(renderer):
    # only pixels whose samples are all equal (the sky) can settle; they add
    # their exact mean, so the film is the plain one
    from oracle import adaptive
    sc = S.config0(width=40, height=32)
    inst = renderer.install(sc)
    plain = inst.render(sc.render_params(samples=24, seed=9, adaptive=False))
    p = sc.render_params(samples=24, seed=9)
    p.adaptive_threshold, p.adaptive_min_samples = 1e-30, 4
    assert np.array_equal(inst.render(p), plain)
    stopped = inst.stats()["stopped_samples"]
    _, traced, _, _ = adaptive.render(_per_sample_values(inst, sc, 24, 9), 4, 1e-30)
    assert stopped == traced.size - traced.sum()


def test_adaptive_range_must_continue_its_exposure(renderer):
    from vork_amd.native import VorkError
    sc = S.config0(width=16, height=16)
    inst = renderer.install(sc)
    p = sc.render_params(samples=32, seed=2)
    p.sample_begin = 20
    with pytest.raises(VorkError) as e:
        inst.render(p)
    assert e.value.status == 2
    q = sc.render_params(samples=32, seed=2)
    q.sample_end = 20
    inst.render(q)
    r = sc.render_params(samples=32, seed=3)  # another seed is another exposure
    r.sample_begin = 20
    with pytest.raises(VorkError):
        inst.render(r)


def test_objects_keep_their_device_open_in_any_release_order():
    # the device handle released first: the scene still renders, and the
    # objects release in dependency-free order afterwards
    from vork_amd import native
    sc = S.config0(width=24, height=16, spp=2)
    r = native.Renderer(0)
    inst = r.install(sc)
    want = inst.render(sc.render_params(adaptive=False))
    r.device.close()
    got = inst.render(sc.render_params(adaptive=False))
    assert np.array_equal(got, want)
    for b in inst.blases:
        b.close()
    inst.tlas.close()
    inst.close()


@pytest.mark.parametrize("make,spp", [(lambda: S.config0(width=48, height=40), 4),
                                      (lambda: S.small_config1(width=64, height=36), 3)])
def test_denoiser_guides_match_the_oracle(renderer, make, spp):
    sc = make()
    inst = renderer.install(sc)
    p = sc.render_params(samples=spp, seed=5, adaptive=False)
    film, g = inst.render_guides(p)
    assert np.array_equal(film, inst.render(p))     # guides leave the film as it is
    _, go = oracle.OracleScene(sc).render_guides(p)
    assert np.array_equal(g[..., 0, 3], go[..., 0, 3])   # the same camera hits
    close = np.abs(g - go) <= 1e-4 + 1e-4 * np.abs(go)
    assert close.mean() > 0.999, close.mean()


def test_graph_replays_equal_direct_launches(renderer):
    """Every batch replays the stream's recorded graph of its launch sequence
    (the kernels read the batch from a device-side launch block); with
    VH_RENDER_PROFILE the same kernels are launched one by one instead.  The
    films agree bit for bit across batch sizes, adaptive batches (a decide
    pass before each late sample), two scenes alternating on one stream, and
    replays separated by other device work."""
    torch = pytest.importorskip("torch")
    from vork_amd import abi
    scenes = [renderer.install(S.config0(width=40, height=32)), renderer.install(S.small_config1(width=48, height=27))]
    for rep in range(3):
        for k, inst in enumerate(scenes):
            for batch, adaptive in ((0, False), (1, False), (3, False), (0, True)):
                p = inst.scene.render_params(samples=20, seed=5 + k, adaptive=adaptive)
                p.batch_samples = batch
                if adaptive:
                    p.adaptive_threshold, p.adaptive_min_samples = 0.1, 6
                graph = inst.render(p)
                st_graph = inst.stats()
                p.flags |= abi.VH_RENDER_PROFILE
                direct = inst.render(p)
                assert np.array_equal(graph, direct), (rep, k, batch, adaptive)
                st = inst.stats()
                assert st["paths"] == st_graph["paths"] and st["extension_rays"] == st_graph["extension_rays"]
                assert st["extend_launches"] > 0 and st["extend_camera_launches"] >= 1
            x = torch.randn(1 << 20, device="cuda")
            torch.cuda.synchronize()
            assert float((x * 0).sum()) == 0.0


@pytest.mark.parametrize("make", [lambda: S.config0(width=48, height=40),
                                  lambda: S.small_config1(width=64, height=36),
                                  lambda: S.config2(width=48, height=27, spp=4, birds=3000)])
def test_material_sorted_shading_gives_the_same_film(renderer, make):
    """Hits sorted into per-closure queues (diffuse, dielectric GGX, metal
    GGX; misses shaded by the classify pass) shade every path exactly as the
    arrival-order queue does: the films are bit-identical, with guides, with
    several batches and adaptively."""
    from vork_amd import abi
    sc = make()
    inst = renderer.install(sc)
    for batch, adaptive in ((0, False), (2, False), (0, True)):
        films = []
        for order in (abi.VH_RENDER_SHADE_ARRIVAL, abi.VH_RENDER_SHADE_SORTED):
            p = sc.render_params(samples=12, seed=4, adaptive=adaptive)
            p.batch_samples = batch
            p.flags |= order
            if adaptive:
                p.adaptive_threshold, p.adaptive_min_samples = 0.1, 4
            films.append(inst.render_guides(p))
            stats = inst.stats()
            assert stats["dropped_samples"] == 0 or batch != 0
        assert np.array_equal(films[0][0], films[1][0]), (batch, adaptive)
        assert np.array_equal(films[0][1], films[1][1]), (batch, adaptive)


def test_shared_wavefront_queue_films_equal_each_job_alone(renderer):
    """vh_render_shared: jobs of different scenes, sizes, sample counts, sample
    ranges and bounce limits with their paths in one wavefront (one extension,
    one shading and one shadow queue, each path traversing its own job's TLAS)
    give each job exactly the film and guides vh_render_guides gives it alone.
    The 1080p x 64-spp job does not fit one batch beside the others, so the
    call runs several batches."""
    torch = pytest.importorskip("torch")
    from vork_amd import native
    specs = [(S.config0(width=64, height=48), 16, 0, 16, 4, True),
             (S.small_config1(width=96, height=54), 8, 0, 8, 8, False),
             (S.config2(width=48, height=27, spp=4, birds=3000), 4, 0, 4, 3, True),
             (S.config0(seed=2, width=32, height=32), 16, 3, 11, 4, False),
             (S.small_config1(seed=3, width=1920, height=1080), 64, 0, 64, 8, False)]
    jobs, alone = [], []
    for sc, spp, b, e, depth, guides in specs:
        inst = renderer.install(sc)
        p = sc.render_params(samples=spp, seed=7, adaptive=False)
        p.sample_begin, p.sample_end, p.max_bounces = b, e, depth
        film = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
        g = torch.zeros((sc.height, sc.width, 2, 4), dtype=torch.float32, device="cuda") if guides else None
        jobs.append((inst, p, film, g))
        f1 = torch.zeros_like(film)
        if guides:
            g1 = torch.zeros_like(g)
            inst.render_guides(p, f1, g1, device=True)
        else:
            g1 = None
            inst.render(p, device_film=f1)
        alone.append((f1, g1))
    native.render_shared(jobs)
    torch.cuda.synchronize()
    for (inst, p, film, g), (f1, g1) in zip(jobs, alone):
        assert torch.equal(film, f1), (p.width, p.height)
        if g is not None:
            assert torch.equal(g, g1)
        assert inst.stats()["paths"] == p.width * p.height * (p.sample_end - p.sample_begin)
    assert float(jobs[-1][2][..., 3].min()) == 64.0


def test_one_scene_renders_on_two_streams_at_once(renderer):
    """A scene is not written by rendering it: two threads, each with its own
    stream, render one scene at once, each an adaptive exposure split into
    sample ranges (the moments of an exposure in progress belong to the scene
    and stream), and a shared call may name the scene twice.  Every film
    equals the same render made alone."""
    import threading
    torch = pytest.importorskip("torch")
    from vork_amd import native
    sc = S.config0(width=64, height=48)
    inst = renderer.install(sc)

    def exposure(seed, cut, stream=None):
        p1 = sc.render_params(samples=32, seed=seed)
        p1.adaptive_threshold, p1.sample_end = 0.05, cut
        p2 = sc.render_params(samples=32, seed=seed)
        p2.adaptive_threshold, p2.sample_begin = 0.05, cut
        return inst.render(p2, film=inst.render(p1, stream=stream), stream=stream)

    alone = {seed: exposure(seed, 12) for seed in (3, 4)}
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    got, errors = {}, []

    def worker(k, seed):
        try:
            for rep in range(4):
                got[(seed, rep)] = exposure(seed, 12, streams[k].cuda_stream)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k, seed)) for k, seed in enumerate((3, 4))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for (seed, rep), f in got.items():
        assert np.array_equal(f, alone[seed]), (seed, rep)
    # one scene in two jobs of one shared wavefront
    plain = sc.render_params(samples=8, seed=6, adaptive=False)
    films = [torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda") for _ in range(2)]
    native.render_shared([(inst, plain, films[0], None), (inst, plain, films[1], None)])
    torch.cuda.synchronize()
    ref = inst.render(plain)
    assert np.array_equal(films[0].cpu().numpy(), ref) and np.array_equal(films[1].cpu().numpy(), ref)
