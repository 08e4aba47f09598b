"""Parity at the workloads' own sizes.

Config 1 (the benchmarked scene): both BLASes of the 7,340,032-triangle scene
built on the GPU from device-resident vertex data, two at once on two streams,
exactly as bench.py builds them every step.  A build this large runs dozens of
clustering rounds with periodic host read-backs, the single-block tail and a
deep collapse, which the small scenes of test_gpu_trace.py never reach.
  * 1,000,000 rays (camera, surface-leaving and light-sample shadow rays)
    traced against it give bit-identical (instance, primitive, t, u, v) and
    occlusion to the oracle's BVH over the same triangles;
  * the 1920x1080 x 64-spp film of the benchmarked call (auto batch: one
    wavefront batch of all 64 samples, camera rays as packets) matches the
    oracle's 64-sample sums on 8,192 random pixels, and single-sample renders
    match the oracle's per-(pixel, sample) values on 8,192 random pairs, far
    below the noise level.
Config 2: 1,000,001 instances (a million birds and the ground) and 1,000,000
rays against the oracle, bit-identical, with the traversal-stack depth of
every ray read back to show that the global-memory part of the per-lane stack
(entries beyond the 12 kept in LDS) is reached and still gives exact hits.

The oracle restates the reference's FP32 watertight test
(crates/vork-render/shaders/vork_medium_domain.slang:44-66,110-118) and its
estimator (path.slang:1313-1969); the hit rule is (t, instance, primitive).
"""
import math

import numpy as np
import pytest

import oracle
from vork_amd import abi, native, scene as S
from tests.helpers import rel_mse
from tests.test_gpu_trace import _compare

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _install_on_streams(renderer, sc):
    """BLASes from device-resident vertex data, each from its own host thread
    on its own stream (the benchmark's build), then the TLAS and scene."""
    import concurrent.futures
    dev = renderer.device
    gm = []
    for m in sc.meshes:
        mats = m.triangle_materials if m.triangle_materials is not None else np.zeros(m.triangle_count, np.uint32)
        gm.append((torch.from_numpy(np.ascontiguousarray(m.positions)).cuda(),
                   torch.from_numpy(np.ascontiguousarray(m.indices).view(np.int32)).cuda(),
                   torch.from_numpy(np.ascontiguousarray(m.normals)).cuda(),
                   torch.from_numpy(np.ascontiguousarray(mats).view(np.int32)).cuda()))
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream() for _ in gm]
    with concurrent.futures.ThreadPoolExecutor(len(gm)) as pool:
        blases = list(pool.map(lambda k: native.Blas(dev, *gm[k], device_pointers=True,
                                                     stream=streams[k].cuda_stream), range(len(gm))))
    tlas = native.Tlas(dev, blases, sc.instance_array())
    return native.InstalledScene(dev, sc, blases, tlas)


@pytest.fixture(scope="module")
def config1(renderer):
    sc = S.config1()
    assert sc.triangle_count == 7_340_032
    inst = _install_on_streams(renderer, sc)
    ora = oracle.OracleScene(sc)
    yield sc, inst, ora
    inst.close()
    ora.close()


def _surface_rays(rng, prim, hits, lights, n_bounce, n_shadow):
    """Rays leaving the surfaces the primary rays hit: bounce rays in random
    and grazing directions from the hit point (some nudged back along the
    incoming ray, some bounded), and shadow rays from the hit point to random
    points on the lights, bounded short of the light."""
    ok = hits["t"] > 0
    prim, hits = prim[ok], hits[ok]
    p = prim[:, :3].astype(np.float64) + prim[:, 4:7].astype(np.float64) * hits["t"][:, None].astype(np.float64)
    k = rng.integers(0, p.shape[0], n_bounce)
    pb, din = p[k], prim[k, 4:7].astype(np.float64)
    d = rng.normal(size=pb.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    g = np.cross(din, rng.normal(size=pb.shape))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    graze = rng.uniform(size=n_bounce) < 0.3
    d[graze] = g[graze]
    nudge = (rng.uniform(size=(n_bounce, 1)) < 0.5) * rng.choice([1e-6, 1e-4, 1e-3], size=(n_bounce, 1))
    bounce = np.zeros((n_bounce, 8), np.float32)
    bounce[:, :3], bounce[:, 4:7], bounce[:, 7] = pb - din * nudge, d, 1e30
    bounded = rng.uniform(size=n_bounce) < 0.3
    bounce[bounded, 7] = rng.uniform(0.01, 8.0, size=bounded.sum())
    k = rng.integers(0, p.shape[0], n_shadow)
    ps = p[k]
    li = rng.integers(0, len(lights), n_shadow)
    c = np.array([l["position"] for l in lights], np.float64)[li]
    u = np.array([l["u"] for l in lights], np.float64)[li]
    v = np.array([l["v"] for l in lights], np.float64)[li]
    target = c + u * rng.uniform(-0.5, 0.5, (n_shadow, 1)) + v * rng.uniform(-0.5, 0.5, (n_shadow, 1))
    d = target - ps
    dist = np.linalg.norm(d, axis=1)
    shadow = np.zeros((n_shadow, 8), np.float32)
    shadow[:, :3], shadow[:, 4:7] = ps, d / dist[:, None]
    shadow[:, 7] = dist * (1 - 1e-4)
    return bounce, shadow


def test_config1_million_rays_bit_identical(config1):
    sc, inst, ora = config1
    rng = np.random.default_rng(101)
    p = sc.render_params(adaptive=False)
    n_cam = 300_000
    pix = rng.integers(0, sc.width * sc.height, n_cam)
    cam = ora.camera_rays(p, pix, rng.integers(0, sc.spp, n_cam))
    oh = ora.trace_closest(cam, abi.VH_RAY_CAMERA)
    _compare(inst.tlas.trace_closest(cam, abi.VH_RAY_CAMERA), oh)
    assert (oh["t"] > 0).mean() > 0.5
    bounce, shadow = _surface_rays(rng, cam, oh, sc.lights, 400_000, 300_000)
    _compare(inst.tlas.trace_closest(bounce), ora.trace_closest(bounce))
    _compare(inst.tlas.trace_closest(shadow), ora.trace_closest(shadow))
    for rays in (bounce, shadow):
        occ = inst.tlas.trace_any(rays, abi.VH_RAY_SHADOW)
        assert np.array_equal(occ, ora.trace_any(rays, abi.VH_RAY_SHADOW))
    assert 0.05 < occ.mean() < 0.95
    info = [b.info() for b in inst.blases]
    assert [i["triangle_count"] for i in info] == [5_242_880, 2_097_152]


def _film_bar(g, o, noise_ref):
    """Image parity: the GPU-vs-oracle relMSE of the same samples is far below
    the relMSE between two independent oracle renders (the noise level)."""
    noise = rel_mse(o, noise_ref)
    diff = rel_mse(g, o)
    close = np.all(np.abs(g[:, :3] - o[:, :3]) <= 1e-3 * np.maximum(np.abs(o[:, :3]), 1e-2), axis=1)
    return diff, noise, close.mean()


def test_config1_film_at_1080p_64spp_matches_oracle(config1):
    sc, inst, ora = config1
    p = sc.render_params(adaptive=False)
    assert (p.width, p.height, p.samples) == (1920, 1080, 64)
    film = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
    inst.render(p, device_film=film)
    torch.cuda.synchronize()
    st = inst.stats()
    assert st["paths"] == 1920 * 1080 * 64 and st["dropped_samples"] == 0
    g = film.reshape(-1, 4).cpu().numpy()
    assert np.isfinite(g).all() and np.array_equal(g[:, 3], np.full(g.shape[0], 64.0, np.float32))
    rng = np.random.default_rng(7)
    pix = rng.choice(sc.width * sc.height, 8192, replace=False)
    o = ora.render_pixels(p, pix)
    other = ora.render_pixels(sc.render_params(seed=4242, adaptive=False), pix)
    diff, noise, close = _film_bar(g[pix] / 64, o / 64, other / 64)
    print(f"config1 64-spp film: relMSE vs oracle {diff:.3e}, noise {noise:.3e}, pixels within 1e-3: {close:.4f}")
    assert diff < 0.05 * noise, (diff, noise)
    # the same 64 paths per pixel: only a pixel one of whose samples took
    # another branch (a roulette or threshold draw at a rounding boundary)
    # differs by more than the approximate shading arithmetic's last bits
    assert close > 0.99, close


def test_config1_single_samples_match_oracle(config1):
    sc, inst, ora = config1
    rng = np.random.default_rng(8)
    film = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
    gs, os_, ns = [], [], []
    for k in (0, 41):
        p = sc.render_params(adaptive=False)
        p.sample_begin, p.sample_end = k, k + 1
        film.zero_()
        inst.render(p, device_film=film)
        torch.cuda.synchronize()
        g = film.reshape(-1, 4).cpu().numpy()
        pix = rng.choice(sc.width * sc.height, 4096, replace=False)
        gs.append(g[pix])
        os_.append(ora.sample_radiance_many(p, pix, k))
        ns.append(ora.sample_radiance_many(sc.render_params(seed=777, adaptive=False), pix, k))
    g, o, n = np.concatenate(gs), np.concatenate(os_), np.concatenate(ns)
    assert np.array_equal(g[:, 3], o[:, 3])
    diff, noise, close = _film_bar(g, o, n)
    print(f"config1 single samples: relMSE vs oracle {diff:.3e}, noise {noise:.3e}, within 1e-3: {close:.4f}")
    assert diff < 0.05 * noise, (diff, noise)
    assert close > 0.999, close


@pytest.fixture(scope="module")
def config2(renderer):
    sc = S.config2()
    assert len(sc.instances) == 1_000_001
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    yield sc, inst, ora
    inst.close()
    ora.close()


def test_config2_million_instances_bit_identical_with_deep_stacks(config2):
    sc, inst, ora = config2
    rng = np.random.default_rng(202)
    p = sc.render_params(adaptive=False)
    n_cam = 400_000
    cam = ora.camera_rays(p, rng.integers(0, sc.width * sc.height, n_cam), rng.integers(0, sc.spp, n_cam))
    oc = ora.trace_closest(cam, abi.VH_RAY_CAMERA)
    _compare(inst.tlas.trace_closest(cam, abi.VH_RAY_CAMERA), oc)
    # rays from inside the ribbon in every direction: they cross the densest
    # part of the instance hierarchy
    n_in = 400_000
    o = rng.normal(size=(n_in, 3)) * np.array([40.0, 3.0, 6.0]) + np.array([0.0, 22.0, 0.0])
    d = rng.normal(size=(n_in, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    ribbon = np.zeros((n_in, 8), np.float32)
    ribbon[:, :3], ribbon[:, 4:7], ribbon[:, 7] = o, d, 1e30
    orib = ora.trace_closest(ribbon)
    _compare(inst.tlas.trace_closest(ribbon), orib)
    assert np.array_equal(inst.tlas.trace_any(ribbon), ora.trace_any(ribbon))
    bounce, shadow = _surface_rays(rng, ribbon, orib, sc.lights, 100_000, 100_000)
    _compare(inst.tlas.trace_closest(bounce), ora.trace_closest(bounce))
    assert np.array_equal(inst.tlas.trace_any(shadow, abi.VH_RAY_SHADOW), ora.trace_any(shadow, abi.VH_RAY_SHADOW))
    assert (oc["t"] > 0).mean() > 0.05 and (orib["t"] > 0).mean() > 0.2
    # stack depth of the same rays (the counting variant traverses like the
    # closest-hit kernel): some rays go past the 12 LDS entries into the
    # global part of the stack, and their hits were among the exact ones above
    depth = np.concatenate([inst.tlas.trace_counts(r)[:, 2] for r in (cam, ribbon, bounce)])
    deep = int((depth > 12).sum())
    print(f"config2: deepest stack {depth.max()} entries, {deep} rays beyond the 12 in LDS, "
          f"mean {depth.mean():.2f}")
    assert deep > 0, depth.max()
