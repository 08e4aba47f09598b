"""GPU parity of the acceleration structures and trace kernels.

The oracle is brute force over every triangle of every instance with the same
FP32 watertight test; the HIP path builds its BVH on the device and must
return the same closest (instance, primitive) and bit-identical t and
barycentrics, equal-t ties included (lowest instance, then primitive).
"""
import math

import numpy as np
import pytest

import oracle
from vork_amd import abi, native, scene as S
from tests.helpers import random_rays

pytestmark = pytest.mark.gpu


def _compare(g, o):
    gh, oh = g["t"] >= 0, o["t"] >= 0
    assert np.array_equal(gh, oh), f"hit/miss differs on {np.sum(gh != oh)} rays"
    assert np.array_equal(g["instance"][gh], o["instance"][oh])
    assert np.array_equal(g["primitive"][gh], o["primitive"][oh])
    assert np.array_equal(g["custom_index"][gh], o["custom_index"][oh])
    assert np.array_equal(g["t"].view(np.uint32), o["t"].view(np.uint32)), "t differs"
    assert np.array_equal(g["u"][gh].view(np.uint32), o["u"][oh].view(np.uint32))
    assert np.array_equal(g["v"][gh].view(np.uint32), o["v"][oh].view(np.uint32))


def test_closest_and_any_match_brute_force_config0(renderer):
    sc = S.config0()
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    rays = random_rays(sc, 6000, seed=1)
    _compare(inst.tlas.trace_closest(rays), ora.trace_closest(rays, brute_force=True))
    occ = inst.tlas.trace_any(rays)
    assert np.array_equal(occ, ora.trace_any(rays, brute_force=True))


def test_rebuilt_scenes_reuse_device_blocks(renderer):
    """A released install's BLAS, TLAS and scene storage goes back to the
    device block cache and the next install of the same size takes it: each
    rebuilt scene, with other vertex positions written over the reused
    blocks, still traces like brute force."""
    rng = np.random.default_rng(5)
    for k in range(3):
        sc = S.config0(width=64, height=64)
        for m in sc.meshes:
            m.positions = (m.positions + rng.normal(0.0, 0.02, m.positions.shape)).astype(np.float32)
        inst = renderer.install(sc)
        ora = oracle.OracleScene(sc)
        rays = random_rays(sc, 3000, seed=10 + k)
        _compare(inst.tlas.trace_closest(rays), ora.trace_closest(rays, brute_force=True))
        assert np.array_equal(inst.tlas.trace_any(rays), ora.trace_any(rays, brute_force=True))
        inst.close()
        inst.tlas.close()
        for b in inst.blases:
            b.close()


def test_camera_rays_match_oracle_bvh(renderer):
    sc = S.config0(width=96, height=96)
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    p = sc.render_params()
    rays = np.array([ora.camera_ray(p, x, y, 0) for y in range(96) for x in range(96)], np.float32)
    rays = rays[:, [0, 1, 2, 7, 4, 5, 6, 3]]  # origin, tmin(0), direction, tmax
    rays[:, 3] = 0.0
    _compare(inst.tlas.trace_closest(rays, mask=abi.VH_RAY_CAMERA), ora.trace_closest(rays, abi.VH_RAY_CAMERA))


def _rot(axis, angle):
    axis = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    x, y, z = axis
    c, s, C = math.cos(angle), math.sin(angle), 1 - math.cos(angle)
    return np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                     [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                     [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])


def test_instanced_transforms_masks_and_culling(renderer):
    rng = np.random.default_rng(5)
    meshes = [S.uv_sphere(1.0, 24, 12), S.box((1.0, 2.0, 0.5)), S.icosphere_mesh(2)]
    instances = []
    for k in range(40):
        R = _rot(rng.normal(size=3), rng.uniform(0, 6.28)) * rng.uniform(0.3, 1.7, size=3)
        T = rng.uniform(-6, 6, size=3)
        m = np.concatenate([R, T[:, None]], 1).reshape(-1)
        instances.append(dict(transform=m.tolist(), blas=int(k % 3), material=0, custom=1000 + k,
                              mask=int(rng.choice([0xFF, 0x01, 0x04])),
                              flags=abi.VH_INSTANCE_CULL_BACK_FACES if k % 5 == 0 else 0))
    # a singular instance must be ignored
    instances.append(dict(transform=[0.0] * 12, blas=0, material=0))
    sc = S.Scene(meshes, instances, [dict(base_color=(0.5, 0.5, 0.5))], [], dict(eye=(0, 0, 10), target=(0, 0, 0)),
                 dict(kind=0, color=(0, 0, 0)))
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    rays = random_rays(S.Scene(meshes, instances[:-1], [], [], {}, {}), 4000, seed=3)
    for mask in (0xFF, 0x01, 0x04):
        _compare(inst.tlas.trace_closest(rays, mask), ora.trace_closest(rays, mask, brute_force=True))
        assert np.array_equal(inst.tlas.trace_any(rays, mask), ora.trace_any(rays, mask, brute_force=True))


def test_shared_edges_are_hit_exactly_once(renderer):
    """A ray through a shared edge of a closed mesh crosses it exactly once per
    side (watertightness): with culling off, the closest hit exists for every
    ray aimed at an interior edge point."""
    m = S.icosphere_mesh(3)
    sc = S.Scene([m], [dict(transform=S.translate(0, 0, 0), blas=0, material=0)], [dict(base_color=(1, 1, 1))], [],
                 dict(eye=(0, 0, 5), target=(0, 0, 0)), dict(kind=0, color=(0, 0, 0)))
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    rng = np.random.default_rng(11)
    P, I = m.positions.astype(np.float64), m.indices
    n = 3000
    t = I[rng.integers(0, I.shape[0], n)]
    w = rng.uniform(0.05, 0.95, size=(n, 1))
    target = P[t[:, 0]] * (1 - w) + P[t[:, 1]] * w
    origin = target * 3.0 + rng.normal(size=(n, 3)) * 0.3
    d = target - origin
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = origin, d, 1e30
    g = inst.tlas.trace_closest(rays)
    assert (g["t"] > 0).all(), "a ray through an edge missed both triangles"
    _compare(g, ora.trace_closest(rays, brute_force=True))


def test_traversal_counts_are_consistent(renderer):
    sc = S.config0()
    inst = renderer.install(sc)
    rays = random_rays(sc, 2000, seed=9)
    counts = inst.tlas.trace_counts(rays)
    hits = inst.tlas.trace_closest(rays)
    assert (counts[:, 0] >= 1).all()
    assert (counts[hits["t"] >= 0, 1] >= 1).all()
    info = inst.blases[0].info()
    assert info["triangle_count"] == 20000
    assert info["node_count"] >= 20000 // (8 * 4)


def test_blas_from_device_pointers_matches_host(renderer):
    torch = pytest.importorskip("torch")
    m = S.uv_sphere(1.0, 32, 16)
    b_host = native.Blas(renderer.device, m.positions, m.indices, m.normals)
    pos = torch.from_numpy(m.positions).cuda()
    idx = torch.from_numpy(m.indices.astype(np.int32)).cuda()
    nrm = torch.from_numpy(m.normals).cuda()
    b_dev = native.Blas(renderer.device, pos, idx, nrm, device_pointers=True)
    sc = S.Scene([m], [dict(transform=S.translate(0, 0, 0), blas=0, material=0)], [dict(base_color=(1, 1, 1))], [],
                 dict(eye=(0, 0, 5), target=(0, 0, 0)), dict(kind=0, color=(0, 0, 0)))
    rays = random_rays(sc, 1000, seed=2)
    t1 = native.Tlas(renderer.device, [b_host], sc.instance_array())
    t2 = native.Tlas(renderer.device, [b_dev], sc.instance_array())
    _compare(t1.trace_closest(rays), t2.trace_closest(rays))


def test_concurrent_builds_on_two_streams_match_serial(renderer):
    """Two BLASes built at once from two threads on two streams (each stream
    has its own build scratch) equal the same builds made one after the
    other on the device's stream: same trees, same hits."""
    torch = pytest.importorskip("torch")
    import concurrent.futures
    meshes = [S.uv_sphere(1.0, 256, 128), S.uv_sphere(0.5, 96, 48)]
    dev_meshes = [(torch.from_numpy(m.positions).cuda(), torch.from_numpy(m.indices.astype(np.int32)).cuda(),
                   torch.from_numpy(m.normals).cuda()) for m in meshes]
    torch.cuda.synchronize()
    serial = [native.Blas(renderer.device, *d, device_pointers=True) for d in dev_meshes]
    streams = [torch.cuda.Stream() for _ in meshes]
    with concurrent.futures.ThreadPoolExecutor(2) as pool:
        for _ in range(2):  # the second round reuses the streams' scratch
            both = list(pool.map(lambda k: native.Blas(renderer.device, *dev_meshes[k], device_pointers=True,
                                                        stream=streams[k].cuda_stream), range(2)))
            for a, b in zip(serial, both):
                ia, ib = a.info(), b.info()
                for key in ("triangle_count", "node_count", "binary_node_count", "max_depth"):
                    assert ia[key] == ib[key], key
                # summed with device atomics in any order
                assert ib["sah_cost"] == pytest.approx(ia["sah_cost"], rel=1e-12)
            insts = [dict(transform=S.translate(0, 0, 0), blas=0, material=0),
                     dict(transform=S.translate(1.2, 0.3, 0), blas=1, material=0)]
            sc = S.Scene(meshes, insts, [dict(base_color=(1, 1, 1))], [],
                         dict(eye=(0, 0, 5), target=(0, 0, 0)), dict(kind=0, color=(0, 0, 0)))
            rays = random_rays(sc, 4000, seed=3)
            t1 = native.Tlas(renderer.device, serial, sc.instance_array())
            t2 = native.Tlas(renderer.device, both, sc.instance_array())
            _compare(t1.trace_closest(rays), t2.trace_closest(rays))


def test_invalid_inputs_raise(renderer):
    with pytest.raises(native.VorkError) as e:
        native.Blas(renderer.device, np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))
    assert e.value.status == abi.VH_ERR_SCENE
    with pytest.raises(native.VorkError) as e:
        native.Blas(renderer.device, np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    assert e.value.status == abi.VH_ERR_SCENE


def _secondary_rays(sc, ora, n, seed):
    """Rays that start on surfaces, as bounce and shadow rays do: at the hit
    point itself, nudged along the face normal, and grazing along it."""
    rng = np.random.default_rng(seed)
    prim = random_rays(sc, n, seed=seed, edge_fraction=0.0)
    h = ora.trace_closest(prim)
    ok = h["t"] > 0
    prim, h = prim[ok], h[ok]
    p = prim[:, :3].astype(np.float64) + prim[:, 4:7].astype(np.float64) * h["t"][:, None]
    d = rng.normal(size=p.shape)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a third graze: nearly perpendicular to the incoming ray
    g = np.cross(prim[:, 4:7].astype(np.float64), rng.normal(size=p.shape))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    k = np.arange(p.shape[0]) % 3
    d[k == 2] = g[k == 2]
    nudge = (rng.uniform(size=(p.shape[0], 1)) < 0.5) * rng.choice([1e-6, 1e-4, 1e-3], size=(p.shape[0], 1))
    o = p - prim[:, 4:7].astype(np.float64) * nudge
    rays = np.zeros((p.shape[0], 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 1e30
    bounded = rng.uniform(size=p.shape[0]) < 0.3
    rays[bounded, 7] = rng.uniform(0.01, 5.0, size=bounded.sum())
    return rays


def test_secondary_rays_match_oracle_on_ggx_scene(renderer):
    sc = S.config1(width=64, height=36, spp=4, sphere_level=6, terrain=128, materials=8)
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    rays = _secondary_rays(sc, ora, 60000, seed=21)
    _compare(inst.tlas.trace_closest(rays), ora.trace_closest(rays))
    assert np.array_equal(inst.tlas.trace_any(rays), ora.trace_any(rays))
    sub = rays[:3000]
    _compare(inst.tlas.trace_closest(sub), ora.trace_closest(sub, brute_force=True))


def test_flock_instances_match_oracle_bvh(renderer):
    sc = S.config2(width=32, height=32, spp=1, birds=5000)
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    rng = np.random.default_rng(4)
    n = 20000
    o = rng.normal(size=(n, 3)) * np.array([40.0, 3.0, 6.0]) + np.array([0.0, 22.0, 0.0])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 1e30
    g = inst.tlas.trace_closest(rays)
    assert (g["t"] > 0).mean() > 0.05
    _compare(g, ora.trace_closest(rays))
    assert np.array_equal(inst.tlas.trace_any(rays), ora.trace_any(rays))


def test_axis_parallel_rays_match_brute_force_and_stay_cheap(renderer):
    """Rays with exactly zero direction components: same hits as brute force,
    and no loss of pruning on the zero axis (a horizontal ray through the
    flock must not enter every instance along its xz path)."""
    sc = S.config0()
    inst = renderer.install(sc)
    ora = oracle.OracleScene(sc)
    rng = np.random.default_rng(8)
    n = 3000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-6, 6, size=(n, 3)) * np.array([1.0, 0.4, 1.0]) + np.array([0.0, 1.0, 0.0])
    d = rng.normal(size=(n, 3))
    zero = rng.integers(0, 3, size=n)
    d[np.arange(n), zero] = 0.0
    two = rng.uniform(size=n) < 0.3
    d[two, (zero[two] + 1) % 3] = 0.0
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 4:7] = d
    rays[:, 7] = 1e30
    rays[::7, 1] = 0.0  # some exactly in the grid's plane
    _compare(inst.tlas.trace_closest(rays), ora.trace_closest(rays, brute_force=True))
    assert np.array_equal(inst.tlas.trace_any(rays), ora.trace_any(rays, brute_force=True))

    flock = S.config2(width=32, height=32, spp=1, birds=20000)
    fi = renderer.install(flock)
    m = 512
    flat = np.zeros((m, 8), np.float32)
    flat[:, 0] = -150.0
    flat[:, 1] = np.linspace(16.0, 28.0, m)
    flat[:, 2] = np.linspace(-4.0, 4.0, m)
    flat[:, 4] = 1.0
    flat[:, 7] = 1e30
    tilted = flat.copy()
    tilted[:, 5] = 1e-6
    tilted[:, 4:7] /= np.linalg.norm(tilted[:, 4:7], axis=1, keepdims=True)
    c_flat = fi.tlas.trace_counts(flat)[:, 0].astype(np.int64)
    c_tilt = fi.tlas.trace_counts(tilted)[:, 0].astype(np.int64)
    assert c_flat.sum() <= 2 * c_tilt.sum() + 1000, (c_flat.sum(), c_tilt.sum())
