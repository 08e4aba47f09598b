"""Oracle intersection and BVH, pinned by closed forms and brute force.

The reference's hit IDs come from the Vulkan driver and are not pinned by any
reference fixture (parity unpinned for driver IDs).  What pins the oracle's
own test: the watertight crossing of crates/vork-render/shaders/
vork_medium_domain.slang:40-66 (a ray through a shared edge of a closed mesh
crosses exactly one of the two triangles), t against the closed form of a
plane, and the BVH traversal returning exactly the brute-force answer.
"""
import numpy as np

import oracle
from vork_amd import scene as S
from tests.helpers import random_rays


def test_bvh_matches_brute_force():
    sc = S.config0()
    o = oracle.OracleScene(sc)
    rays = random_rays(sc, 3000, seed=4)
    a, b = o.trace_closest(rays), o.trace_closest(rays, brute_force=True)
    for f in ("t", "instance", "primitive", "u", "v"):
        assert np.array_equal(a[f], b[f]), f
    assert np.array_equal(o.trace_any(rays), o.trace_any(rays, brute_force=True))


def _tri(a, b, c):
    return [np.asarray(x, np.float32) for x in (a, b, c)]


def _isect(a, b, c, o, d, tmin=0.0, tmax=1e30, cull=0):
    L = oracle.lib()
    t, u, v = (np.zeros(1, np.float32) for _ in range(3))
    args = [oracle.ptr(np.asarray(x, np.float32)) for x in (a, b, c, o, d)]
    hit = L.vo_intersect(*args, tmin, tmax, cull, oracle.ptr(t), oracle.ptr(u), oracle.ptr(v))
    return hit, float(t[0]), float(u[0]), float(v[0])


def test_plane_t_closed_form_and_barycentrics():
    a, b, c = _tri((0, 0, 0), (4, 0, 0), (0, 0, 4))
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = np.array([rng.uniform(0.1, 1.5), 0.0, rng.uniform(0.1, 1.5)])
        o = p + np.array([rng.uniform(-1, 1), rng.uniform(0.5, 3), rng.uniform(-1, 1)])
        d = (p - o) / np.linalg.norm(p - o)
        hit, t, u, v = _isect(a, b, c, o, d)
        assert hit
        assert abs(t - np.linalg.norm(p - o)) <= 1e-5 * np.linalg.norm(p - o)
        assert abs(u - p[0] / 4) < 1e-5 and abs(v - p[2] / 4) < 1e-5


def test_tmin_tmax_are_exclusive_and_culling_uses_winding():
    a, b, c = _tri((0, 0, 0), (0, 0, 1), (1, 0, 0))  # winding normal +y
    o, d = np.array([0.2, 1.0, 0.2]), np.array([0.0, -1.0, 0.0])
    assert _isect(a, b, c, o, d)[0]
    assert not _isect(a, b, c, o, d, tmax=1.0)[0]
    assert _isect(a, b, c, o, d, tmax=np.nextafter(np.float32(1.0), np.float32(2.0)))[0]
    assert not _isect(a, b, c, o, d, tmin=1.0)[0]
    # front face (ray against the winding normal) survives culling, back face does not
    assert _isect(a, b, c, o, d, cull=1)[0]
    assert not _isect(a, b, c, np.array([0.2, -1.0, 0.2]), -d, cull=1)[0]
    assert _isect(a, b, c, np.array([0.2, -1.0, 0.2]), -d, cull=0)[0]


def test_shared_edge_crossed_exactly_once():
    """Two triangles sharing an edge, consistently wound: rays through points of
    the shared edge (and its end vertices) hit exactly one of them."""
    a, b, c, d4 = (np.array(x, np.float32) for x in ((0, 0, 0), (1, 0, 0.3), (0.2, 0, 1), (1.3, 0, 1.1)))
    rng = np.random.default_rng(1)
    for k in range(500):
        w = rng.choice([0.0, 1.0, rng.uniform()])
        p = b * (1 - w) + c * w
        o = p + np.array([rng.uniform(-1, 1), rng.uniform(0.2, 2), rng.uniform(-1, 1)], np.float32)
        dirn = (p - o) / np.linalg.norm(p - o)
        h1 = _isect(a, b, c, o, dirn)[0]
        h2 = _isect(b, d4, c, o, dirn)[0]
        if 0.0 < w < 1.0:
            assert h1 + h2 == 1, (k, w)
        else:
            assert h1 + h2 <= 2


def test_degenerate_triangle_never_hits():
    a, b, c = _tri((0, 0, 0), (1, 0, 0), (2, 0, 0))
    assert not _isect(a, b, c, np.array([1.0, 1.0, 0.0]), np.array([0.0, -1.0, 0.0]))[0]
