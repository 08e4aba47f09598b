"""Properties of the synthetic workload generators (CPU only)."""
import numpy as np

from vork_amd import abi, scene as S


def test_bird_has_two_thousand_triangles_and_consistent_normals():
    b = S.bird()
    assert b.triangle_count == 2000
    assert b.triangle_materials.shape == (2000,)
    assert np.isfinite(b.normals).all()
    assert np.allclose(np.linalg.norm(b.normals, axis=1), 1.0, atol=1e-5)


def test_curl_velocity_is_divergence_free():
    rng = np.random.default_rng(0)
    p = rng.normal(size=(64, 3)) * 20
    h = 1e-4
    div = np.zeros(64)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        div += (S.curl_velocity(p + e, 0.3)[:, k] - S.curl_velocity(p - e, 0.3)[:, k]) / (2 * h)
    assert np.abs(div).max() < 1e-6


def test_flock_instances_are_yawed_scaled_and_advected():
    sc = S.config2(width=32, height=32, birds=2000)
    inst = sc.instances
    assert inst.dtype == abi.INSTANCE_DTYPE and inst.shape == (2001,)
    assert sc.triangle_count == 2 * 4 * 4 + 2000 * 2000
    m = inst["transform"][1:].reshape(-1, 3, 4).astype(np.float64)
    R = m[:, :, :3]
    s = np.cbrt(np.linalg.det(R))
    assert ((s > 0.9 - 1e-5) & (s < 1.1 + 1e-5)).all()
    RtR = np.einsum("nij,nik->njk", R, R) / (s ** 2)[:, None, None]
    assert np.allclose(RtR, np.eye(3)[None], atol=1e-5)
    # the bird's +z axis points along the velocity
    pos = m[:, :, 3]
    v = S.curl_velocity(pos, 0.0)
    fwd = R[:, :, 2] / s[:, None]
    vh = v.copy()
    vh[:, 1] = 0
    vh /= np.linalg.norm(vh, axis=1, keepdims=True)
    assert np.allclose(fwd, vh, atol=1e-4)
    frames = S.flock_frames(sc, 3)
    moved = np.linalg.norm(frames[2]["transform"][1:, [3, 7, 11]] - inst["transform"][1:, [3, 7, 11]], axis=1)
    assert (moved > 0).all() and moved.max() < 2.0
    assert np.array_equal(frames[0], inst)
