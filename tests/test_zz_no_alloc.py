"""vh_trace_closest_static, the closest-hit entry point that allocates nothing,
against vh_trace_closest through the C ABI: the same hits, bit for bit, on
two calls in a row (the second with fewer rays, so stale scratch or stale
claims from the first call would show), plus its argument checks."""
import ctypes as C

import numpy as np
import pytest

from vork_amd import abi, native, scene as S
from tests.helpers import random_rays

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _static(tlas, rays_dev, hits_dev, count, mask=0xFF):
    L = native.lib()
    return L.vh_trace_closest_static(tlas.handle, C.c_void_p(rays_dev.data_ptr()), count, mask, None,
                                     C.c_void_p(hits_dev.data_ptr()))


def _trace(tlas, rays, mask):
    """The existing entry point with device pointers, and the static one."""
    n = rays.shape[0]
    rd = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    want = torch.zeros((n, 6), dtype=torch.int32, device="cuda")
    native.check(native.lib().vh_trace_closest(tlas.handle, C.c_void_p(rd.data_ptr()), n, mask,
                                               abi.VH_DEVICE_POINTERS, None, C.c_void_p(want.data_ptr())))
    got = torch.full((n, 6), -7, dtype=torch.int32, device="cuda")
    native.check(_static(tlas, rd, got, n, mask))
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


def test_static_entry_point_matches_the_allocating_one(renderer):
    sc = S.small_config1(width=64, height=36)
    inst = renderer.install(sc)
    flock = renderer.install(S.config2(width=32, height=32, spp=1, birds=20000))
    big = random_rays(sc, 200_000, seed=31, edge_fraction=0.1)
    small = random_rays(sc, 3_001, seed=32, edge_fraction=0.3)
    for tlas, rays in ((inst.tlas, big), (inst.tlas, small)):
        got, want = _trace(tlas, rays, 0xFF)
        assert np.array_equal(got, want)
        assert (got.view(np.float32)[:, 0] > 0).mean() > 0.2
    # a two-level scene whose deep traversals use the overflow stacks, then a
    # smaller call again
    rng = np.random.default_rng(4)
    n = 100_000
    fr = np.zeros((n, 8), np.float32)
    fr[:, :3] = rng.normal(size=(n, 3)) * np.array([40.0, 3.0, 6.0]) + np.array([0.0, 22.0, 0.0])
    d = rng.normal(size=(n, 3))
    fr[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    fr[:, 7] = 1e30
    for rays, mask in ((fr, 0xFF), (fr[:777], abi.VH_RAY_CAMERA)):
        got, want = _trace(flock.tlas, rays, mask)
        assert np.array_equal(got, want)


def test_static_entry_point_checks_its_arguments(renderer):
    L = native.lib()
    sc = S.config0(width=16, height=16)
    inst = renderer.install(sc)
    rays = torch.zeros((4, 8), dtype=torch.float32, device="cuda")
    hits = torch.zeros((4, 6), dtype=torch.int32, device="cuda")
    assert _static(inst.tlas, rays, hits, 0) == abi.VH_OK
    assert L.vh_trace_closest_static(None, None, 4, 0xFF, None, None) == abi.VH_ERR_USAGE
    too_many = (1 << 31) + 1
    assert _static(inst.tlas, rays, hits, too_many) == abi.VH_ERR_USAGE
    buf = C.create_string_buffer(256)
    L.vh_last_error(buf, 256)
    assert b"2^31" in buf.value
    # a ray array that is not 16-byte aligned
    assert L.vh_trace_closest_static(inst.tlas.handle, C.c_void_p(rays.data_ptr() + 4), 2, 0xFF, None,
                                     C.c_void_p(hits.data_ptr())) == abi.VH_ERR_USAGE
