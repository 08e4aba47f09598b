"""The device sampler draws the oracle's bits exactly.

Every path kernel takes its random numbers from group_bits / path_random in
vork_amd/csrc/vh_shade.h; vh_sampler_bits runs those same functions for a
list of queries.  The first test is the reference's own GPU check,
shader_matches_cpu_reference (crates/vork-render/src/path_sampler.rs:262-318),
with its invocation layout: 47 x 64 invocations, sample index id % 300 of a
300-sample block, pixel (7 p, 13 p + 1) for p = id / 300, stream 0x5eed, the
camera group and the five groups of vertex id % 5.  The oracle restates the
reference's CPU side (path_sampler.rs:36-80, pinned by its published Sobol
points), so equal bits here mean the device draws the reference's numbers.
"""
import numpy as np
import pytest

import oracle
from vork_amd import abi

pytestmark = pytest.mark.gpu


def _oracle_bits(q, samples, seed):
    return oracle.sample_group_bits_many(q[:, 0], q[:, 1], q[:, 2], samples, seed, q[:, 3])


def test_shader_layout_of_the_reference_check(renderer):
    ids = np.arange(64 * 47, dtype=np.uint32)
    index, p = ids % 300, ids // 300
    px, py = p * 7, p * 13 + 1
    vertex = ids % 5
    queries = []
    for slot in range(-1, 5):   # camera, then scatter, light, mesh, environment, decisions
        group = np.zeros_like(ids) if slot < 0 else 1 + vertex * 5 + slot
        queries.append(np.stack([px, py, index, group], 1))
    q = np.concatenate(queries).astype(np.uint32)
    got = renderer.device.sampler_bits(q, 300, 0x5EED)
    want = _oracle_bits(q, 300, 0x5EED)
    assert np.array_equal(got, want), f"{np.sum(np.any(got != want, axis=1))} queries differ"
    # the floats the kernels use, as the reference compares them
    assert np.array_equal(oracle.sample_unit(got), oracle.sample_unit(want))


@pytest.mark.parametrize("samples", [1, 2, 3, 16, 64, 300, 1024])
def test_groups_counts_and_pixels(renderer, samples):
    rng = np.random.default_rng(samples)
    n = 40_000
    q = np.stack([rng.integers(0, 4096, n), rng.integers(0, 2160, n), rng.integers(0, samples, n),
                  rng.integers(0, 41, n)], 1).astype(np.uint32)
    for seed in (1, 0x5EED, 0xFFFFFFFF):
        got = renderer.device.sampler_bits(q, samples, seed)
        assert np.array_equal(got, _oracle_bits(q, samples, seed)), (samples, seed)


def test_path_random_stream(renderer):
    rng = np.random.default_rng(7)
    n = 20_000
    q = np.stack([rng.integers(0, 1920, n), rng.integers(0, 1080, n), rng.integers(0, 4096, n),
                  np.full(n, abi.VH_SAMPLER_PATH_RANDOM)], 1).astype(np.uint32)
    got = renderer.device.sampler_bits(q, 64, 9)
    L = oracle.lib()
    import ctypes as C
    for i in range(0, n, 97):
        state = C.c_uint32(L.vo_screen_ray_hash(int(q[i, 0]), int(q[i, 1]), int(q[i, 2]), 9))
        want = [state.value]
        for _ in range(3):
            L.vo_path_random(C.byref(state))
            want.append(state.value)
        assert got[i].tolist() == want, i
