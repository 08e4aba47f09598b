"""The render server on the GPU: requests of uneven scene, resolution and
sample count rendered alone (one slot, one at a time) and in the mix (four
slots on four HIP streams, all submitted at once, with a budget that forces
evictions) give bit-identical films."""
import numpy as np
import pytest

from vork_amd import scene as S
from vork_amd import serve

pytestmark = pytest.mark.gpu


def requests():
    scenes = {"c0": S.config0(width=64, height=48, spp=4), "c0b": S.config0(seed=2, width=48, height=48, spp=4),
              "g": S.small_config1(width=96, height=54, spp=4)}
    mix = [("c0", 64, 48, 4, 1, False, False), ("g", 96, 54, 8, 2, False, False), ("c0b", 32, 32, 1, 3, False, False),
           ("c0", 128, 72, 2, 4, False, False), ("g", 48, 27, 16, 5, False, False), ("c0b", 80, 60, 4, 6, False, False),
           ("c0", 64, 48, 24, 7, True, False), ("g", 96, 54, 20, 8, True, False), ("c0", 64, 48, 4, 1, False, False),
           ("g", 96, 54, 4, 9, False, True), ("c0b", 80, 60, 4, 10, False, True)]
    return scenes, mix


def run(backend, slots, budget=None, shared=False):
    scenes, mix = requests()
    s = serve.Server(backend, slots=slots, budget=budget, shared=shared)
    try:
        reqs = [s.submit(k, scenes[k], w, h, spp, seed, adaptive, denoise)
                for k, w, h, spp, seed, adaptive, denoise in mix]
        s.wait_all(120)
        return [r.result() for r in reqs], s.report()
    finally:
        s.close()


def test_results_are_the_same_alone_and_in_the_mix():
    backend = serve.GpuBackend(0, slots=4)
    try:
        alone, rep1 = run(backend, 1)
        for k in list(backend.scenes):
            backend.release(k)
        mixed, rep4 = run(backend, 4)
        assert rep1["requests"] == rep4["requests"] == len(alone) and rep4["failed"] == 0
        for a, m in zip(alone, mixed):
            assert np.isfinite(a).all() and a[..., :3].mean() > 0
            assert np.array_equal(a, m)
        assert (alone[-1][..., 3] == 1).all()   # denoised requests return the mean radiance
        assert np.array_equal(alone[0], alone[8])  # a repeat request of a resident scene
        # every scene counted as 1 GiB against a budget with room for one and a
        # half beside the largest films and scratch: one scene resident at a time
        for k in list(backend.scenes):
            backend.release(k)
        scenes, mix = requests()
        state = max(backend.state_bytes(w, h, spp) for _, w, h, spp, _, _, _ in mix)
        film = max(w * h * (serve.FILM_BYTES + serve.MOMENT_BYTES + serve.DENOISE_BYTES) for _, w, h, *_ in mix)
        backend.scene_bytes = lambda scene: 1 << 30
        tight = 4 * (backend.slot_base_bytes() + state + film) + (3 << 29)
        evicted, rep = run(backend, 4, budget=tight)
        assert rep["evictions"] > 0
        for a, m in zip(alone, evicted):
            assert np.array_equal(a, m)
    finally:
        backend.close()


def test_a_1080p_64spp_request_on_a_large_scene_beside_small_ones():
    """Config 3 at its largest request: a 1920x1080 x 64-spp render of a
    1,835,008-triangle scene admitted with small requests of other scenes;
    every film is the same alone (one slot, in order) and in the mix (four
    slots at once)."""
    big = S.config1(seed=3, sphere_level=8, terrain=512, materials=16)
    assert big.triangle_count >= 1_800_000
    scenes = {"big": big, "c0": S.config0(width=64, height=48, spp=4), "g": S.small_config1(width=96, height=54, spp=4)}
    mix = [("c0", 256, 256, 4, 1), ("big", 1920, 1080, 64, 2), ("g", 512, 512, 1, 3), ("c0", 64, 48, 16, 4),
           ("g", 96, 54, 64, 5), ("big", 256, 256, 16, 6), ("c0", 1280, 720, 1, 7)]
    backend = serve.GpuBackend(0, slots=4)
    try:
        results = []
        for slots in (1, 4):
            s = serve.Server(backend, slots=slots)
            try:
                reqs = [s.submit(k, scenes[k], w, h, spp, seed, False, False) for k, w, h, spp, seed in mix]
                s.wait_all(300)
                results.append([r.result() for r in reqs])
                assert s.report()["failed"] == 0
            finally:
                s.close()
        alone, mixed = results
        for (k, w, h, spp, _), a, m in zip(mix, alone, mixed):
            assert a.shape == (h, w, 4) and np.isfinite(a).all()
            assert np.array_equal(a[..., 3], np.full((h, w), float(spp), np.float32))
            assert np.array_equal(a, m), (k, w, h, spp)
    finally:
        backend.close()


def test_shared_wavefront_server_gives_the_results_of_each_request_alone():
    """The shared-queue server (every admitted request's next slice of samples
    in one vh_render_shared call, arrivals joining the next call) returns each
    request's film, adaptive and denoised ones included, bit-identical to the
    request rendered alone."""
    backend = serve.GpuBackend(0, slots=4)
    try:
        alone, _ = run(backend, 1)
        for k in list(backend.scenes):
            backend.release(k)
        for slots in (4, 16):
            shared, rep = run(backend, slots, shared=True)
            assert rep["failed"] == 0 and rep["requests"] == len(alone)
            for a, m in zip(alone, shared):
                assert np.array_equal(a, m)
    finally:
        backend.close()
