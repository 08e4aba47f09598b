"""The serving scheduler on the CPU with a stand-in backend: slots bound the
renders in flight, scenes install once and stay resident across repeat
requests, admission respects the HBM budget with least-recently-used
eviction of idle scenes, first-fit admission lets small requests pass one
that must wait, and failures reach every request they concern."""
import threading
import time

import numpy as np
import pytest

from vork_amd import serve


class Scene:
    def __init__(self, name, nbytes=100, width=8, height=4, spp=2, seed=1):
        self.name, self.nbytes = name, nbytes
        self.width, self.height, self.spp, self.seed = width, height, spp, seed


class FakeBackend:
    def __init__(self, render_s=0.02, fail_install=()):
        self.render_s, self.fail_install = render_s, set(fail_install)
        self.lock = threading.Lock()
        self.running = self.peak = 0
        self.installs, self.releases = [], []

    def memory_budget(self):
        return 10 ** 9

    def slot_base_bytes(self):
        return 10

    def state_bytes(self, width, height, spp):
        return width * height

    def scene_bytes(self, scene):
        return scene.nbytes

    def install(self, key, scene):
        time.sleep(0.005)
        if key in self.fail_install:
            raise RuntimeError(f"install of {key} failed")
        with self.lock:
            self.installs.append(key)
        return scene.nbytes

    def release(self, key):
        with self.lock:
            self.releases.append(key)

    def render(self, slot, key, scene, width, height, spp, seed, adaptive, denoise=False):
        with self.lock:
            self.running += 1
            self.peak = max(self.peak, self.running)
        time.sleep(self.render_s)
        with self.lock:
            self.running -= 1
        return np.full((height, width, 4), hash((key, seed)) % 1000, np.float32), {"slot": slot}


class FakeSharedBackend(FakeBackend):
    """Shared-wavefront stand-in: a film per request summing the sample
    slices each call hands it (value = sample count), and a log of the calls'
    jobs."""

    def __init__(self, batch_paths=256, **kw):
        super().__init__(**kw)
        self.batch, self.calls, self.films = batch_paths, [], {}

    def shared_batch_paths(self):
        return self.batch

    def shared_state_bytes(self):
        return self.batch

    def render_shared(self, items):
        time.sleep(self.render_s)
        self.calls.append([(rid, s0, s1) for _, _, _, rid, _, _, _, _, s0, s1, _ in items])
        seats = [seat for seat, *_ in items]
        assert len(set(seats)) == len(seats)
        for seat, key, scene, rid, w, h, spp, seed, s0, s1, denoise in items:
            f = self.films.setdefault(rid, np.zeros((h, w, 4), np.float32))
            f += s1 - s0

    def finish_shared(self, rid, spp, denoise):
        return self.films.pop(rid), {}

    def drop_shared(self, rid):
        self.films.pop(rid, None)


def test_shared_wavefront_slices_samples_and_admits_arrivals_between_calls():
    b = FakeSharedBackend(batch_paths=256, render_s=0.01)
    s = serve.Server(b, slots=4, shared=True)
    big = Scene("big", width=16, height=16, spp=8)       # 256 px: one sample per call beside others
    small = Scene("small", width=4, height=4, spp=2)
    r_big = s.submit("big", big)
    time.sleep(0.015)
    r_small = [s.submit("small", small, seed=k) for k in range(6)]
    s.wait_all(10)
    s.close()
    assert np.all(r_big.result() == 8) and all(np.all(r.result() == 2) for r in r_small)
    # every call renders at most 4 requests, each slice inside its request's
    # exposure, slices of one request consecutive; the big one shares calls
    # with small ones that arrived after it started
    seen = {}
    for call in b.calls:
        assert len(call) <= 4
        for rid, s0, s1 in call:
            assert s0 == seen.get(rid, 0) and s1 > s0
            seen[rid] = s1
    assert any(len(c) > 1 and r_big.rid in [rid for rid, _, _ in c] for c in b.calls)
    assert s.report()["requests"] == 7


def test_slots_bound_the_renders_in_flight_and_scenes_install_once():
    b = FakeBackend()
    s = serve.Server(b, slots=3)
    a = Scene("a")
    reqs = [s.submit("a", a, seed=k) for k in range(10)]
    s.wait_all(10)
    s.close()
    assert b.peak == 3 and b.installs == ["a"]
    for k, r in enumerate(reqs):
        assert np.array_equal(r.result(), np.full((4, 8, 4), hash(("a", k)) % 1000, np.float32))
    rep = s.report()
    assert rep["requests"] == 10 and rep["samples"] == 10 * 8 * 4 * 2
    assert rep["latency_ms"]["p99"] >= rep["latency_ms"]["p50"] > 0


def test_resident_scenes_are_evicted_least_recently_used_first():
    b = FakeBackend(render_s=0.0)
    # overflow 2 x 10 + scratch 2 x 32 + one 512-byte film: room for two 1000-byte scenes, not three
    s = serve.Server(b, slots=2, budget=3000)
    A, B, C = Scene("A", 1000), Scene("B", 1000), Scene("C", 1000)
    for key, sc in (("A", A), ("B", B), ("A", None), ("C", C)):
        s.submit(key, sc).result(5)
    s.wait_all(5)
    assert b.installs == ["A", "B", "C"] and b.releases == ["B"] and s.evictions == 1
    assert list(s.resident) == ["A", "C"]
    with pytest.raises(KeyError):
        s.submit("B")
    s.submit("B", B).result(5)
    s.close()
    assert b.installs == ["A", "B", "C", "B"]


def test_first_fit_lets_small_requests_pass_and_over_budget_is_refused():
    b = FakeBackend(render_s=0.05)
    s = serve.Server(b, slots=2, budget=2 * 10 + 1000 + 100 + 64 * 4 * 16 + 64 * 4)
    big = Scene("big", 1000, width=16, height=16)
    s.submit("big", big)                                    # holds its scene
    small = Scene("s", 0, width=2, height=2)
    wide = Scene("w", 0, width=64, height=4)
    blocked = s.submit("w", wide)                           # its film does not fit beside the running one
    passing = s.submit("s", small)
    passing.result(5)
    assert not blocked.done.is_set() or blocked.started >= passing.started
    blocked.result(5)
    with pytest.raises(MemoryError):
        s.submit("huge", Scene("huge", 10 ** 7))
    s.close()


def test_an_install_failure_reaches_every_request_of_its_scene():
    b = FakeBackend(fail_install={"bad"})
    s = serve.Server(b, slots=2)
    bad = Scene("bad")
    r1, r2 = s.submit("bad", bad), s.submit("bad", bad)
    ok = s.submit("good", Scene("good"))
    s.wait_all(5)
    for r in (r1, r2):
        with pytest.raises(RuntimeError):
            r.result(1)
    assert ok.result(1).shape == (4, 8, 4)
    assert "bad" not in s.resident and s.report()["failed"] == 2
    s.close()


def test_batch_paths_follow_the_library_sizing():
    assert serve.batch_paths(1920, 1080, 64) == 1920 * 1080 * 64
    assert serve.batch_paths(1920, 1080, 128) == 1920 * 1080 * 64
    assert serve.batch_paths(256, 256, 4) == 256 * 256 * 4
    assert serve.batch_paths(256, 256, 4096) == 1 << 27
