"""Continuous-batching render server on one GPU.

A persistent request loop over the C-ABI: scenes stay resident in HBM (BLASes,
TLAS, scene tables) across repeat requests, and renders of uneven resolution
and sample count are admitted as slots and memory free up.  Each slot is a
HIP stream with its own wavefront scratch (the library keeps one per stream),
driven by its own host thread: the renders of the admitted requests run at
once and their kernels share the CUs (waves of different requests are
resident together), so small requests fill the chip that one alone leaves
idle.  A request's result depends only on its scene and options: the
counter-based sampler fixes every path, and each pixel's film sums its
samples in sample order, so a result is the same alone and in the mix.

Admission is first-fit over the queue in arrival order: a request is admitted
when a slot is free and the HBM it needs fits the budget after evicting
resident scenes no admitted request uses (least recently used first): its
film, its scene if not resident, and the growth of the slot's stream scratch
(kept at its high-water mark, as the library keeps it).

With `shared=True` the admitted requests instead share one wavefront: a
single loop repeatedly hands every admitted request's next slice of samples
(an equal share of one batch's path slots) to vh_render_shared, whose
extension, shading and shadow queues hold the paths of all of them, each
path traversing and shading its own scene; requests arriving meanwhile join
at the next call, and a request is done once its last slice is.  Adaptive
requests, which vh_render_shared does not take, render alone in the same
loop.

`Server(backend, ...)` holds the scheduling; `GpuBackend` installs and renders
through vork_amd.native.  A request's latency is finish - submit; the server
reports p50/p99 latency beside throughput (samples per second over the run).
"""
import collections
import threading
import time

import numpy as np

PATH_BYTES = 262                   # path state 134 + extension queues 2x32 + shadow queues 2x32
FILM_BYTES = 16                    # RGBA32F per pixel
MOMENT_BYTES = 48                  # adaptive moment per pixel
DENOISE_BYTES = 32 + 16 + 1100     # guides, denoised output, U-Net activations (bf16, all levels) per pixel
def _tuned(knob):
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")) as f:
        return json.load(f)["knobs"][knob]


BATCH_PATHS = 1 << _tuned("VH_BATCH_PATHS_LOG2")   # paths per wavefront batch the library aims for
MAX_BATCH_PATHS = (1 << 30) - 1    # paths of one batch (30-bit path ids)


def batch_paths(width, height, spp):
    """Paths in flight per batch, as the library sizes them by default."""
    pixels = width * height
    batch = min(max(1, BATCH_PATHS // pixels), spp, MAX_BATCH_PATHS // pixels)
    return pixels * batch


class Request:
    def __init__(self, rid, key, scene, width, height, spp, seed, adaptive, denoise=False):
        self.rid, self.key, self.scene = rid, key, scene
        self.width, self.height, self.spp, self.seed, self.adaptive = width, height, spp, seed, adaptive
        self.denoise = denoise
        self.submitted = time.perf_counter()
        self.started = self.finished = None
        self.slot = None
        self.film = self.stats = self.error = None
        self.done = threading.Event()
        self.bytes = 0

    @property
    def samples(self):
        return self.width * self.height * self.spp

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.submitted

    def result(self, timeout=None):
        if not self.done.wait(timeout):
            raise TimeoutError(f"request {self.rid} not finished")
        if self.error is not None:
            raise self.error
        return self.film


class _Scene:
    def __init__(self, key, scene, nbytes):
        self.key, self.scene, self.bytes = key, scene, nbytes
        self.users = 0
        self.ready = threading.Event()
        self.error = None
        self.installing = False


class Server:
    def __init__(self, backend, slots=4, budget=None, shared=False):
        self.backend = backend
        self.shared = shared
        if shared and not 1 <= slots <= 16:
            raise ValueError("a shared wavefront takes 1 to 16 requests at once")
        self.slots = slots
        self.budget = budget if budget is not None else backend.memory_budget()
        self.cv = threading.Condition()
        self.queue = collections.deque()
        self.free = list(range(slots))
        # the streams' traversal overflow stacks (shared: one stream and its whole-batch wavefront state)
        self.base_bytes = (backend.slot_base_bytes() + backend.shared_state_bytes() if shared
                           else backend.slot_base_bytes() * slots)
        self.slot_bytes = [0] * slots                           # wavefront scratch high-water mark per slot
        self.resident = collections.OrderedDict()               # key -> _Scene, least recently used first
        self.active = {}                                        # rid -> Request
        self.finished = []
        self.evictions = 0
        self.next_id = 0
        self.running = True
        self.work = [collections.deque() for _ in range(slots)]
        if shared:
            self.threads = [threading.Thread(target=self._shared_loop, daemon=True)]
        else:
            self.threads = [threading.Thread(target=self._slot_loop, args=(k,), daemon=True) for k in range(slots)]
        for t in self.threads:
            t.start()

    # -------------------------------------------------------------- API
    def submit(self, key, scene=None, width=None, height=None, spp=None, seed=None, adaptive=False, denoise=False):
        """Queues a render of scene `key` (the Scene object is needed unless
        the scene is resident or queued); `denoise` returns the denoised mean
        radiance instead of the film sums.  Returns a Request."""
        with self.cv:
            known = key in self.resident or any(r.key == key and r.scene is not None for r in self.queue)
            if scene is None and not known:
                raise KeyError(f"scene {key!r} is not resident; pass the scene")
            base = scene if scene is not None else self._scene_of(key)
            req = Request(self.next_id, key, scene, width or base.width, height or base.height, spp or base.spp,
                          base.seed if seed is None else seed, adaptive, denoise)
            self.next_id += 1
            # what it needs with every other request finished and every other scene evicted
            sb, total = self._scene_bytes(req), sum(self.slot_bytes)
            need = self.base_bytes + min(total + self._need(req, hw, sb) for hw in self.slot_bytes)
            if need > self.budget:
                raise MemoryError(f"request {req.rid} needs {need} bytes, over the {self.budget}-byte budget")
            self.queue.append(req)
            self._admit()
            return req

    def wait_all(self, timeout=None):
        end = None if timeout is None else time.perf_counter() + timeout
        with self.cv:
            while self.queue or self.active:
                left = None if end is None else end - time.perf_counter()
                if left is not None and left <= 0:
                    raise TimeoutError("requests still running")
                self.cv.wait(left)

    def close(self):
        with self.cv:
            self.running = False
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        self.threads = []   # the threads' targets refer back to the server

    def report(self):
        """Throughput and latency of the finished requests."""
        done = [r for r in self.finished if r.error is None]
        if not done:
            return {"requests": 0}
        lat = np.array([r.latency for r in done]) * 1e3
        t0 = min(r.submitted for r in done)
        t1 = max(r.finished for r in done)
        samples = sum(r.samples for r in done)
        return {"requests": len(done), "failed": len(self.finished) - len(done), "samples": samples,
                "wall_s": t1 - t0, "msamples_per_s": samples / (t1 - t0) / 1e6,
                "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                               "mean": float(lat.mean())},
                "evictions": self.evictions}

    # -------------------------------------------------------------- admission
    def _scene_of(self, key):
        if key in self.resident:
            return self.resident[key].scene
        for r in self.queue:
            if r.key == key and r.scene is not None:
                return r.scene
        raise KeyError(key)

    def _scene_bytes(self, req):
        if req.key in self.resident:
            return 0
        return self.backend.scene_bytes(req.scene if req.scene is not None else self._scene_of(req.key))

    def _need(self, req, slot_hw, scene_bytes):
        state = 0 if self.shared and not req.adaptive else self.backend.state_bytes(req.width, req.height, req.spp)
        return self._film_bytes(req) + max(0, state - slot_hw) + scene_bytes

    @staticmethod
    def _film_bytes(req):
        per = FILM_BYTES + (MOMENT_BYTES if req.adaptive else 0) + (DENOISE_BYTES if req.denoise else 0)
        return req.width * req.height * per

    def _used(self):
        return (self.base_bytes + sum(s.bytes for s in self.resident.values()) + sum(self.slot_bytes) +
                sum(r.bytes for r in self.active.values()))

    def _admit(self):
        """Admits queued requests, first fit in arrival order (holds cv)."""
        progress = True
        while progress and self.free and self.queue:
            progress = False
            for req in list(self.queue):
                if not self.free:
                    break
                state = 0 if self.shared and not req.adaptive else \
                    self.backend.state_bytes(req.width, req.height, req.spp)
                # prefer a slot whose scratch already holds this request
                slot = min(self.free, key=lambda k: (self.slot_bytes[k] < state, -self.slot_bytes[k]))
                sb = self._scene_bytes(req)
                need = self._need(req, self.slot_bytes[slot], sb)
                if self._used() + need > self.budget and not self._evict(self._used() + need - self.budget, req.key):
                    continue
                if req.key not in self.resident:
                    self.resident[req.key] = _Scene(req.key, req.scene if req.scene is not None
                                                    else self._scene_of(req.key), sb)
                self.queue.remove(req)
                self.free.remove(slot)
                self.slot_bytes[slot] = max(self.slot_bytes[slot], state)
                req.bytes = self._film_bytes(req)
                entry = self.resident[req.key]
                entry.users += 1
                self.resident.move_to_end(req.key)
                req.slot = slot
                self.active[req.rid] = req
                self.work[slot].append(req)
                self.cv.notify_all()
                progress = True

    def _evict(self, amount, keep):
        """Drops idle resident scenes, least recently used first, until
        `amount` bytes are free; False (and nothing dropped) if it cannot."""
        idle = [s for k, s in self.resident.items() if s.users == 0 and k != keep and s.ready.is_set()]
        if sum(s.bytes for s in idle) < amount:
            return False
        freed = 0
        for s in idle:
            if freed >= amount:
                break
            del self.resident[s.key]
            self.backend.release(s.key)
            self.evictions += 1
            freed += s.bytes
        return True

    # -------------------------------------------------------------- slots
    def _ensure_installed(self, req):
        """Installs the request's scene unless it is (or is being) installed;
        raises the install's error."""
        with self.cv:
            entry = self.resident[req.key]
            install = not entry.installing and not entry.ready.is_set()
            if install:
                entry.installing = True
        if install:
            try:
                nbytes = self.backend.install(req.key, entry.scene)
                with self.cv:
                    entry.bytes = nbytes
            except Exception as e:  # the install failed: every request of the scene fails with it
                entry.error = e
            finally:
                entry.ready.set()
        entry.ready.wait()
        if entry.error is not None:
            raise entry.error
        return entry

    def _finish(self, req, seat):
        req.finished = time.perf_counter()
        with self.cv:
            entry = self.resident[req.key]
            entry.users -= 1
            if entry.error is not None and entry.users == 0 and self.resident.get(req.key) is entry:
                del self.resident[req.key]
            del self.active[req.rid]
            self.finished.append(req)
            self.free.append(seat)
            self._admit()
            self.cv.notify_all()
        req.done.set()

    def _shared_loop(self):
        """Every admitted request's next slice of samples in one
        vh_render_shared call, again and again; arrivals join the next call."""
        done_samples = {}
        while True:
            with self.cv:
                while self.running and not self.active:
                    self.cv.wait()
                if not self.active:
                    return
                batch = sorted(self.active.values(), key=lambda r: r.rid)
                for req in batch:
                    for q in self.work:
                        if req in q:
                            q.remove(req)
            jobs = []
            for req in batch:
                try:
                    self._ensure_installed(req)
                except Exception as e:
                    req.error = e
                    self._finish(req, req.slot)
                    continue
                if req.started is None:
                    req.started = time.perf_counter()
                if req.adaptive:   # vh_render_shared renders plain sampling only
                    try:
                        entry = self.resident[req.key]
                        req.film, req.stats = self.backend.render(0, req.key, entry.scene, req.width, req.height,
                                                                  req.spp, req.seed, True, req.denoise)
                    except Exception as e:
                        req.error = e
                    self._finish(req, req.slot)
                    continue
                jobs.append(req)
            if not jobs:
                continue
            share = self.backend.shared_batch_paths() // len(jobs)
            items = []
            for req in jobs:
                s0 = done_samples.get(req.rid, 0)
                s1 = min(req.spp, s0 + max(1, share // (req.width * req.height)))
                items.append((req, s0, s1))
            try:
                self.backend.render_shared([(req.slot, req.key, self.resident[req.key].scene, req.rid, req.width,
                                             req.height, req.spp, req.seed, s0, s1, req.denoise)
                                            for req, s0, s1 in items])
            except Exception as e:
                for req, _, _ in items:
                    req.error = e
                    self.backend.drop_shared(req.rid)
                    self._finish(req, req.slot)
                continue
            for req, s0, s1 in items:
                done_samples[req.rid] = s1
                if s1 == req.spp:
                    del done_samples[req.rid]
                    try:
                        req.film, req.stats = self.backend.finish_shared(req.rid, req.spp, req.denoise)
                    except Exception as e:
                        req.error = e
                    self._finish(req, req.slot)

    def _slot_loop(self, k):
        while True:
            with self.cv:
                while self.running and not self.work[k]:
                    self.cv.wait()
                if not self.work[k]:
                    return
                req = self.work[k].popleft()
                entry = self.resident[req.key]
                install = not entry.installing and not entry.ready.is_set()
                if install:
                    entry.installing = True
            try:
                if install:
                    try:
                        nbytes = self.backend.install(req.key, entry.scene)
                        with self.cv:
                            entry.bytes = nbytes
                    except Exception as e:  # the install failed: every request of the scene fails with it
                        entry.error = e
                    finally:
                        entry.ready.set()
                entry.ready.wait()
                if entry.error is not None:
                    raise entry.error
                req.started = time.perf_counter()
                req.film, req.stats = self.backend.render(k, req.key, entry.scene, req.width, req.height, req.spp,
                                                          req.seed, req.adaptive, req.denoise)
            except Exception as e:
                req.error = e
            req.finished = time.perf_counter()
            with self.cv:
                entry.users -= 1
                if entry.error is not None and entry.users == 0 and self.resident.get(req.key) is entry:
                    del self.resident[req.key]
                del self.active[req.rid]
                self.finished.append(req)
                self.free.append(k)
                self._admit()
                self.cv.notify_all()
            req.done.set()


class GpuBackend:
    """Installs and renders through vork_amd.native: one HIP stream per slot,
    one vh_scene per (scene, slot, resolution) over the scene's resident
    BLASes and TLAS, device films."""

    def __init__(self, ordinal=0, slots=4):
        import torch
        from . import native
        self.torch, self.native = torch, native
        self.ordinal = ordinal
        self.renderer = native.Renderer(ordinal)
        self.streams = [torch.cuda.Stream(device=ordinal) for _ in range(slots)]
        props = torch.cuda.get_device_properties(ordinal)
        self.total = props.total_memory
        self.cus = props.multi_processor_count
        self.scenes = {}
        self.lock = threading.Lock()
        self.denoisers = [None] * slots
        self.shared_films = {}   # request id -> (film, guides or None) on the device, summed over slices

    def memory_budget(self):
        return int(0.85 * self.total)

    def slot_base_bytes(self):
        # traversal overflow stacks of one stream: the persistent grid's threads x 244 8-byte entries
        return self.cus * _tuned("VH_GRID_BLOCKS_PER_CU") * _tuned("VH_BLOCK") * 244 * 8

    def state_bytes(self, width, height, spp):
        return batch_paths(width, height, spp) * PATH_BYTES

    def shared_batch_paths(self):
        """Path slots of one vh_render_shared batch."""
        return min(BATCH_PATHS, 1 << 27)

    def shared_state_bytes(self):
        return self.shared_batch_paths() * PATH_BYTES

    def scene_bytes(self, scene):
        tris = scene.triangle_count if not isinstance(scene.instances, np.ndarray) else \
            sum(m.triangle_count for m in scene.meshes)
        verts = sum(m.positions.shape[0] for m in scene.meshes)
        n_inst = len(scene.instances)
        # leaf-order triangles (36 B) + ids (4) + shading copies (index 12, material 4) + vertices (24) + nodes
        return tris * (36 + 4 + 16 + 20) + verts * 24 + n_inst * 200 + (1 << 20)

    def install(self, key, scene):
        native = self.native
        blases = [native.Blas(self.renderer.device, m.positions, m.indices, m.normals, m.triangle_materials)
                  for m in scene.meshes]
        tlas = native.Tlas(self.renderer.device, blases, scene.instance_array())
        with self.lock:
            self.scenes[key] = {"blases": blases, "tlas": tlas, "views": {}}
        return self.scene_bytes(scene)

    def release(self, key):
        with self.lock:
            entry = self.scenes.pop(key, None)
        if entry:
            for v in entry["views"].values():
                v.close()
            entry["tlas"].close()
            for b in entry["blases"]:
                b.close()

    def close(self):
        """Releases every resident scene, the denoisers and the device."""
        for key in list(self.scenes):
            self.release(key)
        for d in self.denoisers:
            if d is not None:
                d.close()
        self.renderer.device.close()

    def _view(self, slot, key, scene, width, height):
        import copy
        with self.lock:
            entry = self.scenes[key]
            v = entry["views"].get((slot, width, height))
        if v is None:
            sc = copy.copy(scene)
            sc.width, sc.height = width, height
            v = self.native.InstalledScene(self.renderer.device, sc, entry["blases"], entry["tlas"])
            with self.lock:
                entry["views"][(slot, width, height)] = v
        return v

    def render_shared(self, items):
        """Samples [s0, s1) of each item (seat, key, scene, request id, width,
        height, spp, seed, s0, s1, denoise) in one vh_render_shared call,
        added to the request's device film (and guides when it is denoised)."""
        torch = self.torch
        dev = f"cuda:{self.ordinal}"
        stream = self.streams[0]
        jobs = []
        with torch.cuda.stream(stream):
            for seat, key, scene, rid, width, height, spp, seed, s0, s1, denoise in items:
                view = self._view(seat, key, scene, width, height)
                if rid not in self.shared_films:
                    film = torch.zeros((height, width, 4), dtype=torch.float32, device=dev)
                    guides = torch.zeros((height, width, 2, 4), dtype=torch.float32, device=dev) if denoise else None
                    self.shared_films[rid] = (film, guides, seat, view)
                film, guides, _, _ = self.shared_films[rid]
                p = view.scene.render_params(samples=spp, seed=seed, adaptive=False)
                p.sample_begin, p.sample_end = s0, s1
                jobs.append((view, p, film, guides))
            self.native.render_shared(jobs, stream=stream.cuda_stream)

    def finish_shared(self, rid, spp, denoise):
        """The request's result (film sums, or the denoised mean radiance) and
        its view's stats; frees its device film."""
        torch = self.torch
        film, guides, seat, view = self.shared_films.pop(rid)
        stream = self.streams[0]
        with torch.cuda.stream(stream):
            if not denoise:
                return film.cpu().numpy(), view.stats()
            from .denoise import Denoiser
            if self.denoisers[0] is None:
                self.denoisers[0] = Denoiser(self.renderer.device)
            out = self.denoisers[0](film, guides, spp, stream=stream.cuda_stream)
            return out.cpu().numpy(), view.stats()

    def drop_shared(self, rid):
        self.shared_films.pop(rid, None)

    def render(self, slot, key, scene, width, height, spp, seed, adaptive, denoise=False):
        torch = self.torch
        view = self._view(slot, key, scene, width, height)
        p = view.scene.render_params(samples=spp, seed=seed, adaptive=adaptive)
        stream = self.streams[slot]
        dev = f"cuda:{self.ordinal}"
        with torch.cuda.stream(stream):
            film = torch.zeros((height, width, 4), dtype=torch.float32, device=dev)
            if not denoise:
                view.render(p, device_film=film, stream=stream.cuda_stream)
                return film.cpu().numpy(), view.stats()
            from .denoise import Denoiser
            if self.denoisers[slot] is None:
                self.denoisers[slot] = Denoiser(self.renderer.device)
            guides = torch.zeros((height, width, 2, 4), dtype=torch.float32, device=dev)
            view.render_guides(p, film, guides, stream=stream.cuda_stream, device=True)
            out = self.denoisers[slot](film, guides, spp, stream=stream.cuda_stream)
            return out.cpu().numpy(), view.stats()
