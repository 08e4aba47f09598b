"""ctypes binding of libvork_hip.so and a thin host-side mirror of the
reference's install/render flow.

``Renderer.install(scene)`` plays the part of ``Renderer::install``
(crates/vork-render/src/lib.rs:1464): it builds one BLAS per mesh on the
device and the TLAS over the scene's instances.  ``InstalledScene.render``
plays ``render_frame`` + the path pass (crates/vork-render/src/install.rs:3201,
crates/vork-render/src/passes/path.rs:539-725) for a range of exposure
samples, accumulating into an RGBA32F film.

There is no CPU fallback: if the HIP library is missing or no GPU is present,
every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# VORK_HIP_LIB names another build of the library (the autotuner measures its
# variants this way); the default is the in-tree build.
LIB_PATH = os.environ.get("VORK_HIP_LIB") or os.path.join(HERE, "libvork_hip.so")

EXPORTS = [
    "vh_abi_version", "vh_last_error", "vh_device_create", "vh_device_release", "vh_device_synchronize",
    "vh_blas_build", "vh_blas_get_info", "vh_blas_release", "vh_tlas_build", "vh_tlas_release",
    "vh_trace_closest", "vh_trace_any", "vh_trace_closest_counts", "vh_scene_create", "vh_scene_set_camera",
    "vh_scene_release", "vh_render_params_default", "vh_render", "vh_render_guides", "vh_render_get_stats",
    "vh_denoiser_create", "vh_denoise", "vh_denoiser_release", "vh_sampler_bits", "vh_trace_closest_static",
    "vh_trainer_create", "vh_trainer_step", "vh_trainer_gradients", "vh_trainer_export", "vh_trainer_release",
    "vh_render_shared",
]

_lib = None


class VorkError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"[{status}] {message}")
        self.status = status


def lib():
    """Loads libvork_hip.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: when torch is installed, let it load its
    # runtime first so that the library and torch share it (device tensors
    # passed with VH_DEVICE_POINTERS must come from the same runtime).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    L.vh_abi_version.restype = u32
    L.vh_last_error.restype = i32
    L.vh_last_error.argtypes = [C.c_char_p, C.c_size_t]
    L.vh_device_create.restype = i32
    L.vh_device_create.argtypes = [i32, C.POINTER(vp)]
    L.vh_device_release.argtypes = [vp]
    L.vh_device_synchronize.restype = i32
    L.vh_device_synchronize.argtypes = [vp]
    L.vh_blas_build.restype = i32
    L.vh_blas_build.argtypes = [vp, C.POINTER(abi.Geometry), u32, vp, C.POINTER(vp)]
    L.vh_blas_get_info.restype = i32
    L.vh_blas_get_info.argtypes = [vp, C.POINTER(abi.BlasInfo)]
    L.vh_blas_release.argtypes = [vp]
    L.vh_tlas_build.restype = i32
    L.vh_tlas_build.argtypes = [vp, C.POINTER(vp), u32, vp, u32, u32, u32, vp, C.POINTER(vp)]
    L.vh_tlas_release.argtypes = [vp]
    for name in ("vh_trace_closest", "vh_trace_any", "vh_trace_closest_counts"):
        f = getattr(L, name)
        f.restype = i32
        f.argtypes = [vp, vp, u32, u32, u32, vp, vp]
    L.vh_trace_closest_static.restype = i32
    L.vh_trace_closest_static.argtypes = [vp, vp, u32, u32, vp, vp]
    L.vh_scene_create.restype = i32
    L.vh_scene_create.argtypes = [vp, vp, C.POINTER(abi.SceneDesc), C.POINTER(vp)]
    L.vh_scene_set_camera.restype = i32
    L.vh_scene_set_camera.argtypes = [vp, C.POINTER(abi.Camera)]
    L.vh_scene_release.argtypes = [vp]
    L.vh_render_params_default.argtypes = [C.POINTER(abi.RenderParams), u32, u32, u32]
    L.vh_render.restype = i32
    L.vh_render.argtypes = [vp, C.POINTER(abi.RenderParams), u32, vp, vp]
    L.vh_render_shared.restype = i32
    L.vh_render_shared.argtypes = [vp, u32, vp]
    L.vh_render_guides.restype = i32
    L.vh_render_guides.argtypes = [vp, C.POINTER(abi.RenderParams), u32, vp, vp, vp]
    L.vh_denoiser_create.restype = i32
    L.vh_denoiser_create.argtypes = [vp, vp, C.c_size_t, vp]
    L.vh_denoise.restype = i32
    L.vh_denoise.argtypes = [vp, u32, u32, C.c_float, vp, vp, u32, vp, vp]
    L.vh_denoiser_release.restype = None
    L.vh_denoiser_release.argtypes = [vp]
    L.vh_sampler_bits.restype = i32
    L.vh_sampler_bits.argtypes = [vp, vp, u32, u32, u32, u32, vp, vp]
    L.vh_trainer_create.restype = i32
    L.vh_trainer_create.argtypes = [vp, vp, C.c_size_t, u32, vp]
    L.vh_trainer_step.restype = i32
    L.vh_trainer_step.argtypes = [vp, u32, u32, u32, C.c_float, vp, vp, vp, C.c_float, u32, vp, vp]
    L.vh_trainer_gradients.restype = i32
    L.vh_trainer_gradients.argtypes = [vp, vp, C.c_size_t, vp, C.c_size_t]
    L.vh_trainer_export.restype = i32
    L.vh_trainer_export.argtypes = [vp, vp, C.c_size_t, vp]
    L.vh_trainer_release.restype = None
    L.vh_trainer_release.argtypes = [vp]
    L.vh_render_get_stats.restype = i32
    L.vh_render_get_stats.argtypes = [vp, C.POINTER(abi.RenderStats)]
    _lib = L
    return L


def check(status):
    if status != abi.VH_OK:
        buf = C.create_string_buffer(1024)
        lib().vh_last_error(buf, len(buf))
        raise VorkError(status, buf.value.decode(errors="replace"))


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def _dev_ptr(t):
    """Address of a torch CUDA tensor (device memory owned by the caller)."""
    return C.c_void_p(t.data_ptr())


class Device:
    def __init__(self, ordinal=0):
        self.handle = C.c_void_p()
        check(lib().vh_device_create(ordinal, C.byref(self.handle)))

    def synchronize(self):
        check(lib().vh_device_synchronize(self.handle))

    def close(self):
        if self.handle:
            lib().vh_device_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


    def sampler_bits(self, queries, samples, seed):
        """queries (n, 4) uint32: px, py, sample index, group (or
        abi.VH_SAMPLER_PATH_RANDOM).  Returns (n, 4) uint32 drawn on the GPU."""
        q = np.ascontiguousarray(queries, np.uint32).reshape(-1, 4)
        out = np.zeros_like(q)
        check(lib().vh_sampler_bits(self.handle, _ptr(q), q.shape[0], samples, seed, 0, None, _ptr(out)))
        return out


class Blas:
    def __init__(self, device, positions, indices, normals=None, triangle_materials=None, device_pointers=False,
                 stream=None):
        """positions (V,3) float32, indices (T,3) uint32; numpy arrays, or torch
        CUDA tensors when device_pointers is True.  `stream` (a HIP stream
        handle, e.g. torch.cuda.Stream().cuda_stream) orders the build on that
        stream; builds on different streams run concurrently from several
        threads.  The call returns with the BLAS complete."""
        self.device = device
        g = abi.Geometry()
        if device_pointers:
            g.positions = positions.data_ptr()
            g.indices = indices.data_ptr()
            g.normals = normals.data_ptr() if normals is not None else None
            g.triangle_materials = triangle_materials.data_ptr() if triangle_materials is not None else None
            nv, nt = positions.shape[0], indices.shape[0]
            g.position_stride = positions.stride(0) * 4
            g.normal_stride = normals.stride(0) * 4 if normals is not None else 0
        else:
            positions = np.ascontiguousarray(positions, np.float32)
            indices = np.ascontiguousarray(indices, np.uint32)
            self._keep = [positions, indices]
            g.positions = positions.ctypes.data
            g.indices = indices.ctypes.data
            if normals is not None:
                normals = np.ascontiguousarray(normals, np.float32)
                self._keep.append(normals)
                g.normals = normals.ctypes.data
                g.normal_stride = 12
            if triangle_materials is not None:
                triangle_materials = np.ascontiguousarray(triangle_materials, np.uint32)
                self._keep.append(triangle_materials)
                g.triangle_materials = triangle_materials.ctypes.data
            nv, nt = positions.shape[0], indices.shape[0]
            g.position_stride = 12
        g.vertex_count = nv
        g.triangle_count = nt
        self.handle = C.c_void_p()
        flags = abi.VH_DEVICE_POINTERS if device_pointers else 0
        check(lib().vh_blas_build(device.handle, C.byref(g), flags, C.c_void_p(stream) if stream else None,
                                  C.byref(self.handle)))
        self._keep = None

    def info(self):
        i = abi.BlasInfo()
        check(lib().vh_blas_get_info(self.handle, C.byref(i)))
        return {f: getattr(i, f) for f, _ in abi.BlasInfo._fields_}

    def close(self):
        if self.handle:
            lib().vh_blas_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tlas:
    def __init__(self, device, blases, instances, device_pointers=False, count=None):
        """instances: a ctypes array of abi.Instance, or with device_pointers a
        device tensor holding `count` of them."""
        self.device = device
        self.blases = list(blases)
        self.handle = C.c_void_p()
        self.build(instances, device_pointers=device_pointers, count=count)

    def build(self, instances, refit=False, device_pointers=False, count=None):
        arr = (C.c_void_p * max(1, len(self.blases)))(*[b.handle.value for b in self.blases])
        n = count if count is not None else len(instances)
        ptr = C.c_void_p(instances.data_ptr()) if device_pointers else C.c_void_p(C.addressof(instances))
        check(lib().vh_tlas_build(self.device.handle, arr, len(self.blases), ptr, n,
                                  abi.VH_TLAS_REFIT if refit else abi.VH_TLAS_REBUILD,
                                  abi.VH_DEVICE_POINTERS if device_pointers else 0, None, C.byref(self.handle)))

    def trace_closest(self, rays, mask=0xFF):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        check(lib().vh_trace_closest(self.handle, _ptr(rays), rays.shape[0], mask, 0, None, _ptr(hits)))
        return hits

    def trace_any(self, rays, mask=0xFF):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], np.uint8)
        check(lib().vh_trace_any(self.handle, _ptr(rays), rays.shape[0], mask, 0, None, _ptr(occ)))
        return occ

    def trace_counts(self, rays, mask=0xFF):
        """(n, 4) uint32 per ray: nodes visited, triangles tested, deepest
        traversal-stack entry, 0."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        counts = np.zeros((rays.shape[0], 4), np.uint32)
        check(lib().vh_trace_closest_counts(self.handle, _ptr(rays), rays.shape[0], mask, 0, None, _ptr(counts)))
        return counts

    def close(self):
        if self.handle:
            lib().vh_tlas_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


HIT_DTYPE = np.dtype([("t", np.float32), ("instance", np.uint32), ("primitive", np.uint32),
                      ("custom_index", np.uint32), ("u", np.float32), ("v", np.float32)])


class InstalledScene:
    """A scene resident on the device: BLASes, TLAS, materials, lights."""

    def __init__(self, device, scene, blases, tlas):
        self.device, self.scene, self.blases, self.tlas = device, scene, blases, tlas
        keep = []
        desc = scene.scene_desc(keep)
        self.handle = C.c_void_p()
        check(lib().vh_scene_create(device.handle, tlas.handle, C.byref(desc), C.byref(self.handle)))

    def render(self, params=None, film=None, device_film=None, stream=None):
        """Adds samples [params.sample_begin, params.sample_end) to `film`
        (numpy (H,W,4) float32, created if None) or, with `device_film` (a torch
        CUDA tensor), to device memory.  `stream` is a HIP stream handle (int)
        or None for the device's own stream.  Returns the film."""
        p = params or self.scene.render_params()
        s = C.c_void_p(stream) if stream else None
        if device_film is not None:
            check(lib().vh_render(self.handle, C.byref(p), abi.VH_DEVICE_POINTERS, s, _dev_ptr(device_film)))
            return device_film
        if film is None:
            film = np.zeros((p.height, p.width, 4), np.float32)
        check(lib().vh_render(self.handle, C.byref(p), 0, s, _ptr(film)))
        return film

    def render_guides(self, params=None, film=None, guides=None, stream=None, device=False):
        """Like render, also adding the denoiser guides: `guides` (H, W, 2, 4)
        float32 (albedo rgb + coverage, normal xyz + depth).  With device=True
        both are torch CUDA tensors.  Returns (film, guides)."""
        p = params or self.scene.render_params()
        s = C.c_void_p(stream) if stream else None
        if device:
            check(lib().vh_render_guides(self.handle, C.byref(p), abi.VH_DEVICE_POINTERS, s, _dev_ptr(film),
                                         _dev_ptr(guides)))
            return film, guides
        if film is None:
            film = np.zeros((p.height, p.width, 4), np.float32)
        if guides is None:
            guides = np.zeros((p.height, p.width, 2, 4), np.float32)
        check(lib().vh_render_guides(self.handle, C.byref(p), 0, s, _ptr(film), _ptr(guides)))
        return film, guides

    def stats(self):
        s = abi.RenderStats()
        check(lib().vh_render_get_stats(self.handle, C.byref(s)))
        return {f: getattr(s, f) for f, _ in abi.RenderStats._fields_}

    def close(self):
        if self.handle:
            lib().vh_scene_release(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_shared(jobs, stream=None):
    """Renders several jobs with their paths in one wavefront
    (vh_render_shared): `jobs` is a list of (InstalledScene, params, film,
    guides or None), films and guides torch CUDA tensors that the samples are
    added to."""
    arr = (abi.RenderJob * len(jobs))()
    for k, (inst, p, film, guides) in enumerate(jobs):
        arr[k].scene = inst.handle.value
        arr[k].params = p
        arr[k].film = film.data_ptr()
        arr[k].guides = guides.data_ptr() if guides is not None else None
    check(lib().vh_render_shared(arr, len(jobs), C.c_void_p(stream) if stream else None))


class Renderer:
    def __init__(self, ordinal=0):
        self.device = Device(ordinal)

    def install(self, scene):
        blases = [Blas(self.device, m.positions, m.indices, m.normals, m.triangle_materials) for m in scene.meshes]
        tlas = Tlas(self.device, blases, scene.instance_array())
        return InstalledScene(self.device, scene, blases, tlas)
