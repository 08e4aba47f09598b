"""ctypes bindings of the CPU oracle (oracle/vork_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke entry and the CPU
baseline of bench.py as the checker.  The product (vork_amd) never imports
this package.
"""
import ctypes as C
import os

import numpy as np

from vork_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle library missing at {LIB_PATH}; run `make -C oracle`")
    L = C.CDLL(LIB_PATH)
    u32, f32, vp = C.c_uint32, C.c_float, C.c_void_p
    L.vo_pixel_hash.restype = u32
    L.vo_pixel_hash.argtypes = [u32]
    L.vo_pixel_hash_many.argtypes = [vp, u32, vp]
    L.vo_screen_ray_hash.restype = u32
    L.vo_screen_ray_hash.argtypes = [u32, u32, u32, u32]
    L.vo_sobol.argtypes = [u32, vp]
    L.vo_sobol_directions.argtypes = [vp]
    L.vo_sample_group_bits.argtypes = [u32, u32, u32, u32, u32, u32, vp]
    L.vo_sample_group.argtypes = [u32, u32, u32, u32, u32, u32, vp]
    L.vo_sample_group_bits_batch.argtypes = [u32, u32, u32, u32, u32, vp, u32, vp]
    L.vo_sample_group_bits_many.argtypes = [vp, vp, vp, u32, u32, vp, u32, vp]
    L.vo_path_random.restype = f32
    L.vo_path_random.argtypes = [vp]
    L.vo_filter_table.argtypes = [vp]
    L.vo_filter_cdf.restype = C.c_double
    L.vo_filter_cdf.argtypes = [C.c_double]
    L.vo_filter_offset.restype = f32
    L.vo_filter_offset.argtypes = [vp, f32]
    L.vo_intersect.restype = C.c_int
    L.vo_intersect.argtypes = [vp, vp, vp, vp, vp, f32, f32, C.c_int, vp, vp, vp]
    L.vo_volume_crosses.restype = C.c_int
    L.vo_volume_crosses.argtypes = [vp, vp, vp, vp, vp]
    L.vo_scene_create.restype = vp
    L.vo_scene_create.argtypes = [vp, u32, vp, u32, vp]
    L.vo_scene_release.argtypes = [vp]
    L.vo_invert_transform.restype = C.c_int
    L.vo_invert_transform.argtypes = [vp, vp]
    L.vo_trace_closest.argtypes = [vp, vp, u32, u32, vp, C.c_int]
    L.vo_trace_any.argtypes = [vp, vp, u32, u32, vp, C.c_int]
    L.vo_brdf_moments.argtypes = [f32, f32, vp]
    L.vo_dielectric_albedo.restype = f32
    L.vo_dielectric_albedo.argtypes = [f32, f32, f32]
    L.vo_dielectric_directional_albedo.restype = f32
    L.vo_dielectric_directional_albedo.argtypes = [f32, f32, f32, u32]
    L.vo_bake_moments_texel.argtypes = [u32, u32, vp]
    L.vo_bsdf_sample.argtypes = [vp, vp, vp, vp, vp, vp]
    L.vo_bsdf_eval.argtypes = [vp, vp, vp, vp, vp, u32, vp]
    L.vo_modulated_ior.restype = f32
    L.vo_modulated_ior.argtypes = [f32, f32]
    L.vo_fresnel.restype = f32
    L.vo_fresnel.argtypes = [f32, f32]
    L.vo_light_pick.restype = u32
    L.vo_light_pick.argtypes = [vp, vp, f32, f32, vp]
    L.vo_light_probability.restype = f32
    L.vo_light_probability.argtypes = [vp, u32, vp]
    L.vo_rect_sample.argtypes = [vp, vp, vp, vp]
    L.vo_rect_eval.argtypes = [vp, vp, vp, vp]
    L.vo_sample_radiance.argtypes = [vp, vp, u32, u32, u32, vp]
    L.vo_render.argtypes = [vp, vp, vp, vp, C.c_int]
    L.vo_render_guides.argtypes = [vp, vp, vp, vp, vp, C.c_int]
    L.vo_camera_ray.argtypes = [vp, vp, u32, u32, u32, vp]
    L.vo_render_stats.argtypes = [vp, vp]
    L.vo_camera_rays.argtypes = [vp, vp, vp, vp, u32, vp]
    L.vo_render_pixels.argtypes = [vp, vp, vp, u32, vp, C.c_int]
    L.vo_sample_radiance_many.argtypes = [vp, vp, vp, vp, u32, vp, C.c_int]
    _lib = L
    return L


def ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class MeshDesc(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("indices", C.c_void_p),
                ("triangle_materials", C.c_void_p), ("vertex_count", C.c_uint32), ("triangle_count", C.c_uint32)]


class OracleScene:
    """The oracle's copy of a vork_amd.scene.Scene (meshes, instances, tables)."""

    def __init__(self, scene):
        L = lib()
        self.scene = scene
        self._keep = []
        meshes = (MeshDesc * len(scene.meshes))()
        for k, m in enumerate(scene.meshes):
            pos = np.ascontiguousarray(m.positions, np.float32)
            nrm = np.ascontiguousarray(m.normals, np.float32) if m.normals is not None else None
            idx = np.ascontiguousarray(m.indices, np.uint32)
            mat = np.ascontiguousarray(m.triangle_materials, np.uint32) if m.triangle_materials is not None else None
            self._keep += [pos, nrm, idx, mat]
            meshes[k].positions = pos.ctypes.data
            meshes[k].normals = nrm.ctypes.data if nrm is not None else None
            meshes[k].indices = idx.ctypes.data
            meshes[k].triangle_materials = mat.ctypes.data if mat is not None else None
            meshes[k].vertex_count = pos.shape[0]
            meshes[k].triangle_count = idx.shape[0]
        inst = scene.instance_array()
        desc = scene.scene_desc(self._keep)
        self._keep += [meshes, inst, desc]
        self.handle = L.vo_scene_create(C.addressof(meshes), len(scene.meshes), C.addressof(inst),
                                        len(scene.instances), C.byref(desc))

    def close(self):
        if self.handle:
            lib().vo_scene_release(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_closest(self, rays, mask=0xFF, brute_force=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(rays.shape[0], dtype=HIT_DTYPE)
        lib().vo_trace_closest(self.handle, ptr(rays), rays.shape[0], mask, ptr(hits), int(brute_force))
        return hits

    def trace_any(self, rays, mask=0xFF, brute_force=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], np.uint8)
        lib().vo_trace_any(self.handle, ptr(rays), rays.shape[0], mask, ptr(occ), int(brute_force))
        return occ

    def render(self, params, threads=0, film32=True, film64=True):
        n = params.width * params.height * 4
        f32 = np.zeros(n, np.float32) if film32 else None
        f64 = np.zeros(n, np.float64) if film64 else None
        lib().vo_render(self.handle, C.byref(params), ptr(f32), ptr(f64), threads)
        shape = (params.height, params.width, 4)
        return (f32.reshape(shape) if f32 is not None else None, f64.reshape(shape) if f64 is not None else None)

    def render_guides(self, params, threads=0):
        """(film32 (H, W, 4), guides (H, W, 2, 4)): sums over the sample range."""
        n = params.width * params.height
        f32 = np.zeros(n * 4, np.float32)
        g = np.zeros(n * 8, np.float32)
        lib().vo_render_guides(self.handle, C.byref(params), ptr(f32), None, ptr(g), threads)
        return f32.reshape(params.height, params.width, 4), g.reshape(params.height, params.width, 2, 4)

    def stats(self):
        out = np.zeros(4, np.uint64)
        lib().vo_render_stats(self.handle, ptr(out))
        return out

    def sample_radiance(self, params, px, py, index):
        out = np.zeros(4, np.float32)
        lib().vo_sample_radiance(self.handle, C.byref(params), px, py, index, ptr(out))
        return out

    def camera_rays(self, params, pixels, indices):
        """(n, 8) float32 rays (origin, tmin 0, direction, tmax) of the camera
        samples (row-major pixel, sample index)."""
        pix = np.ascontiguousarray(pixels, np.uint32)
        idx = np.ascontiguousarray(np.broadcast_to(np.asarray(indices, np.uint32), pix.shape), np.uint32)
        out = np.zeros((pix.size, 8), np.float32)
        lib().vo_camera_rays(self.handle, C.byref(params), ptr(pix), ptr(idx), pix.size, ptr(out))
        return out

    def render_pixels(self, params, pixels, threads=0):
        """Film sums (n, 4) float32 over the params' sample range for the
        row-major pixel indices `pixels`."""
        pix = np.ascontiguousarray(pixels, np.uint32)
        out = np.zeros((pix.size, 4), np.float32)
        lib().vo_render_pixels(self.handle, C.byref(params), ptr(pix), pix.size, ptr(out), threads)
        return out

    def sample_radiance_many(self, params, pixels, indices, threads=0):
        """One sample's RGBA value (n, 4) per (row-major pixel, sample index)."""
        pix = np.ascontiguousarray(pixels, np.uint32)
        idx = np.ascontiguousarray(np.broadcast_to(np.asarray(indices, np.uint32), pix.shape), np.uint32)
        out = np.zeros((pix.size, 4), np.float32)
        lib().vo_sample_radiance_many(self.handle, C.byref(params), ptr(pix), ptr(idx), pix.size, ptr(out), threads)
        return out

    def camera_ray(self, params, px, py, index):
        out = np.zeros(8, np.float32)
        lib().vo_camera_ray(self.handle, C.byref(params), px, py, index, ptr(out))
        return out


HIT_DTYPE = np.dtype([("t", np.float32), ("instance", np.uint32), ("primitive", np.uint32),
                      ("custom_index", np.uint32), ("u", np.float32), ("v", np.float32)])


def sample_group_bits(px, py, count, stream, group, n=None, indices=None):
    L = lib()
    if indices is not None:
        idx = np.ascontiguousarray(indices, np.uint32)
        n = idx.size
    else:
        idx = None
    out = np.zeros((n, 4), np.uint32)
    L.vo_sample_group_bits_batch(px, py, count, stream, group, ptr(idx), n, ptr(out))
    return out


def sample_group_bits_many(px, py, index, count, stream, group):
    px, py, index, group = (np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.uint32), np.shape(index)), np.uint32)
                            for a in (px, py, index, group))
    out = np.zeros((index.size, 4), np.uint32)
    lib().vo_sample_group_bits_many(ptr(px), ptr(py), ptr(index), count, stream, ptr(group), index.size, ptr(out))
    return out


def pixel_hash_many(values):
    v = np.ascontiguousarray(values, np.uint32)
    out = np.zeros_like(v)
    lib().vo_pixel_hash_many(ptr(v), v.size, ptr(out))
    return out


def sample_unit(bits):
    return ((np.asarray(bits, np.uint32) >> np.uint32(9)).astype(np.float32) + np.float32(0.5)) * np.float32(1.0 / 8388608.0)


def filter_table():
    t = np.zeros(1025, np.float32)
    lib().vo_filter_table(ptr(t))
    return t
