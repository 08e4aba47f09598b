"""ctypes mirror of include/vork_hip.h (the C ABI of the HIP library).

Field order and sizes follow the header exactly; ``check_layout`` asserts the
sizes the header's comments promise so that a drift between the two fails at
import time rather than as corrupted records.
"""
import ctypes as C

import numpy as np

VH_OK, VH_ERR_USAGE, VH_ERR_SCENE, VH_ERR_DEVICE, VH_ERR_IO = 0, 2, 3, 4, 5
VH_DEVICE_POINTERS = 1
VH_RAY_CAMERA, VH_RAY_SHADOW, VH_RAY_DIFFUSE, VH_RAY_SPECULAR, VH_RAY_TRANSMISSION = 1, 2, 4, 8, 16
VH_INSTANCE_CULL_BACK_FACES = 1
VH_TLAS_REBUILD, VH_TLAS_REFIT = 0, 1
VH_SKY_CONSTANT, VH_SKY_PROCEDURAL = 0, 1
VH_SAMPLER_PATH_RANDOM = 0xFFFFFFFF
VH_ABI_VERSION = 2

F3 = C.c_float * 3


class Geometry(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("position_stride", C.c_uint32), ("vertex_count", C.c_uint32),
                ("indices", C.c_void_p), ("triangle_count", C.c_uint32), ("normal_stride", C.c_uint32),
                ("normals", C.c_void_p), ("triangle_materials", C.c_void_p)]


class BlasInfo(C.Structure):
    _fields_ = [("triangle_count", C.c_uint32), ("node_count", C.c_uint32), ("binary_node_count", C.c_uint32),
                ("max_depth", C.c_uint32), ("sah_cost", C.c_double), ("build_ms", C.c_double),
                ("device_bytes", C.c_uint64)]


class Instance(C.Structure):
    _fields_ = [("transform", C.c_float * 12), ("blas", C.c_uint32), ("custom_index", C.c_uint32),
                ("material", C.c_uint32), ("mask", C.c_uint8), ("flags", C.c_uint8), ("reserved", C.c_uint16)]


# numpy view of vh_instance, for instance tables too large for Python lists
INSTANCE_DTYPE = np.dtype([("transform", "<f4", (12,)), ("blas", "<u4"), ("custom_index", "<u4"),
                           ("material", "<u4"), ("mask", "u1"), ("flags", "u1"), ("reserved", "<u2")])
assert INSTANCE_DTYPE.itemsize == C.sizeof(Instance) == 64


class Ray(C.Structure):
    _fields_ = [("origin", F3), ("tmin", C.c_float), ("direction", F3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("instance", C.c_uint32), ("primitive", C.c_uint32),
                ("custom_index", C.c_uint32), ("u", C.c_float), ("v", C.c_float)]


class Material(C.Structure):
    _fields_ = [("base_color", F3), ("base_metalness", C.c_float), ("specular_weight", C.c_float),
                ("specular_roughness", C.c_float), ("specular_ior", C.c_float), ("reserved", C.c_float)]


class RectLight(C.Structure):
    _fields_ = [("position", F3), ("reserved0", C.c_float), ("u", F3), ("reserved1", C.c_float),
                ("v", F3), ("reserved2", C.c_float), ("normal", F3), ("reserved3", C.c_float),
                ("radiance", F3), ("reserved4", C.c_float)]


class Camera(C.Structure):
    _fields_ = [("eye", F3), ("tan_half_x", C.c_float), ("forward", F3), ("tan_half_y", C.c_float),
                ("right", F3), ("clip_start", C.c_float), ("up", F3), ("clip_end", C.c_float)]


class Sky(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("color", F3), ("zenith", F3), ("sun_exponent", C.c_float),
                ("ground", F3), ("reserved", C.c_float), ("sun_direction", F3), ("reserved2", C.c_float),
                ("sun_radiance", F3), ("reserved3", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("materials", C.c_void_p), ("material_count", C.c_uint32), ("lights", C.c_void_p),
                ("light_count", C.c_uint32), ("camera", Camera), ("sky", Sky)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples", C.c_uint32),
                ("sample_begin", C.c_uint32), ("sample_end", C.c_uint32), ("seed", C.c_uint32),
                ("max_bounces", C.c_uint32), ("diffuse_bounces", C.c_uint32), ("glossy_bounces", C.c_uint32),
                ("caustics", C.c_uint32), ("clamp_direct", C.c_float), ("clamp_indirect", C.c_float),
                ("light_threshold", C.c_float), ("batch_samples", C.c_uint32), ("flags", C.c_uint32),
                ("adaptive_min_samples", C.c_uint32), ("adaptive_threshold", C.c_float),
                ("reserved", C.c_uint32 * 2)]


class RenderJob(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", RenderParams), ("film", C.c_void_p), ("guides", C.c_void_p)]


class RenderStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("dropped_samples", C.c_uint64), ("device_ms", C.c_double),
                ("extend_ms", C.c_double), ("shade_ms", C.c_double), ("shadow_ms", C.c_double),
                ("other_ms", C.c_double), ("extend_launches", C.c_uint64), ("shade_launches", C.c_uint64),
                ("shadow_launches", C.c_uint64), ("extend_node_visits", C.c_uint64),
                ("extend_triangle_tests", C.c_uint64), ("shadow_node_visits", C.c_uint64),
                ("shadow_triangle_tests", C.c_uint64), ("stopped_samples", C.c_uint64),
                ("extend_camera_ms", C.c_double), ("extend_camera_launches", C.c_uint64),
                ("extend_camera_rays", C.c_uint64), ("extend_camera_node_visits", C.c_uint64),
                ("extend_camera_triangle_tests", C.c_uint64)]


VH_RENDER_PROFILE, VH_RENDER_COUNT_VISITS, VH_RENDER_ADAPTIVE = 1, 2, 4
VH_RENDER_SHADE_SORTED, VH_RENDER_SHADE_ARRIVAL = 8, 16


def default_params(width, height, samples, seed=1, max_bounces=12, adaptive=True):
    """Production transport policy (crates/vork-render/src/constants.rs:256-267,
    crates/vork-render/src/passes/path.rs:577-585): bounces 12/4/4, caustics on,
    direct clamp off, indirect clamp 10, light threshold 0.01, adaptive
    sampling on (threshold 0.01, minimum max(16, sqrt(samples)))."""
    p = RenderParams()
    p.width, p.height, p.samples = width, height, samples
    p.sample_begin, p.sample_end, p.seed = 0, samples, seed
    p.max_bounces, p.diffuse_bounces, p.glossy_bounces = max_bounces, 4, 4
    p.caustics = 3
    p.clamp_direct, p.clamp_indirect, p.light_threshold = 0.0, 10.0, 0.01
    p.batch_samples = 0
    p.flags = VH_RENDER_ADAPTIVE if adaptive else 0
    return p


def check_layout():
    sizes = {Geometry: 48, BlasInfo: 40, Instance: 64, Ray: 32, Hit: 24, Material: 32, RectLight: 80,
             Camera: 64, Sky: 80, RenderParams: 76, RenderStats: 176, RenderJob: 104}
    for t, n in sizes.items():
        if C.sizeof(t) != n:
            raise AssertionError(f"{t.__name__} is {C.sizeof(t)} bytes, header layout says {n}")


check_layout()
