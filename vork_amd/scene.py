"""Procedural scenes of the benchmark workloads, as plain numpy arrays.

The std geometry follows the reference's generators in
packages/ts/src/lib/shapes.ts (box 60-78, UV sphere 80-98, grid 239-256):
same vertex order, winding and normals.  Everything random is drawn from a
counter-based hash of (seed, stream, index), so a scene is a pure function of
its seed.
"""
from dataclasses import dataclass, field
import math

import numpy as np

from . import abi

_MIX = np.uint32(0x45D9F3B)


def hash_u32(x):
    """The reference's 32-bit avalanche (crates/vork-slang/src/vork_sampling.slang:3-7)."""
    x = np.asarray(x, dtype=np.uint32)
    with np.errstate(over="ignore"):
        v = (x ^ (x >> np.uint32(16))) * _MIX
        v = (v ^ (v >> np.uint32(16))) * _MIX
    return v ^ (v >> np.uint32(16))


def uniform(seed, stream, n):
    """n counter-based uniforms in [0, 1) for (seed, stream)."""
    i = np.arange(n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = hash_u32(np.uint32(seed & 0xFFFFFFFF) ^ hash_u32(np.uint32(stream & 0xFFFFFFFF)))
        h = hash_u32((i.astype(np.uint32) * np.uint32(0x9E3779B9)) ^ key)
        h = hash_u32(h ^ np.uint32(0x85EBCA6B))
    return (h >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


@dataclass
class Mesh:
    positions: np.ndarray          # (V, 3) float32
    normals: np.ndarray            # (V, 3) float32
    indices: np.ndarray            # (T, 3) uint32
    triangle_materials: np.ndarray = None  # (T,) uint32 or None
    name: str = "mesh"

    @property
    def triangle_count(self):
        return int(self.indices.shape[0])


@dataclass
class Scene:
    meshes: list
    instances: object              # list of dict(transform=(12,), blas=int, material=int, mask=int, flags=int,
                                   # custom=int), or a numpy array of abi.INSTANCE_DTYPE
    materials: np.ndarray          # structured list of abi.Material-compatible tuples
    lights: list
    camera: dict
    sky: dict
    width: int = 256
    height: int = 256
    spp: int = 16
    max_depth: int = 4
    seed: int = 1
    name: str = "scene"
    notes: dict = field(default_factory=dict)

    @property
    def triangle_count(self):
        if isinstance(self.instances, np.ndarray):
            per = np.array([m.triangle_count for m in self.meshes], np.int64)
            return int(per[self.instances["blas"]].sum())
        return sum(self.meshes[i["blas"]].triangle_count for i in self.instances)

    # ---------------------------------------------------------------- ABI
    def instance_array(self):
        if isinstance(self.instances, np.ndarray):
            a = np.ascontiguousarray(self.instances, dtype=abi.INSTANCE_DTYPE)
            self.instances = a
            return (abi.Instance * a.shape[0]).from_buffer(a)
        arr = (abi.Instance * len(self.instances))()
        for k, inst in enumerate(self.instances):
            arr[k].transform[:] = [float(v) for v in inst["transform"]]
            arr[k].blas = inst["blas"]
            arr[k].custom_index = inst.get("custom", k)
            arr[k].material = inst.get("material", 0)
            arr[k].mask = inst.get("mask", 0xFF)
            arr[k].flags = inst.get("flags", 0)
        return arr

    def material_array(self):
        arr = (abi.Material * len(self.materials))()
        for k, m in enumerate(self.materials):
            arr[k].base_color[:] = m["base_color"]
            arr[k].base_metalness = m.get("base_metalness", 0.0)
            arr[k].specular_weight = m.get("specular_weight", 1.0)
            arr[k].specular_roughness = m.get("specular_roughness", 0.5)
            arr[k].specular_ior = m.get("specular_ior", 1.5)
        return arr

    def light_array(self):
        arr = (abi.RectLight * max(1, len(self.lights)))()
        for k, l in enumerate(self.lights):
            arr[k].position[:] = l["position"]
            arr[k].u[:] = l["u"]
            arr[k].v[:] = l["v"]
            arr[k].normal[:] = l["normal"]
            arr[k].radiance[:] = l["radiance"]
        return arr

    def camera_struct(self):
        c = abi.Camera()
        cam = self.camera
        eye = np.asarray(cam["eye"], np.float64)
        if "target" in cam:
            f = np.asarray(cam["target"], np.float64) - eye
            up = np.asarray(cam.get("up", (0, 1, 0)), np.float64)
        else:  # an explicit frame: forward and the camera's up axis
            f = np.asarray(cam["forward"], np.float64)
            up = np.asarray(cam["up"], np.float64)
        f /= np.linalg.norm(f)
        r = np.cross(f, up)
        r /= np.linalg.norm(r)
        u = np.cross(r, f)
        aspect = self.width / self.height
        # libraries/std/camera.slang:25-33: the sensor side named by `fit`
        # spans `span` (Auto: the longer image side), lens in mm.
        lens = cam.get("lens", 18.0 / math.tan(math.radians(22.5)))
        fit = cam.get("fit", "auto")
        span = aspect if fit == "horizontal" else 1.0 if fit == "vertical" else (aspect if aspect > 1 else 1.0)
        sensor = cam.get("sensor_height", 24.0) if fit == "vertical" else cam.get("sensor_width", 36.0)
        height = sensor / span
        screen_x = 2 * lens / (height * aspect)
        screen_y = 2 * lens / height
        c.eye[:] = eye.tolist()
        c.forward[:] = f.tolist()
        c.right[:] = r.tolist()
        c.up[:] = u.tolist()
        c.tan_half_x = 1.0 / screen_x
        c.tan_half_y = 1.0 / screen_y
        c.clip_start = cam.get("clip_start", 0.1)
        c.clip_end = cam.get("clip_end", 1000.0)
        return c

    def sky_struct(self):
        s = abi.Sky()
        sky = self.sky
        s.kind = sky.get("kind", abi.VH_SKY_CONSTANT)
        s.color[:] = sky["color"]
        s.zenith[:] = sky.get("zenith", sky["color"])
        s.ground[:] = sky.get("ground", sky["color"])
        sd = np.asarray(sky.get("sun_direction", (0, 1, 0)), np.float64)
        s.sun_direction[:] = (sd / np.linalg.norm(sd)).tolist()
        s.sun_radiance[:] = sky.get("sun_radiance", (0, 0, 0))
        s.sun_exponent = sky.get("sun_exponent", 1.0)
        return s

    def scene_desc(self, keep):
        """abi.SceneDesc; `keep` (a list) holds the arrays it points to."""
        mats, lights = self.material_array(), self.light_array()
        keep += [mats, lights]
        d = abi.SceneDesc()
        d.materials = C_addr(mats)
        d.material_count = len(self.materials)
        d.lights = C_addr(lights)
        d.light_count = len(self.lights)
        d.camera = self.camera_struct()
        d.sky = self.sky_struct()
        return d

    def render_params(self, samples=None, seed=None, width=None, height=None, adaptive=True):
        p = abi.default_params(width or self.width, height or self.height, samples or self.spp,
                               seed=self.seed if seed is None else seed, max_bounces=self.max_depth,
                               adaptive=adaptive)
        return p


def C_addr(arr):
    import ctypes
    return ctypes.addressof(arr)


# ------------------------------------------------------------------ shapes
def translate(x, y, z):
    return [1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z]


def box(size=(2.0, 2.0, 2.0)):
    """shapes.ts:60-78: centred box, 4 vertices and 2 triangles per face."""
    faces = [((1, 0, 0), (0, 0, -1), (0, 1, 0)), ((-1, 0, 0), (0, 0, 1), (0, 1, 0)),
             ((0, 1, 0), (1, 0, 0), (0, 0, -1)), ((0, -1, 0), (1, 0, 0), (0, 0, 1)),
             ((0, 0, 1), (1, 0, 0), (0, 1, 0)), ((0, 0, -1), (-1, 0, 0), (0, 1, 0))]
    pos, nrm, idx = [], [], []
    for n, u, v in faces:
        ids = []
        for y in (-1, 1):
            for x in (-1, 1):
                p = [(n[k] + u[k] * x + v[k] * y) * size[k] / 2 for k in range(3)]
                ids.append(len(pos))
                pos.append(p)
                nrm.append(n)
        idx += [(ids[0], ids[1], ids[3]), (ids[0], ids[3], ids[2])]
    return Mesh(np.asarray(pos, np.float32), np.asarray(nrm, np.float32), np.asarray(idx, np.uint32), name="box")


def uv_sphere(radius=1.0, longitude=64, rings=32):
    """shapes.ts:80-98: Y-up UV sphere with separate seam and pole vertices."""
    pos, nrm = [], []
    tau = 2 * math.pi
    for y in range(rings + 1):
        for x in range(longitude + 1):
            a = y * math.pi / rings
            az = (0 if x == longitude else x) * tau / longitude
            c, s = math.cos(az), math.sin(az)
            if y == 0:
                n = (0.0, 1.0, 0.0)
            elif y == rings:
                n = (0.0, -1.0, 0.0)
            else:
                n = (math.sin(a) * c, math.cos(a), math.sin(a) * s)
            pos.append([v * radius for v in n])
            nrm.append(n)
    idx = []
    for y in range(rings):
        for x in range(longitude):
            a = y * (longitude + 1) + x
            b = a + longitude + 1
            if y < rings - 1:
                idx.append((a, b + 1, b))
            if y > 0:
                idx.append((a, a + 1, b + 1))
    return Mesh(np.asarray(pos, np.float32), np.asarray(nrm, np.float32), np.asarray(idx, np.uint32), name="sphere")


def grid(size=(2.0, 2.0), segments=(1, 1), heights=None):
    """shapes.ts:239-256: XZ grid facing +Y; optional per-vertex heights
    (row-major (nz+1, nx+1)) lift it into a heightfield with smooth normals."""
    nx, nz = segments
    xs = (np.arange(nx + 1, dtype=np.float64) / nx - 0.5) * size[0]
    zs = (np.arange(nz + 1, dtype=np.float64) / nz - 0.5) * size[1]
    X, Z = np.meshgrid(xs, zs)
    Y = np.zeros_like(X) if heights is None else heights
    pos = np.stack([X, Y, Z], -1).reshape(-1, 3).astype(np.float32)
    a = (np.arange(nz)[:, None] * (nx + 1) + np.arange(nx)[None, :]).reshape(-1)
    c = a + nx + 1
    idx = np.empty((a.size * 2, 3), np.uint32)
    idx[0::2] = np.stack([a, c, a + 1], -1)
    idx[1::2] = np.stack([a + 1, c, c + 1], -1)
    if heights is None:
        nrm = np.tile(np.asarray([[0, 1, 0]], np.float32), (pos.shape[0], 1))
    else:
        nrm = vertex_normals(pos, idx)
    return Mesh(pos, nrm, idx, name="grid")


def vertex_normals(pos, idx):
    p = pos.astype(np.float64)
    f = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        for c in range(3):
            n[:, c] += np.bincount(idx[:, k], weights=f[:, c], minlength=p.shape[0])
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    unused = ln[:, 0] == 0
    ln[unused] = 1
    n = n / ln
    n[unused] = (0.0, 1.0, 0.0)  # vertices no triangle references
    return n.astype(np.float32)


# --------------------------------------------------------------- noise
def value_noise(p, seed, octaves=5, frequency=1.0, gain=0.5):
    """Fractal value noise in [-1, 1] at points p (N, 3): lattice values from
    the counter hash, smoothstep-trilinear interpolation, octaves summed."""
    p = np.asarray(p, np.float64)
    total = np.zeros(p.shape[0])
    amplitude, norm = 1.0, 0.0
    for o in range(octaves):
        q = p * frequency * (2.0 ** o)
        i = np.floor(q).astype(np.int64)
        f = q - i
        w = f * f * (3 - 2 * f)
        acc = np.zeros(p.shape[0])
        for corner in range(8):
            dx, dy, dz = corner & 1, (corner >> 1) & 1, (corner >> 2) & 1
            key = ((i[:, 0] + dx) * 73856093) ^ ((i[:, 1] + dy) * 19349663) ^ ((i[:, 2] + dz) * 83492791)
            h = hash_u32((key & 0xFFFFFFFF).astype(np.uint32) ^ np.uint32((seed * 2654435761 + o) & 0xFFFFFFFF))
            v = (h >> np.uint32(8)).astype(np.float64) * (2.0 / 16777216.0) - 1.0
            wx = w[:, 0] if dx else 1 - w[:, 0]
            wy = w[:, 1] if dy else 1 - w[:, 1]
            wz = w[:, 2] if dz else 1 - w[:, 2]
            acc += v * wx * wy * wz
        total += amplitude * acc
        norm += amplitude
        amplitude *= gain
    return total / norm


def icosphere(level):
    """Unit icosphere by midpoint subdivision (20 * 4**level triangles),
    vertices welded through an edge table so the mesh stays watertight."""
    t = (1 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    verts = [v]
    nv = v.shape[0]
    for _ in range(level):
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        e = np.concatenate([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)])
        lo, hi = np.minimum(e[:, 0], e[:, 1]), np.maximum(e[:, 0], e[:, 1])
        key = lo * (1 << 32) + hi
        uniq, inv = np.unique(key, return_inverse=True)
        ulo, uhi = uniq >> 32, uniq & 0xFFFFFFFF
        allv = np.concatenate(verts) if len(verts) > 1 else verts[0]
        mid = allv[ulo] + allv[uhi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        verts = [allv, mid]
        m = inv.reshape(3, -1) + nv
        nv += uniq.size
        ab, bc, ca = m[0], m[1], m[2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    allv = np.concatenate(verts) if len(verts) > 1 else verts[0]
    return allv, f.astype(np.uint32)


def icosphere_mesh(level, radius=1.0):
    v, f = icosphere(level)
    return Mesh((v * radius).astype(np.float32), v.astype(np.float32), f, name="icosphere")


# ------------------------------------------------------------ workloads
def _lambert(seed, k):
    u = uniform(seed, 1000 + k, 3)
    return dict(base_color=(0.2 + 0.6 * u).tolist(), base_metalness=0.0, specular_weight=0.0,
                specular_roughness=0.5, specular_ior=1.5)


def config0(seed=1, width=256, height=256, spp=16, max_depth=4):
    """Plain path tracer, CPU-runnable: a 100x100-quad ground grid spanning
    +-10 (20,000 triangles), a 64x32 UV sphere of radius 1 at y=1 (3,968
    triangles from the reference generator), a unit box (12 triangles),
    seeded Lambertian albedo U[0.2,0.8]^3 (specular weight 0), a 1x1
    rectangle light of radiance 10 at y=5 facing down, a constant sky of 0.1,
    a 45-degree pinhole camera, 256x256 at 16 spp, depth 4."""
    meshes = [grid((20.0, 20.0), (100, 100)), uv_sphere(1.0, 64, 32), box((1.0, 1.0, 1.0))]
    instances = [dict(transform=translate(0, 0, 0), blas=0, material=0),
                 dict(transform=translate(0, 1, 0), blas=1, material=1),
                 dict(transform=translate(-2.2, 0.5, 0.9), blas=2, material=2)]
    materials = [_lambert(seed, k) for k in range(3)]
    e = 10.0 * math.pi * 1.0  # radiance 10 over area 1 (emitter.slang:142: L = colour*energy/(pi*A))
    lights = [dict(position=(0.0, 5.0, 0.0), u=(1.0, 0.0, 0.0), v=(0.0, 0.0, 1.0), normal=(0.0, -1.0, 0.0),
                   radiance=(e, e, e))]
    camera = dict(eye=(0.0, 2.6, 7.0), target=(0.0, 0.8, 0.0))
    sky = dict(kind=abi.VH_SKY_CONSTANT, color=(0.1, 0.1, 0.1))
    return Scene(meshes, instances, materials, lights, camera, sky, width, height, spp, max_depth, seed, "config0",
                 notes=dict(sphere_triangles_reference_generator=3968, sphere_triangles_stated=4032))


def config1(seed=1, width=1920, height=1080, spp=64, max_depth=8, sphere_level=9, terrain=1024, materials=32):
    """Single large mesh: an icosphere subdivided to 20*4**9 = 5,242,880
    triangles with fractal value-noise radial displacement of amplitude 0.05,
    a 1024x1024-cell noise heightfield (2,097,152 triangles), GGX materials
    with roughness U[0.05,0.8] and metalness drawn from {0,1} with p=0.5, two
    rectangle lights and a procedural sky; 1920x1080, 64 spp, depth 8."""
    v, f = icosphere(sphere_level)
    r = 1.0 + 0.05 * value_noise(v * 2.0, seed, octaves=6, frequency=2.0)
    radius = 2.0
    pos = (v * (r * radius)[:, None] + np.array([0.0, 2.6, 0.0])).astype(np.float32)
    sphere = Mesh(pos, vertex_normals(pos, f), f, name="icosphere")
    # material patches: the level-2 ancestor face of every triangle
    anc = (np.arange(f.shape[0]) // (4 ** max(sphere_level - 2, 0))) if sphere_level >= 2 else np.arange(f.shape[0])
    sphere.triangle_materials = (hash_u32(anc.astype(np.uint32) ^ np.uint32(seed)) % np.uint32(materials // 2)).astype(np.uint32)
    n = terrain
    xs = np.linspace(-0.5, 0.5, n + 1) * 40.0
    X, Z = np.meshgrid(xs, xs)
    P = np.stack([X, np.zeros_like(X), Z], -1).reshape(-1, 3)
    H = (1.2 * value_noise(P * 0.08, seed + 7, octaves=6, frequency=1.0)).reshape(n + 1, n + 1)
    ground = grid((40.0, 40.0), (n, n), heights=H)
    cell = n // 16
    tz, tx = np.divmod(np.arange(n * n), n)
    patch = ((tz // cell) * 16 + tx // cell).astype(np.uint32)
    ground.triangle_materials = np.repeat(
        (materials // 2 + hash_u32(patch ^ np.uint32(seed * 3 + 1)) % np.uint32(materials - materials // 2)).astype(np.uint32), 2)
    mats = []
    for k in range(materials):
        u = uniform(seed, 2000 + k, 5)
        mats.append(dict(base_color=(0.2 + 0.6 * u[:3]).tolist(), base_metalness=1.0 if u[3] < 0.5 else 0.0,
                         specular_weight=1.0, specular_roughness=float(0.05 + 0.75 * u[4]), specular_ior=1.5))
    e1, e2 = 12.0 * math.pi * 4.0, 8.0 * math.pi * 2.0
    lights = [dict(position=(4.0, 9.0, 3.0), u=(2.0, 0.0, 0.0), v=(0.0, 0.0, 2.0), normal=(0.0, -1.0, 0.0),
                   radiance=(e1, e1 * 0.9, e1 * 0.8)),
              # tilted 11.3 degrees down; the sides stay perpendicular to the normal
              dict(position=(-7.0, 5.0, -2.0), u=(2.0 * 0.2 / math.sqrt(1.04), 2.0 / math.sqrt(1.04), 0.0),
                   v=(0.0, 0.0, 1.0), normal=(1.0, -0.2, 0.0), radiance=(e2 * 0.7, e2 * 0.8, e2))]
    camera = dict(eye=(0.0, 4.5, 11.0), target=(0.0, 2.0, 0.0))
    sky = dict(kind=abi.VH_SKY_PROCEDURAL, color=(0.45, 0.5, 0.6), zenith=(0.12, 0.25, 0.6), ground=(0.08, 0.07, 0.06),
               sun_direction=(0.3, 0.6, 0.25), sun_radiance=(6.0, 5.5, 4.5), sun_exponent=48.0)
    return Scene([sphere, ground], [dict(transform=translate(0, 0, 0), blas=0, material=0),
                                    dict(transform=translate(0, 0, 0), blas=1, material=0)],
                 mats, lights, camera, sky, width, height, spp, max_depth, seed, "config1")


def small_config1(seed=1, width=96, height=54, spp=16):
    """config1 at test size: the same generators, a level-4 icosphere and a
    64x64 terrain."""
    return config1(seed, width, height, spp, 8, sphere_level=4, terrain=64, materials=8)


# ------------------------------------------------------------ config 2
def bird(seed=1):
    """A 2,000-triangle bird: a UV-sphere body (32 x 16, 960 triangles)
    stretched along +z, and two 13 x 20-cell wings (520 triangles each) with
    a slight dihedral; about 2.2 units of wingspan."""
    body = uv_sphere(1.0, 32, 16)
    bp = body.positions.astype(np.float64) * np.array([0.22, 0.18, 0.6])
    parts_p, parts_i = [bp], [body.indices.astype(np.int64)]
    off = bp.shape[0]
    for side in (-1.0, 1.0):
        w = grid((1.0, 0.5), (13, 20))
        wp = w.positions.astype(np.float64)
        x = (wp[:, 0] + 0.5) * side * 1.0 + side * 0.12      # 0.12 .. 1.12 out from the body
        z = wp[:, 2] * (1.0 - 0.5 * (wp[:, 0] + 0.5))         # tapering chord
        y = 0.15 * np.abs(x) + 0.02
        q = np.stack([x, y, z], -1)
        idx = w.indices.astype(np.int64)
        if side < 0:
            idx = idx[:, [0, 2, 1]]
        parts_p.append(q)
        parts_i.append(idx + off)
        off += q.shape[0]
    pos = np.concatenate(parts_p).astype(np.float32)
    idx = np.concatenate(parts_i).astype(np.uint32)
    m = Mesh(pos, vertex_normals(pos, idx), idx, name="bird")
    m.triangle_materials = np.concatenate([np.zeros(960, np.uint32), np.ones(1040, np.uint32)])
    assert m.triangle_count == 2000
    return m


def curl_velocity(p, t):
    """Divergence-free velocity: the curl of a sum-of-sines vector potential,
    evaluated analytically (units per second)."""
    x, y, z = p[:, 0], p[:, 1], p[:, 2]
    a, b, c = 0.071, 0.23, 0.13
    # potential psi = (sin(b y + t), sin(c z + 0.7 t), sin(a x + 1.3 t)) * s
    s = np.array([6.0, 4.0, 5.0])
    dpz_dy = 0.0
    dpy_dz = s[1] * c * np.cos(c * z + 0.7 * t)
    dpx_dz = 0.0
    dpz_dx = s[2] * a * np.cos(a * x + 1.3 * t)
    dpy_dx = 0.0
    dpx_dy = s[0] * b * np.cos(b * y + t)
    return np.stack([dpz_dy - dpy_dz, dpx_dz - dpz_dx, dpy_dx - dpx_dy], -1) + np.array([4.0, 0.0, 0.0])


def flock_instances(positions, velocities, scales, seed=1):
    """Instance table of the flock: yaw from the velocity, uniform scale."""
    n = positions.shape[0]
    inst = np.zeros(n, abi.INSTANCE_DTYPE)
    yaw = np.arctan2(velocities[:, 0], velocities[:, 2])
    c, s = np.cos(yaw) * scales, np.sin(yaw) * scales
    tr = np.zeros((n, 12))
    tr[:, 0], tr[:, 2], tr[:, 3] = c, s, positions[:, 0]
    tr[:, 5], tr[:, 7] = scales, positions[:, 1]
    tr[:, 8], tr[:, 10], tr[:, 11] = -s, c, positions[:, 2]
    inst["transform"] = tr.astype(np.float32)
    inst["blas"] = 1
    inst["custom_index"] = np.arange(n, dtype=np.uint32) & 0xFFFFFF
    inst["material"] = 0
    inst["mask"] = 0xFF
    return inst


def flock_frames(scene, frames, dt=1.0 / 24.0, progress=None):
    """Instance tables of `frames` successive animation frames of a config2
    scene (frame 0 is the scene's own table): each frame advects the birds by
    the curl-noise velocity for `dt` and re-yaws them.  `progress(k)` is
    called after every frame made."""
    base = scene.instances
    birds = base[1:]
    tr = birds["transform"].astype(np.float64)
    pos = tr[:, [3, 7, 11]].copy()
    scales = tr[:, 5].copy()
    t = scene.notes.get("frame", 0) * dt
    out = [base.copy()]
    for _ in range(frames - 1):
        pos = pos + curl_velocity(pos, t) * dt
        t += dt
        f = base.copy()
        f[1:] = flock_instances(pos, curl_velocity(pos, t), scales)
        out.append(f)
        if progress:
            progress(len(out))
    return out


def config2(seed=1, width=1920, height=1080, spp=4, max_depth=3, birds=1_000_000, frame=0, dt=1.0 / 24.0):
    """Instanced flock: `birds` copies of the 2,000-triangle bird placed from a
    Gaussian ribbon with sigma (40, 3, 6), advected by curl noise for
    `frame` steps of `dt`, yawed along their velocity and scaled U[0.9, 1.1];
    a ground plane, one rectangle light and a procedural sky; 1920x1080,
    4 spp, depth 3."""
    g = np.random.default_rng(seed)
    pos = g.normal(size=(birds, 3)) * np.array([40.0, 3.0, 6.0]) + np.array([0.0, 22.0, 0.0])
    scales = g.uniform(0.9, 1.1, size=birds)
    t = 0.0
    for _ in range(frame):
        pos = pos + curl_velocity(pos, t) * dt
        t += dt
    vel = curl_velocity(pos, t)
    inst = flock_instances(pos, vel, scales, seed)
    ground = grid((600.0, 600.0), (4, 4))
    b = bird(seed)
    ground_inst = np.zeros(1, abi.INSTANCE_DTYPE)
    ground_inst["transform"] = np.asarray(translate(0, 0, 0), np.float32)
    ground_inst["blas"] = 0
    ground_inst["material"] = 2
    ground_inst["mask"] = 0xFF
    ground_inst["custom_index"] = 0xFFFFFF
    mats = [dict(base_color=(0.55, 0.5, 0.45), base_metalness=0.0, specular_weight=1.0, specular_roughness=0.6),
            dict(base_color=(0.2, 0.22, 0.25), base_metalness=0.0, specular_weight=1.0, specular_roughness=0.35),
            dict(base_color=(0.35, 0.4, 0.3), base_metalness=0.0, specular_weight=1.0, specular_roughness=0.9)]
    e = 40.0 * math.pi * 400.0
    lights = [dict(position=(-10.0, 80.0, 20.0), u=(20.0, 0.0, 0.0), v=(0.0, 0.0, 20.0), normal=(0.0, -1.0, 0.0),
                   radiance=(e, e * 0.95, e * 0.9))]
    camera = dict(eye=(0.0, 26.0, 130.0), target=(0.0, 21.0, 0.0))
    sky = dict(kind=abi.VH_SKY_PROCEDURAL, color=(0.45, 0.5, 0.6), zenith=(0.12, 0.25, 0.6), ground=(0.08, 0.07, 0.06),
               sun_direction=(0.3, 0.6, 0.25), sun_radiance=(6.0, 5.5, 4.5), sun_exponent=48.0)
    return Scene([ground, b], np.concatenate([ground_inst, inst]), mats, lights, camera, sky, width, height, spp,
                 max_depth, seed, "config2", notes=dict(frame=frame, birds=birds))
