"""scene.json (vork-scene v2) -> device scene: the subset the path workloads use.

Reads what `vork render SCENE` reads for the std/path hot path and turns it
into a `vork_amd.scene.Scene` (meshes, instances, materials, rectangle lights,
camera, environment) that `native.Renderer.install` puts on the device:

* nodes: glTF TRS (`translation`, `rotation` as [x,y,z,w] or {euler}, `scale`)
  composed parent-to-child (crates/vork-scene/src/v2/node.rs:395-458), `target`
  aiming -Z with world Y up and the rotation as an offset on top
  (node.rs:48-50);
* meshes: named or inline geometry with POSITION / NORMAL / indices buffers,
  inline `data` or little-endian `file` + `offset` + `count`
  (geometry.rs:31-60, 83-125), polygon `face_vertex_counts` fan-split,
  `material_ranges` into the node's material slots, GPU instancing through
  `instances.attributes` TRANSLATION / ROTATION / SCALE (geometry.rs:170-180);
  one BLAS per distinct geometry, one TLAS instance per placement;
* materials: std/openpbr inputs base_weight x base_color, base_metalness,
  specular_weight, specular_roughness, specular_ior, with the OpenPBR 1.1
  defaults and roughness 0.3 for a mesh without materials (node.rs:180-186);
* lights: `rectangle` (light record of crates/vork-render/src/lights.rs:55-140:
  centre, sides along the node's X/Y scaled by size / size_y, normal -Z,
  radiance = colour x energy, default energy 10 W, colour white) and
  `environment` (constant colour x energy; no environment light = RGB 0.05,
  crates/vork-scene/src/v2/light.rs:1-30, constants.rs:63-70);
* camera: std/camera perspective inputs lens, sensor_width/height, fit,
  clip_start/end (libraries/std/camera.slang:4-44), the node named by
  `render.camera` or the first camera;
* render: `resolution`; render-call fields resolution=WxH, samples, seed
  (crates/vork-scene/src/v2/profile.rs:15-27; samples default 64 for
  std/path, at most 4096).

Anything else a scene asks for that this build's hot path does not render
(other light shapes, orthographic or depth-of-field cameras, textures,
emission, coat, transmission, subsurface, curves, points, volumes) is refused
with a scene error naming the field, never silently dropped.
"""
import json
import math
import os

import numpy as np

from . import abi
from .scene import Mesh, Scene, vertex_normals

PATH_SAMPLES = 64            # crates/vork-scene/src/v2/constants.rs:11
PROFILE_DENOISE = True       # libraries/std/profiles/path.json:6
MAX_RENDER_SAMPLES = 4096    # constants.rs:67
RESOLUTION = (1920, 1080)    # constants.rs:28
LOCAL_INTENSITY = 10.0       # constants.rs:53
ENVIRONMENT_INTENSITY = 1.0  # constants.rs:63
ENVIRONMENT_COLOR = (0.05, 0.05, 0.05)  # constants.rs:70
AREA_DIMENSION = 1.0         # constants.rs:46

# OpenPBR Surface 1.1 defaults for the inputs a scene may set to a value the
# closure subset renders; any other input must keep its default.
OPENPBR_RENDERED = {"base_weight": 1.0, "base_color": (0.8, 0.8, 0.8), "base_metalness": 0.0,
                    "specular_weight": 1.0, "specular_roughness": 0.3, "specular_ior": 1.5}
OPENPBR_NEUTRAL = {"base_diffuse_roughness": 0.0, "specular_color": (1.0, 1.0, 1.0),
                   "specular_anisotropy": 0.0, "specular_rotation": 0.0, "transmission_weight": 0.0,
                   "subsurface_weight": 0.0, "coat_weight": 0.0, "fuzz_weight": 0.0, "emission_luminance": 0.0,
                   "thin_film_weight": 0.0, "geometry_opacity": 1.0, "geometry_thin_walled": False}
# std/camera inputs and defaults (libraries/std/camera.slang:6-22); the
# orthographic and aperture inputs only matter with projection ortho / use_dof
CAMERA_INPUTS = {"projection": "persp", "fit": "auto", "lens": 50.0, "sensor_width": 36.0, "sensor_height": 24.0,
                 "ortho_scale": 6.0, "shift": (0.0, 0.0), "clip_start": 0.1, "clip_end": 1000.0, "use_dof": False,
                 "focus_distance": 10.0, "aperture_fstop": 2.8, "aperture_blades": 0, "aperture_rotation": 0.0,
                 "aperture_ratio": 1.0}

ELEMENT = {"f32": (np.float32, 1), "i32": (np.int32, 1), "u32": (np.uint32, 1), "vec2": (np.float32, 2),
           "vec3": (np.float32, 3), "vec4": (np.float32, 4), "uvec4": (np.uint32, 4), "mat4": (np.float32, 16)}


class SceneError(RuntimeError):
    """A scene the loader cannot turn into the hot path's input; `status` is
    the reference's scene failure code (3)."""
    status = abi.VH_ERR_SCENE

    def __init__(self, where, message):
        super().__init__(f"{where}: {message}")
        self.where = where


# ----------------------------------------------------------------- maths
def quat_matrix(q):
    x, y, z, w = (float(v) for v in q)
    n = math.sqrt(x * x + y * y + z * z + w * w)
    if not n > 1e-8:
        raise ValueError("zero-length rotation")
    x, y, z, w = x / n, y / n, z / n, w / n
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def euler_matrix(e):
    """Blender XYZ Euler: R = Rz * Ry * Rx (node.rs:420-428)."""
    x, y, z = (float(v) for v in e)
    cx, sx, cy, sy, cz, sz = math.cos(x), math.sin(x), math.cos(y), math.sin(y), math.cos(z), math.sin(z)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def aim_matrix(eye, target):
    """Rotation whose -Z looks from eye to target with world Y up."""
    f = np.asarray(target, np.float64) - np.asarray(eye, np.float64)
    n = np.linalg.norm(f)
    if not n > 0:
        raise ValueError("target coincides with the node")
    f /= n
    up = np.array([0.0, 1.0, 0.0])
    r = np.cross(f, up)
    if np.linalg.norm(r) < 1e-12:
        r = np.array([1.0, 0.0, 0.0])
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return np.stack([r, u, -f], 1)


def _num(v, where, n=None):
    try:
        a = np.asarray(v, np.float64)
    except (TypeError, ValueError):
        raise SceneError(where, "expected a number or a numeric vector")
    if n is not None and a.shape != (n,):
        raise SceneError(where, f"expected {n} numbers")
    if not np.isfinite(a).all():
        raise SceneError(where, "non-finite value")
    return a


def _scalar(v, where):
    a = _num(v, where)
    if a.shape != ():
        raise SceneError(where, "expected a number")
    return float(a)


# ---------------------------------------------------------------- loader
class _Loader:
    def __init__(self, doc, base_dir, overrides):
        self.doc = doc
        self.base_dir = base_dir
        self.overrides = dict(overrides or {})
        self.warnings = []
        self.buffers = doc.get("buffers", {}) or {}
        self.geometry = doc.get("geometry", {}) or {}
        self.materials_table = doc.get("materials", {}) or {}
        self.meshes, self.mesh_key = [], {}
        self.instances = []
        self.materials = []
        self.lights = []
        self.cameras = []
        self.environment = None
        self.names = set()

    # buffers ----------------------------------------------------------
    def buffer(self, ref, where, want=None):
        spec = self.buffers.get(ref) if isinstance(ref, str) else ref
        if isinstance(ref, str) and spec is None:
            raise SceneError(where, f"unknown buffer '{ref}'")
        if not isinstance(spec, dict) or "type" not in spec:
            raise SceneError(where, "a buffer is {type, data} or {type, file, offset, count}")
        t = spec["type"]
        if t not in ELEMENT:
            raise SceneError(where, f"unknown buffer type '{t}'")
        if want is not None and t not in want:
            raise SceneError(where, f"buffer type {t}, expected {'/'.join(want)}")
        dtype, k = ELEMENT[t]
        if ("data" in spec) == ("file" in spec):
            raise SceneError(where, "a buffer has exactly one of data and file")
        if "data" in spec:
            a = np.asarray(spec["data"], dtype=np.float64)
            if k == 1:
                a = a.reshape(-1)
            elif a.ndim != 2 or a.shape[1] != k:
                raise SceneError(where, f"{t} entries have {k} components")
            if dtype != np.float32 and (a < 0).any() and dtype == np.uint32:
                raise SceneError(where, "negative value in an unsigned buffer")
            return a.astype(dtype)
        path = os.path.join(self.base_dir, spec["file"])
        offset = int(spec.get("offset", 0))
        if offset % 4:
            raise SceneError(where, "buffer offset must be a multiple of 4")
        if "count" not in spec:
            raise SceneError(where, "a file buffer needs count")
        count = int(spec["count"])
        try:
            raw = np.fromfile(path, dtype=np.dtype(dtype).newbyteorder("<"), count=count * k, offset=offset)
        except OSError as e:
            raise SceneError(where, f"cannot read '{spec['file']}': {e}")
        if raw.size != count * k:
            raise SceneError(where, f"'{spec['file']}' holds fewer than {count} {t} elements at offset {offset}")
        raw = raw.astype(dtype)
        return raw if k == 1 else raw.reshape(count, k)

    # geometry ---------------------------------------------------------
    def mesh(self, ref, where):
        key = ref if isinstance(ref, str) else id(ref)
        if key in self.mesh_key:
            return self.mesh_key[key]
        g = self.geometry.get(ref) if isinstance(ref, str) else ref
        if g is None:
            raise SceneError(where, f"unknown geometry '{ref}'")
        if g.get("topology", "triangles") != "triangles":
            raise SceneError(where, f"topology '{g.get('topology')}' is not rendered by this build")
        for k in ("subdivision", "morphs", "curve_vertex_counts", "widths", "count"):
            if g.get(k):
                raise SceneError(f"{where}.{k}", "not rendered by this build")
        attrs = g.get("attributes", {})
        if "POSITION" not in attrs:
            raise SceneError(where, "geometry without POSITION")
        pos = self.buffer(attrs["POSITION"], f"{where}.attributes.POSITION", ("vec3",)).astype(np.float32)
        if "indices" in g and g["indices"] is not None:
            idx = self.buffer(g["indices"], f"{where}.indices", ("u32",)).astype(np.int64)
        else:
            idx = np.arange(pos.shape[0], dtype=np.int64)
        counts = g.get("face_vertex_counts")
        if counts is not None:
            counts = np.asarray(counts, np.int64)
            if (counts < 3).any() or counts.sum() != idx.size:
                raise SceneError(f"{where}.face_vertex_counts", "polygon corner counts must be >= 3 and sum to the index count")
            tris = []
            start = 0
            for c in counts:  # fan split of each polygon
                for j in range(1, int(c) - 1):
                    tris.append((idx[start], idx[start + j], idx[start + j + 1]))
                start += int(c)
            tri = np.asarray(tris, np.int64).reshape(-1, 3)
        else:
            if idx.size % 3:
                raise SceneError(f"{where}.indices", "triangle indices come in threes")
            tri = idx.reshape(-1, 3)
        if tri.size and (tri.min() < 0 or tri.max() >= pos.shape[0]):
            raise SceneError(f"{where}.indices", "index out of range of POSITION")
        if "NORMAL" in attrs:
            nrm = self.buffer(attrs["NORMAL"], f"{where}.attributes.NORMAL", ("vec3",)).astype(np.float32)
            if nrm.shape != pos.shape:
                raise SceneError(f"{where}.attributes.NORMAL", "NORMAL and POSITION differ in length")
        else:
            nrm = vertex_normals(pos, tri.astype(np.uint32))
        m = Mesh(pos, nrm, tri.astype(np.uint32), name=ref if isinstance(ref, str) else "inline")
        slots = np.zeros(tri.shape[0], np.uint32)
        for r in g.get("material_ranges", []) or []:
            first = int(r["start"]) // 3
            slots[first:first + int(r["count"])] = int(r["material_index"])
        m.triangle_materials = slots
        self.meshes.append(m)
        self.mesh_key[key] = len(self.meshes) - 1
        return self.mesh_key[key]

    # materials --------------------------------------------------------
    def material(self, ref, where):
        spec = self.materials_table.get(ref) if isinstance(ref, str) else ref
        if spec is None:
            raise SceneError(where, f"unknown material '{ref}'")
        program = spec.get("program", "std/openpbr")
        if program != "std/openpbr":
            raise SceneError(f"{where}.program", f"program '{program}' is not rendered by this build")
        inputs = spec.get("inputs", {}) or {}
        vals = dict(OPENPBR_RENDERED)
        for k, v in inputs.items():
            w = f"{where}.inputs.{k}"
            if isinstance(v, dict):
                raise SceneError(w, "textures, attributes and other resource inputs are not rendered by this build")
            if k in OPENPBR_RENDERED:
                vals[k] = _num(v, w, 3 if k == "base_color" else None)
            elif k in OPENPBR_NEUTRAL:
                d = OPENPBR_NEUTRAL[k]
                same = (v == d) if isinstance(d, bool) else np.allclose(_num(v, w), d)
                if not same:
                    raise SceneError(w, f"only the neutral value {d} is rendered by this build")
            else:
                raise SceneError(w, "input outside the closure subset this build renders")
        color = np.asarray(vals["base_color"], np.float64) * float(vals["base_weight"])
        return dict(base_color=tuple(float(c) for c in color), base_metalness=float(vals["base_metalness"]),
                    specular_weight=float(vals["specular_weight"]),
                    specular_roughness=float(vals["specular_roughness"]), specular_ior=float(vals["specular_ior"]))

    # nodes ------------------------------------------------------------
    def local(self, node, where, parent):
        t = _num(node.get("translation", (0, 0, 0)), f"{where}.translation", 3)
        s = _num(node.get("scale", (1, 1, 1)), f"{where}.scale", 3)
        rot = node.get("rotation")
        try:
            if rot is None:
                r = np.eye(3)
            elif isinstance(rot, dict):
                r = euler_matrix(_num(rot.get("euler"), f"{where}.rotation.euler", 3))
            else:
                r = quat_matrix(_num(rot, f"{where}.rotation", 4))
        except ValueError as e:
            raise SceneError(f"{where}.rotation", str(e))
        m = np.eye(4)
        m[:3, :3] = r * s[None, :]
        m[:3, 3] = t
        world = parent @ m
        target = node.get("target")
        if target is not None:
            eye = world[:3, 3]
            if isinstance(target, str):
                raise SceneError(f"{where}.target", "named targets are not supported by this loader; give a point")
            try:
                aim = aim_matrix(eye, _num(target, f"{where}.target", 3))
            except ValueError as e:
                raise SceneError(f"{where}.target", str(e))
            world = world.copy()
            world[:3, :3] = (aim @ r) * s[None, :]
        return world

    def walk(self, nodes, parent, prefix):
        for k, node in enumerate(nodes):
            name = node.get("name") or f"{prefix}[{k}]"
            where = f"nodes.{name}"
            if node.get("name"):
                if name in self.names:
                    raise SceneError(where, "duplicate node name")
                self.names.add(name)
            if node.get("active", True) is False:
                continue
            kind = node.get("type", "node")
            world = self.local(node, where, parent)
            if kind == "mesh":
                self.mesh_node(node, where, world)
            elif kind == "light":
                self.light_node(node, where, world)
            elif kind == "camera":
                self.cameras.append((name, node, world))
            elif kind != "node":
                raise SceneError(where, f"node type '{kind}' is not rendered by this build")
            self.walk(node.get("children", []) or [], world, name)

    def mesh_node(self, node, where, world):
        if node.get("skin"):
            raise SceneError(f"{where}.skin", "skinning is not rendered by this build")
        blas = self.mesh(node["geometry"], f"{where}.geometry")
        mats = node.get("materials") or []
        base = len(self.materials)
        if not mats:
            self.materials.append(self.material({}, f"{where}.materials(default)"))
        for j, mref in enumerate(mats):
            self.materials.append(self.material(mref, f"{where}.materials[{j}]"))
        slots = max(1, len(mats))
        if int(self.meshes[blas].triangle_materials.max(initial=0)) >= slots:
            raise SceneError(f"{where}.materials", "a material range names a slot the node does not have")
        placements = [world]
        inst = node.get("instances")
        if inst:
            a = inst.get("attributes", {}) or {}
            n = None
            cols = {}
            for key, want in (("TRANSLATION", "vec3"), ("ROTATION", "vec4"), ("SCALE", "vec3")):
                if key in a:
                    cols[key] = self.buffer(a[key], f"{where}.instances.attributes.{key}", (want,)).astype(np.float64)
                    if n is not None and cols[key].shape[0] != n:
                        raise SceneError(f"{where}.instances", "instance attributes differ in length")
                    n = cols[key].shape[0]
            if n is None:
                raise SceneError(f"{where}.instances", "instances without attributes")
            placements = self.instance_matrices(world, cols, n)
        for m in placements:
            self.instances.append((m, blas, base))

    @staticmethod
    def instance_matrices(world, cols, n):
        t = cols.get("TRANSLATION", np.zeros((n, 3)))
        s = cols.get("SCALE", np.ones((n, 3)))
        q = cols.get("ROTATION", np.tile([0.0, 0.0, 0.0, 1.0], (n, 1)))
        q = q / np.linalg.norm(q, axis=1, keepdims=True)
        x, y, z, w = q.T
        r = np.empty((n, 3, 3))
        r[:, 0, 0], r[:, 0, 1], r[:, 0, 2] = 1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)
        r[:, 1, 0], r[:, 1, 1], r[:, 1, 2] = 2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)
        r[:, 2, 0], r[:, 2, 1], r[:, 2, 2] = 2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)
        m = np.zeros((n, 4, 4))
        m[:, :3, :3] = r * s[:, None, :]
        m[:, :3, 3] = t
        m[:, 3, 3] = 1
        return np.einsum("ij,njk->nik", world, m)

    def light_node(self, node, where, world):
        shape = node.get("shape")
        for k in ("temperature", "image", "medium", "illuminates", "shadowed_by"):
            if node.get(k) is not None:
                raise SceneError(f"{where}.{k}", "not rendered by this build")
        energy = _scalar(node["energy"], f"{where}.energy") if node.get("energy") is not None else None
        if shape == "environment":
            if node.get("material") is not None:
                raise SceneError(f"{where}.material", "environment programs are not rendered by this build")
            if self.environment is not None:
                raise SceneError(where, "more than one environment light")
            color = _num(node.get("color", ENVIRONMENT_COLOR), f"{where}.color", 3)
            e = ENVIRONMENT_INTENSITY if energy is None else energy
            self.environment = tuple(float(c) * e for c in color)
            return
        if shape != "rectangle":
            raise SceneError(f"{where}.shape", f"light shape '{shape}' is not rendered by this build (rectangle, environment)")
        width = _scalar(node.get("size", AREA_DIMENSION), f"{where}.size")
        height = _scalar(node["size_y"], f"{where}.size_y") if node.get("size_y") is not None else width
        color = _num(node.get("color", (1.0, 1.0, 1.0)), f"{where}.color", 3)
        e = LOCAL_INTENSITY if energy is None else energy
        axes = world[:3, :3]
        scale = np.linalg.norm(axes, axis=0)
        frame = axes / np.where(scale > 0, scale, 1.0)[None, :]
        u = frame[:, 0] * scale[0] * width
        v = frame[:, 1] * scale[1] * height
        self.lights.append(dict(position=tuple(world[:3, 3]), u=tuple(u), v=tuple(v), normal=tuple(-frame[:, 2]),
                                radiance=tuple(float(c) * e for c in color)))

    def camera(self, aspect):
        want = self.doc.get("render", {}).get("camera")
        chosen = None
        for name, node, world in self.cameras:
            if want is None or name == want:
                chosen = (name, node, world)
                break
        if chosen is None:
            raise SceneError("render.camera", f"no camera node{' named ' + repr(want) if want else ''}")
        name, node, world = chosen
        where = f"nodes.{name}.material"
        mat = node.get("material") or {}
        if isinstance(mat, str):
            mat = self.materials_table.get(mat, {})
        if mat.get("program", "std/camera") != "std/camera":
            raise SceneError(f"{where}.program", "only std/camera is rendered by this build")
        inputs = dict(CAMERA_INPUTS)
        for k, v in (mat.get("inputs") or {}).items():
            if k not in CAMERA_INPUTS:
                raise SceneError(f"{where}.inputs.{k}", "camera input not rendered by this build")
            inputs[k] = v
        if inputs["projection"] != "persp":
            raise SceneError(f"{where}.inputs.projection", "only perspective projection is rendered by this build")
        if inputs["use_dof"] or tuple(inputs["shift"]) != (0.0, 0.0):
            raise SceneError(f"{where}.inputs", "depth of field and lens shift are not rendered by this build")
        if node.get("focus_object"):
            raise SceneError(f"nodes.{name}.focus_object", "depth of field is not rendered by this build")
        axes = world[:3, :3]
        fwd = -axes[:, 2] / np.linalg.norm(axes[:, 2])
        up = axes[:, 1] / np.linalg.norm(axes[:, 1])
        cam = dict(eye=tuple(world[:3, 3]), forward=tuple(fwd), up=tuple(up), lens=float(inputs["lens"]),
                   sensor_width=float(inputs["sensor_width"]), sensor_height=float(inputs["sensor_height"]),
                   fit=str(inputs["fit"]), clip_start=float(inputs["clip_start"]), clip_end=float(inputs["clip_end"]))
        if node.get("target") is not None:
            cam["target"] = tuple(float(v) for v in node["target"])
        return cam


def parse_set(items):
    """Render-call `--set KEY=VALUE` fields: resolution WxH, samples, seed,
    fps, denoise (crates/vork-cli/src/main.rs:149, profile.rs:15-27)."""
    out = {}
    for item in items or []:
        key, _, value = item.partition("=")
        if key == "resolution":
            w, _, h = value.lower().partition("x")
            out[key] = (int(w), int(h))
        elif key in ("samples", "seed"):
            out[key] = int(value)
        elif key == "fps":
            out[key] = float(value)
        elif key == "denoise":
            out[key] = value.lower() in ("1", "true", "yes")
        else:
            raise SceneError(f"--set {key}", "allowed fields: resolution, fps, seed, samples, denoise")
    return out


def load(source, overrides=None, base_dir=None, max_depth=8):
    """scene.json (a path, a JSON string or a parsed dict) -> (Scene, warnings).

    `overrides` holds render-call fields (`parse_set`).  Raises SceneError for
    a scene outside the rendered subset."""
    if isinstance(source, dict):
        doc, base = source, base_dir or "."
    elif isinstance(source, str) and source.lstrip().startswith("{"):
        doc, base = json.loads(source), base_dir or "."
    else:
        with open(source) as f:
            doc = json.load(f)
        base = base_dir or os.path.dirname(os.path.abspath(source))
    if doc.get("version") != 2:
        raise SceneError("version", "scene.json version 2 is required")
    L = _Loader(doc, base, overrides)
    L.walk(doc.get("nodes", []) or [], np.eye(4), "nodes")
    if not L.instances:
        raise SceneError("nodes", "the scene has no mesh to render")
    ov = L.overrides
    render = doc.get("render", {}) or {}
    if render.get("motion_blur"):
        raise SceneError("render.motion_blur", "not rendered by this build")
    if render.get("aovs"):
        L.warnings.append("render.aovs: AOV outputs are not written by this build (the beauty pass is)")
    w, h = ov.get("resolution", tuple(render.get("resolution", RESOLUTION)))
    samples = int(ov.get("samples", PATH_SAMPLES))
    if not 1 <= samples <= MAX_RENDER_SAMPLES:
        raise SceneError("--set samples", f"1..{MAX_RENDER_SAMPLES}")
    camera = L.camera(w / h)
    n = len(L.instances)
    inst = np.zeros(n, abi.INSTANCE_DTYPE)
    for k, (m, blas, base_mat) in enumerate(L.instances):
        inst[k]["transform"] = np.asarray(m[:3, :4], np.float64).reshape(-1).astype(np.float32)
        inst[k]["blas"] = blas
        inst[k]["custom_index"] = k & 0xFFFFFF
        inst[k]["material"] = base_mat
        inst[k]["mask"] = 0xFF
    sky = dict(kind=abi.VH_SKY_CONSTANT, color=L.environment if L.environment is not None else ENVIRONMENT_COLOR)
    sc = Scene(L.meshes, inst, L.materials, L.lights, camera, sky, int(w), int(h), samples, max_depth,
               int(ov.get("seed", 0)), doc.get("name", "scene"))
    # the std/path profile denoises (libraries/std/profiles/path.json); --set denoise overrides
    sc.notes["denoise"] = bool(ov.get("denoise", PROFILE_DENOISE))
    return sc, L.warnings


# ---------------------------------------------------------------- writer
def _matrix_trs(m):
    """3x4 rows -> (translation, quaternion xyzw, scale) for a rotation x scale."""
    m = np.asarray(m, np.float64).reshape(3, 4)
    a = m[:, :3]
    s = np.linalg.norm(a, axis=0)
    if np.linalg.det(a) < 0:
        s[0] = -s[0]
    r = a / s[None, :]
    t = np.trace(r)
    if t > 0:
        k = math.sqrt(t + 1.0) * 2
        q = ((r[2, 1] - r[1, 2]) / k, (r[0, 2] - r[2, 0]) / k, (r[1, 0] - r[0, 1]) / k, 0.25 * k)
    elif r[0, 0] > r[1, 1] and r[0, 0] > r[2, 2]:
        k = math.sqrt(1.0 + r[0, 0] - r[1, 1] - r[2, 2]) * 2
        q = (0.25 * k, (r[0, 1] + r[1, 0]) / k, (r[0, 2] + r[2, 0]) / k, (r[2, 1] - r[1, 2]) / k)
    elif r[1, 1] > r[2, 2]:
        k = math.sqrt(1.0 + r[1, 1] - r[0, 0] - r[2, 2]) * 2
        q = ((r[0, 1] + r[1, 0]) / k, 0.25 * k, (r[1, 2] + r[2, 1]) / k, (r[0, 2] - r[2, 0]) / k)
    else:
        k = math.sqrt(1.0 + r[2, 2] - r[0, 0] - r[1, 1]) * 2
        q = ((r[0, 2] + r[2, 0]) / k, (r[1, 2] + r[2, 1]) / k, 0.25 * k, (r[1, 0] - r[0, 1]) / k)
    return m[:, 3].tolist(), list(q), s.tolist()


def dump(scene, path, inline_limit=4096):
    """Writes a Scene as scene.json (+ `<stem>.bin` for large buffers) in the
    subset `load` reads: a mesh node per instance (or one instanced mesh node
    per BLAS when the instances share a material base), rectangle lights as
    nodes aimed by rotation, the camera with its target."""
    stem = os.path.splitext(os.path.basename(path))[0]
    binname = stem + ".bin"
    blob = bytearray()
    buffers, geometry, materials = {}, {}, {}

    def put(name, arr, typ):
        arr = np.ascontiguousarray(arr)
        if arr.size <= inline_limit:
            buffers[name] = {"type": typ, "data": arr.tolist()}
        else:
            off = len(blob)
            blob.extend(arr.astype("<f4" if arr.dtype.kind == "f" else "<u4").tobytes())
            while len(blob) % 4:
                blob.append(0)
            count = arr.shape[0] if arr.ndim > 1 else arr.size
            buffers[name] = {"type": typ, "file": binname, "offset": off, "count": int(count)}
        return name

    for k, m in enumerate(scene.meshes):
        g = {"attributes": {"POSITION": put(f"m{k}_position", m.positions.astype(np.float32), "vec3"),
                            "NORMAL": put(f"m{k}_normal", m.normals.astype(np.float32), "vec3")},
             "indices": put(f"m{k}_indices", m.indices.reshape(-1).astype(np.uint32), "u32")}
        tm = m.triangle_materials
        if tm is not None and tm.size and tm.max() > 0:
            ranges, start = [], 0
            for j in range(1, tm.size + 1):
                if j == tm.size or tm[j] != tm[start]:
                    ranges.append({"start": 3 * start, "count": j - start, "material_index": int(tm[start])})
                    start = j
            g["material_ranges"] = ranges
        geometry[f"mesh{k}"] = g
    for k, mat in enumerate(scene.materials):
        materials[f"material{k}"] = {"program": "std/openpbr", "inputs": {
            "base_color": [float(c) for c in mat["base_color"]],
            "base_metalness": float(mat.get("base_metalness", 0.0)),
            "specular_weight": float(mat.get("specular_weight", 1.0)),
            "specular_roughness": float(mat.get("specular_roughness", 0.5)),
            "specular_ior": float(mat.get("specular_ior", 1.5))}}
    inst = scene.instances if isinstance(scene.instances, np.ndarray) else None
    nodes = []
    if inst is None:
        for k, i in enumerate(scene.instances):
            slots = int(scene.meshes[i["blas"]].triangle_materials.max(initial=0)) + 1 \
                if scene.meshes[i["blas"]].triangle_materials is not None else 1
            t, q, s = _matrix_trs(i["transform"])
            nodes.append({"name": f"instance{k}", "type": "mesh", "geometry": f"mesh{i['blas']}",
                          "materials": [f"material{i.get('material', 0) + j}" for j in range(slots)],
                          "translation": t, "rotation": q, "scale": s})
    else:
        for b in np.unique(inst["blas"]):
            rows = inst[inst["blas"] == b]
            for base in np.unique(rows["material"]):
                sel = rows[rows["material"] == base]
                tm = scene.meshes[b].triangle_materials
                slots = int(tm.max(initial=0)) + 1 if tm is not None else 1
                mats = [f"material{int(base) + j}" for j in range(slots)]
                if sel.shape[0] == 1:
                    t, q, s = _matrix_trs(sel[0]["transform"])
                    nodes.append({"name": f"mesh{b}_m{base}", "type": "mesh", "geometry": f"mesh{b}", "materials": mats,
                                  "translation": t, "rotation": q, "scale": s})
                    continue
                trs = [_matrix_trs(x) for x in sel["transform"]]
                nodes.append({"name": f"mesh{b}_m{base}", "type": "mesh", "geometry": f"mesh{b}", "materials": mats,
                              "instances": {"attributes": {
                                  "TRANSLATION": put(f"i{b}_{base}_t", np.asarray([x[0] for x in trs], np.float32), "vec3"),
                                  "ROTATION": put(f"i{b}_{base}_r", np.asarray([x[1] for x in trs], np.float32), "vec4"),
                                  "SCALE": put(f"i{b}_{base}_s", np.asarray([x[2] for x in trs], np.float32), "vec3")}}})
    for k, l in enumerate(scene.lights):
        u, v = np.asarray(l["u"], np.float64), np.asarray(l["v"], np.float64)
        su, sv = np.linalg.norm(u), np.linalg.norm(v)
        n = np.asarray(l["normal"], np.float64)
        n /= np.linalg.norm(n)
        if np.linalg.det(np.stack([u, v, -n], 1)) < 0:
            v = -v  # the same centred rectangle, as a proper rotation of the node frame
        r = np.stack([u / su, v / sv, -n], 1)
        t, q, _ = _matrix_trs(np.concatenate([r, np.asarray(l["position"], np.float64)[:, None]], 1).reshape(-1))
        rad = np.asarray(l["radiance"], np.float64)
        e = float(rad.max()) if rad.max() > 0 else 1.0
        nodes.append({"name": f"light{k}", "type": "light", "shape": "rectangle", "size": su, "size_y": sv,
                      "translation": t, "rotation": q, "color": (rad / e).tolist(), "energy": e})
    sky = scene.sky
    if sky.get("kind", abi.VH_SKY_CONSTANT) != abi.VH_SKY_CONSTANT:
        raise SceneError("sky", "only a constant environment is written as scene.json")
    nodes.append({"name": "environment", "type": "light", "shape": "environment", "color": list(sky["color"]),
                  "energy": 1.0})
    cam = scene.camera
    cnode = {"name": "camera", "type": "camera", "translation": list(cam["eye"]),
             "material": {"program": "std/camera", "inputs": {
                 "lens": float(cam.get("lens", 18.0 / math.tan(math.radians(22.5)))),
                 "sensor_width": float(cam.get("sensor_width", 36.0)),
                 "clip_start": float(cam.get("clip_start", 0.1)), "clip_end": float(cam.get("clip_end", 1000.0))}}}
    if "target" in cam:
        cnode["target"] = list(cam["target"])
    nodes.append(cnode)
    doc = {"version": 2, "name": scene.name, "render": {"resolution": [scene.width, scene.height]},
           "nodes": nodes, "buffers": buffers, "geometry": geometry, "materials": materials}
    if blob:
        with open(os.path.join(os.path.dirname(os.path.abspath(path)), binname), "wb") as f:
            f.write(bytes(blob))
    with open(path, "w") as f:
        json.dump(doc, f)
    return doc
