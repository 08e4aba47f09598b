"""`python -m vork_amd.session`: the kept-alive render worker of
crates/vork-cli/src/session.rs, driven by JSON lines on stdin and stdout.

One request per line, one reply per line, in order:

  {"id": 1, "op": "open", "scene": "scene.json"}
  {"id": 2, "op": "open", "document": {...}, "base": "dir"}
  {"id": 3, "op": "render", "out": "a.png", "set": {"samples": 16}}
  {"id": 4, "op": "close"}

The device and the open document's installation (BLASes, TLAS and the scene
tables, resident in HBM) live for the whole session; editing is `open` again
with the whole document, compared with the open one (worker.rs:176-292):

  reused       the decoded scene is identical: the installation is kept;
  updated      only values outside the acceleration structures changed
               (lights, materials, camera, sky): the kept BLASes and TLAS get
               new scene tables;
  reinstalled  geometry or instances changed (or nothing was installed): the
               acceleration structures are rebuilt on the kept device.

`changed` lists the authored paths at which the documents differ (the rules of
worker.rs:652-732) and `["file", relative]` for a changed external buffer.  An
open that fails keeps the previous revision, which renders on.  Failures reply
{"id", "ok": false, "kind": "usage" | "scene" | "io" | "device", "error"}
(session.rs:101-104).  A render through the session writes the same bytes as
the one-shot `python -m vork_amd.render` of the same document and options.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

REQUESTS = ("requests are open {scene | document, base}, render {out, at, profile, set, channels, provisional}, "
            "ask {questions} and close")
RENDER_FIELDS = ("out", "at", "frame_start", "frame_end", "profile", "set", "channels", "exr_codec", "png_depth",
                 "no_write", "gpu_only", "diagnostics", "provisional")
OPEN_FIELDS = ("scene", "document", "base")
SET_KEYS = ("resolution", "fps", "seed", "samples", "denoise")
PROFILE = "std/path"


class Failure(Exception):
    def __init__(self, kind, message):
        super().__init__(message)
        self.kind = kind


def _usage(msg):
    return Failure("usage", msg)


# ------------------------------------------------------------------ diff
def _name(item):
    n = item.get("name") if isinstance(item, dict) else None
    return n if isinstance(n, str) and n else None


def _is_num(v):
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _same_value(x, y):
    if _is_num(x) and _is_num(y):
        return float(x) == float(y)
    return type(x) is type(y) and x == y


def diff(before, after, path=None, out=None):
    """Authored paths at which two JSON documents differ (worker.rs:652-732):
    objects compare by key; a list holding no objects is one value; items of
    other lists pair by non-empty `name` where both sides have one, the rest by
    position, a left-over item is reported with its name or index, and a list
    whose named items changed order is reported itself; numbers compare by
    value (1 == 1.0)."""
    path = [] if path is None else path
    out = [] if out is None else out
    if isinstance(before, dict) and isinstance(after, dict):
        for key in sorted(set(before) | set(after)):
            path.append(key)
            if key in before and key in after:
                diff(before[key], after[key], path, out)
            else:
                out.append(list(path))
            path.pop()
        return out
    if isinstance(before, list) and isinstance(after, list):
        if not any(isinstance(v, dict) for v in before + after):
            if len(before) != len(after) or any(not _same_value(x, y) for x, y in zip(before, after)):
                out.append(list(path))
            return out
        stacks = {}
        for j in range(len(after) - 1, -1, -1):
            n = _name(after[j])
            if n is not None:
                stacks.setdefault(n, []).append(j)
        taken = [False] * len(after)
        pairs, unnamed = [], []
        for i, item in enumerate(before):
            n = _name(item)
            if n is None:
                unnamed.append(i)
            elif stacks.get(n):
                j = stacks[n].pop()
                taken[j] = True
                pairs.append((i, j))
            else:
                out.append(path + [n])
        rest = iter([j for j in range(len(after)) if not taken[j] and _name(after[j]) is None])
        for i in unnamed:
            j = next(rest, None)
            if j is None:
                out.append(path + [i])
            else:
                taken[j] = True
                pairs.append((i, j))
        for j, item in enumerate(after):
            if not taken[j]:
                n = _name(item)
                out.append(path + [n if n is not None else j])
        pairs.sort()
        if any(a[1] > b[1] for a, b in zip(pairs, pairs[1:])):
            out.append(list(path))
        for i, j in pairs:
            n = _name(after[j])
            path.append(n if n is not None else j)
            diff(before[i], after[j], path, out)
            path.pop()
        return out
    if not _same_value(before, after):
        out.append(list(path))
    return out


# ------------------------------------------------------------------ scene identity
def _digest(h, v):
    if isinstance(v, np.ndarray):
        h.update(str(v.dtype).encode() + str(v.shape).encode())
        h.update(np.ascontiguousarray(v).tobytes())
    elif isinstance(v, dict):
        for k in sorted(v):
            h.update(repr(k).encode())
            _digest(h, v[k])
    elif isinstance(v, (list, tuple)):
        h.update(b"[%d" % len(v))
        for x in v:
            _digest(h, x)
    else:
        h.update(repr(v).encode())


def geometry_key(scene):
    """What the acceleration structures are built from: every mesh and the
    instance records (transforms, BLAS, material base, mask, flags)."""
    h = hashlib.sha256()
    for m in scene.meshes:
        _digest(h, [m.positions, m.normals, m.indices, m.triangle_materials])
    inst = scene.instance_array()
    h.update(bytes(inst))
    return h.hexdigest()


def scene_key(scene):
    """Everything a render of the scene reads."""
    h = hashlib.sha256(geometry_key(scene).encode())
    _digest(h, [scene.materials, scene.lights, scene.camera, scene.sky, scene.width, scene.height, scene.spp,
                scene.max_depth, scene.seed])
    return h.hexdigest()


def _buffer_files(document, base):
    """External buffer files of a document -> content hash."""
    out = {}
    bufs = document.get("buffers") if isinstance(document, dict) else None
    for b in (bufs.values() if isinstance(bufs, dict) else bufs if isinstance(bufs, list) else []):
        f = b.get("file") if isinstance(b, dict) else None
        if isinstance(f, str):
            p = os.path.join(base, f)
            try:
                with open(p, "rb") as fh:
                    out[f] = hashlib.sha256(fh.read()).hexdigest()
            except OSError:
                out[f] = None
    return out


# ------------------------------------------------------------------ worker
class Worker:
    """One device, one open document and its installation (worker.rs:1-15)."""

    def __init__(self, ordinal=0):
        self.ordinal = ordinal
        self.renderer = None
        self.open_state = None      # dict: document, base, path, revision, scene, warnings, files, keys
        self.installed = None       # dict: blases, tlas, scenes {options key: InstalledScene}, geometry
        self.pending = "reused"     # how the next render obtains its installation
        self.installs = 0
        self.denoiser = None        # loaded on the first render that denoises

    # -------------------------------------------------------------- open
    def _device(self):
        from . import native
        created = self.renderer is None
        if created:
            self.renderer = native.Renderer(self.ordinal)
        return not created

    def _decode(self, document, base, path, overrides=None):
        from . import scene_json as J
        try:
            return J.load(document, overrides, base_dir=base)
        except J.SceneError as e:
            raise Failure("usage" if str(e).startswith("--set") else "scene", f"load scene {path}: {e}")
        except OSError as e:
            raise Failure("io", f"load scene {path}: {e}")
        except (ValueError, TypeError, KeyError, IndexError) as e:
            raise Failure("scene", f"load scene {path}: {e}")

    def _install(self, scene):
        from . import native
        try:
            self._device()
            blases = [native.Blas(self.renderer.device, m.positions, m.indices, m.normals, m.triangle_materials)
                      for m in scene.meshes]
            tlas = native.Tlas(self.renderer.device, blases, scene.instance_array())
            inst = native.InstalledScene(self.renderer.device, scene, blases, tlas)
        except native.VorkError as e:
            raise Failure({2: "usage", 3: "scene", 5: "io"}.get(e.status, "device"), str(e))
        self.installs += 1
        return {"blases": blases, "tlas": tlas, "geometry": geometry_key(scene),
                "scenes": {scene_key(scene): inst}}

    def open(self, scene=None, document=None, base=None):
        if (scene is None) == (document is None) or (scene is not None and base is not None):
            raise _usage("open takes scene (a file path) or document (the whole scene as JSON, with base naming "
                         "the directory for its relative files), not both")
        if scene is not None:
            if not isinstance(scene, str):
                raise _usage("open: scene must be a file path")
            path = scene
            try:
                with open(scene, "rb") as f:
                    raw = f.read()
            except OSError as e:
                raise Failure("io", f"load scene {scene}: {e.strerror or e}")
            try:
                document = json.loads(raw)
            except (ValueError, RecursionError) as e:
                raise Failure("scene", f"load scene {scene}: {e}")
            base = os.path.dirname(os.path.abspath(scene))
        else:
            if base is not None and not isinstance(base, str):
                raise _usage("open: base must be a directory path")
            base = base if base is not None else "."
            path = os.path.join(base, "(document)")
        if not isinstance(document, dict):
            raise Failure("scene", f"load scene {path}: the document is not a JSON object")
        decoded, warnings = self._decode(document, base, path)
        files = _buffer_files(document, base)
        prev = self.open_state
        changed, reuse = [], "reinstalled"
        if prev is not None:
            diff(prev["document"], document, [], changed)
            for f, digest in files.items():
                if f in prev["files"] and prev["files"][f] != digest:
                    changed.append(["file", f])
            identical = prev["key"] == scene_key(decoded) and prev["base"] == base and prev["files"] == files
            if identical:
                changed = []
            if self.installed is not None and identical:
                reuse = "reused"
            elif self.installed is not None and self.installed["geometry"] == geometry_key(decoded):
                reuse = "updated"
        # the candidate is installed before it is published, so a failure keeps the open revision
        installed = self.installed
        if reuse == "reinstalled":
            installed = self._install(decoded)
        elif reuse == "updated":
            from . import native
            try:
                inst = native.InstalledScene(self.renderer.device, decoded, installed["blases"], installed["tlas"])
            except native.VorkError as e:
                raise Failure({2: "usage", 3: "scene"}.get(e.status, "device"), str(e))
            installed = dict(installed, scenes={scene_key(decoded): inst})
        revision = prev["revision"] + 1 if prev is not None else 1
        self.open_state = {"document": document, "base": base, "path": path, "revision": revision,
                           "scene": decoded, "warnings": warnings, "files": files, "key": scene_key(decoded),
                           "decoded": {repr([]): (decoded, warnings)}}
        self.installed = installed
        self.pending = reuse
        return {"revision": revision, "reuse": reuse, "changed": changed}

    # -------------------------------------------------------------- render
    def _options(self, req):
        from . import scene_json as J
        unknown = [k for k in req if k not in RENDER_FIELDS]
        if unknown:
            raise _usage(f"unknown field `{unknown[0]}` in render; {REQUESTS}")
        if "out" not in req:
            raise _usage(f"missing field `out` in render; {REQUESTS}")
        outs = req["out"] if isinstance(req["out"], list) else [req["out"]]
        if not outs or not all(isinstance(o, str) for o in outs):
            raise _usage("render: out is a file path or a list of file paths")
        for o in outs:
            if os.path.splitext(o)[1].lower() not in (".png", ".exr"):
                raise _usage(f"render: output {o}: .png or .exr")
        profile = req.get("profile") or PROFILE
        if profile != PROFILE:
            raise _usage(f"profile {profile}: this build renders the {PROFILE} profile only")
        channels = req.get("channels") or []
        if channels not in ([], ["color"]):
            raise _usage("render: channels: only the colour output is written by this build")
        if req.get("at") is not None:
            raise _usage("render: at: animation is not rendered by this build (stills only)")
        fs, fe = req.get("frame_start"), req.get("frame_end")
        if (fs not in (None, 0)) or (fe not in (None, 0, 1)):
            raise _usage("render: frame ranges: this build renders one still frame")
        if req.get("exr_codec") not in (None, "zip"):
            raise _usage("render: exr_codec: zip only")
        if req.get("png_depth") not in (None, "8", 8):
            raise _usage("render: png_depth: 8 only")
        sets = req.get("set") or {}
        if not isinstance(sets, dict):
            raise _usage("render: set is an object of render-call fields")
        overrides = {}
        for key, value in sets.items():
            if key not in SET_KEYS:
                raise _usage(f"unknown render-call field {key}; fields: {', '.join(SET_KEYS)}")
            try:
                if isinstance(value, str):
                    overrides.update(J.parse_set([f"{key}={value}"]))
                elif key == "resolution":
                    w, h = value
                    overrides[key] = (int(w), int(h))
                elif key in ("samples", "seed"):
                    if not _is_num(value) or int(value) != value:
                        raise ValueError(value)
                    overrides[key] = int(value)
                elif key == "fps":
                    overrides[key] = float(value)
                else:
                    overrides[key] = bool(value)
            except (ValueError, TypeError) as e:
                raise _usage(f"render-call field {key}: {e}")
        if req.get("provisional"):
            overrides["samples"] = 1
            overrides["denoise"] = False
        return outs, overrides, bool(req.get("no_write"))

    def render(self, req):
        from . import native
        from .render import render_image, write_render
        t0 = time.perf_counter()
        if self.open_state is None or self.installed is None:
            raise _usage("render: no scene is open (send open first)")
        outs, overrides, no_write = self._options(req)
        st = self.open_state
        device_reused = self._device()
        # the decoded scene under these options: resolution changes the camera
        okey = repr(sorted(overrides.items()))
        if okey not in st["decoded"]:
            st["decoded"][okey] = self._decode(st["document"], st["base"], st["path"], overrides)
        scene, warnings = st["decoded"][okey]
        key = scene_key(scene)
        scenes = self.installed["scenes"]
        install_ms = 0.0
        if key not in scenes:
            ti = time.perf_counter()
            try:
                scenes[key] = native.InstalledScene(self.renderer.device, scene, self.installed["blases"],
                                                    self.installed["tlas"])
            except native.VorkError as e:
                raise Failure({2: "usage", 3: "scene"}.get(e.status, "device"), str(e))
            install_ms = (time.perf_counter() - ti) * 1e3
        inst = scenes[key]
        ready = (time.perf_counter() - t0) * 1e3
        p = scene.render_params()
        tr = time.perf_counter()
        try:
            if scene.notes.get("denoise") and self.denoiser is None:
                from .denoise import Denoiser
                self.denoiser = Denoiser(self.renderer.device)
            img, stats, denoised = render_image(self.renderer.device, inst, scene, p, self.denoiser)
        except native.VorkError as e:
            raise Failure({2: "usage", 3: "scene", 5: "io"}.get(e.status, "device"), str(e))
        loop_ms = (time.perf_counter() - tr) * 1e3
        wall = (time.perf_counter() - t0) * 1e3
        reuse, self.pending = self.pending, "reused"
        sidecars = []
        meta = None
        if not no_write:
            try:
                meta = write_render(scene, outs, img, stats, p.samples, wall, warnings=warnings,
                                    adaptive=bool(p.flags & 4), denoised=denoised)
            except OSError as e:
                raise Failure("io", f"write {outs[0]}: {e.strerror or e}")
            sidecars = [o + ".meta.json" for o in outs]
        return {"revision": st["revision"], "reuse": reuse, "installs": self.installs,
                "provisional": bool(req.get("provisional")), "device_reused": device_reused,
                "install_reused": install_ms == 0.0,
                "result": {"warnings": list(meta["warnings"]) if meta else list(warnings), "exposure": None,
                           "sidecars": sidecars, "frames": 1},
                "timing": {"wall_ms": round(wall, 3), "device_ms": round(stats["device_ms"], 3),
                           "install_ms": round(install_ms, 3), "ready_ms": round(ready, 3),
                           "first_frame_ms": round(wall, 3), "loop_wall_ms": round(loop_ms, 3)},
                "report": {"render_settings": meta["render_settings"]} if meta else None}

    def close(self):
        self.installed = None
        if self.denoiser is not None:
            self.denoiser.close()
            self.denoiser = None
        self.renderer = None


# ------------------------------------------------------------------ protocol
def _failure(rid, kind, message):
    return {"id": rid, "ok": False, "kind": kind, "error": message}


def handle(worker, line):
    """One request line -> (reply, stop)."""
    try:
        value = json.loads(line)
    except RecursionError:
        return _failure(None, "scene", "request: JSON nested too deeply"), False
    except ValueError as e:
        return _failure(None, "usage", f"request is not a JSON object: {e}"), False
    if not isinstance(value, dict):
        return _failure(None, "usage", "request is not a JSON object"), False
    rid = value.pop("id", None)
    op = value.pop("op", None)
    try:
        if op == "open":
            unknown = [k for k in value if k not in OPEN_FIELDS]
            if unknown:
                raise _usage(f"unknown field `{unknown[0]}` in open; {REQUESTS}")
            reply = worker.open(value.get("scene"), value.get("document"), value.get("base"))
        elif op == "render":
            reply = worker.render(value)
        elif op == "close":
            if value:
                raise _usage(f"unknown field `{next(iter(value))}` in close; {REQUESTS}")
            worker.close()
            return {"id": rid, "ok": True}, True
        elif op == "ask":
            raise _usage("ask: scene questions are not answered by this build; " + REQUESTS)
        else:
            raise _usage(f"unknown op {json.dumps(op)}; {REQUESTS}")
    except Failure as e:
        return _failure(rid, e.kind, str(e)), False
    return dict({"id": rid, "ok": True}, **reply), False


def serve(inp, out, worker=None):
    """Serves requests from `inp` until close or end of input."""
    worker = worker or Worker()
    for line in inp:
        line = line.rstrip("\n").rstrip("\r")
        if not line.strip():
            continue
        reply, stop = handle(worker, line)
        out.write(json.dumps(reply) + "\n")
        out.flush()
        if stop:
            break
    worker.close()
    return 0


def main(argv=None):
    return serve(sys.stdin, sys.stdout)


if __name__ == "__main__":
    sys.exit(main())
