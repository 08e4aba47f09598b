"""Film output: the resolve mean, the sRGB display views and the image files
`vork render --out FILE` writes (PNG display codes, scene-linear EXR, the
`.meta.json` sidecar).

* resolve: accumulated RGBA32F sum / sample count
  (crates/vork-render/shaders/resolve.slang:228-238, the count in alpha);
* `standard` view: exposure, the sRGB transfer, creative gamma
  (libraries/std/standard.slang:25-30, libraries/std/view.slang:3-9,34-39);
* `agx` view: with the reference's baked 65^3 AgX LUT (`load_agx_lut` on
  vork's libraries/std/agx.exr) applied exactly as
  libraries/blender/display_lut.slang:2-16 does (log2 index over
  [-12.47393, 12.5260688117], linear first interval, trilinear); without it,
  an analytic AgX (a published fit of Blender's AgX base, see `agx_analytic`)
  whose deviation from the LUT is measured by the tests;
* PNG: 8- or 16-bit RGB(A), zlib, no filtering; EXR: scanline RGBA float32 or
  half, uncompressed or ZIP; both readable by `read_png` / `read_exr`;
* `.meta.json`: version-3 sidecar with the fields
  crates/vork-render/src/shot.rs:1324-1331 writes that apply here.

Quantisation is round-to-nearest without the reference's output dither (its
pattern is the encoder's own), so PNG codes can differ by one step where the
dither would have moved them.
"""
import json
import os
import struct
import zlib

import numpy as np

AGX_N = 65
AGX_MIN_LOG2, AGX_MAX_LOG2 = -12.47393, 12.5260688117


# ---------------------------------------------------------------- resolve
def resolve(film):
    """(H, W, 4) accumulated sum with the sample count in alpha -> (H, W, 3)
    mean; pixels without samples are black."""
    film = np.asarray(film, np.float64)
    n = film[..., 3:4]
    return np.where(n > 0, film[..., :3] / np.maximum(n, 1), 0.0)


# ------------------------------------------------------------------ views
def encode_srgb(c):
    """libraries/std/view.slang:7-9."""
    c = np.asarray(c, np.float64)
    return np.where(c > 0.0031308, 1.055 * np.power(np.maximum(c, 0.0), 1 / 2.4) - 0.055, 12.92 * c)


def decode_srgb(c):
    c = np.asarray(c, np.float64)
    return np.where(c > 0.04045, np.power(np.maximum((c + 0.055) / 1.055, 0.0), 2.4), c / 12.92)


def _finish(codes, gamma):
    if gamma != 1.0:
        codes = np.power(np.maximum(codes, 0.0), 1.0 / gamma)
    return codes


def view_standard(rgb, exposure=0.0, gamma=1.0):
    """Display codes of the Standard view for an sRGB display."""
    return _finish(encode_srgb(np.asarray(rgb, np.float64) * 2.0 ** exposure), gamma)


# A published fit of Blender's AgX base (inset matrix, log2 allocation over
# [-12.47393, 4.026069], sixth-order contrast curve, outset matrix, 2.2 display
# power; "Minimal AgX", B. Wrensch 2023).  Rows apply to column vectors.
_AGX_INSET = np.array([[0.842479062253094, 0.0784335999999992, 0.0792237451477643],
                       [0.0423282422610123, 0.878468636469772, 0.0791661274605434],
                       [0.0423756549057051, 0.0784336, 0.879142973793104]])
_AGX_OUTSET = np.array([[1.19687900512017, -0.0980208811401368, -0.0990297440797205],
                        [-0.0528968517574562, 1.15190312990417, -0.0989611768448433],
                        [-0.0529716355144438, -0.0980434501171241, 1.15107367264116]])
_AGX_EV = (-12.47393, 4.026069)


def _agx_curve(x):
    x2 = x * x
    x4 = x2 * x2
    return 15.5 * x4 * x2 - 40.14 * x4 * x + 31.96 * x4 - 6.868 * x2 * x + 0.4298 * x2 + 0.1191 * x - 0.00232


def agx_analytic(rgb):
    """Display codes of the analytic AgX for an sRGB display."""
    c = np.maximum(np.asarray(rgb, np.float64), 0.0) @ _AGX_INSET.T
    lo, hi = _AGX_EV
    x = (np.clip(np.log2(np.maximum(c, 1e-30)), lo, hi) - lo) / (hi - lo)
    y = _agx_curve(x) @ _AGX_OUTSET.T
    return encode_srgb(np.power(np.maximum(y, 0.0), 2.2))


def agx_lut(rgb, lut):
    """libraries/blender/display_lut.slang:2-16 on a (65, 65, 65, 3) LUT
    indexed [b][g][r] (the library image's band of one look)."""
    n = float(AGX_N)
    step = (AGX_MAX_LOG2 - AGX_MIN_LOG2) / (n - 1)
    first = 2.0 ** (AGX_MIN_LOG2 + step)
    c = np.maximum(np.asarray(rgb, np.float64), 0.0)
    logi = (np.log2(np.maximum(c, 1e-30)) - AGX_MIN_LOG2) / step
    index = np.clip(np.where(c < first, c / first, logi), 0.0, n - 1)
    r, g, b = index[..., 0], index[..., 1], index[..., 2]
    tile = np.minimum(np.floor(b), n - 2).astype(np.int64)

    def bilinear(t):
        # hardware-style bilinear sample of the (g, r) slice `t` at (r, g)
        r0 = np.clip(np.floor(r), 0, n - 2).astype(np.int64)
        g0 = np.clip(np.floor(g), 0, n - 2).astype(np.int64)
        fr, fg = (r - r0)[..., None], (g - g0)[..., None]
        a = lut[t, g0, r0] * (1 - fr) + lut[t, g0, r0 + 1] * fr
        bb = lut[t, g0 + 1, r0] * (1 - fr) + lut[t, g0 + 1, r0 + 1] * fr
        return a * (1 - fg) + bb * fg

    lo, hi = bilinear(tile), bilinear(tile + 1)
    return lo + (hi - lo) * (b - tile)[..., None]


def view_agx(rgb, exposure=0.0, gamma=1.0, lut=None):
    c = np.asarray(rgb, np.float64) * 2.0 ** exposure
    return _finish(agx_lut(c, lut) if lut is not None else agx_analytic(c), gamma)


def load_agx_lut(path, look=0):
    """The 65^3 LUT of `look` (0 = None) from vork's libraries/std/agx.exr
    (x = 65 b + r, y = 65 look + g)."""
    ch = read_exr(path)
    img = np.stack([ch["R"], ch["G"], ch["B"]], -1).astype(np.float64)
    band = img[AGX_N * look:AGX_N * (look + 1)]            # (g, 65*65 (b, r), 3)
    return band.reshape(AGX_N, AGX_N, AGX_N, 3).transpose(1, 0, 2, 3)  # [b][g][r]


# -------------------------------------------------------------------- PNG
def _chunk(kind, data):
    c = struct.pack(">I", len(data)) + kind + data
    return c + struct.pack(">I", zlib.crc32(kind + data) & 0xFFFFFFFF)


def quantize(codes, depth=8):
    top = (1 << depth) - 1
    q = np.floor(np.clip(np.asarray(codes, np.float64), 0.0, 1.0) * top + 0.5)
    return q.astype(np.uint8 if depth == 8 else np.uint16)


def write_png(path, codes, depth=8, alpha=None):
    """Display codes in [0, 1] (H, W, 3) -> PNG; `alpha` (H, W) optional."""
    q = quantize(codes, depth)
    if alpha is not None:
        q = np.concatenate([q, quantize(alpha, depth)[..., None]], -1)
    h, w, ch = q.shape
    raw = q.astype(">u2").tobytes() if depth == 16 else q.tobytes()
    stride = w * ch * (depth // 8)
    rows = b"".join(b"\x00" + raw[y * stride:(y + 1) * stride] for y in range(h))
    color_type = 6 if ch == 4 else 2
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color_type, 0, 0, 0))
    data += _chunk(b"sRGB", b"\x00") + _chunk(b"IDAT", zlib.compress(rows, 6)) + _chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(data)


def read_png(path):
    """Decodes the PNGs write_png produces (filter 0 rows) -> uint array."""
    with open(path, "rb") as f:
        data = f.read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, hdr = 8, b"", None
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        if zlib.crc32(kind + body) & 0xFFFFFFFF != struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]:
            raise ValueError("PNG CRC mismatch")
        if kind == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    w, h, depth, ctype = hdr[0], hdr[1], hdr[2], hdr[3]
    ch = 4 if ctype == 6 else 3
    bpp = ch * depth // 8
    raw = zlib.decompress(idat)
    rows = []
    for y in range(h):
        line = raw[y * (w * bpp + 1):(y + 1) * (w * bpp + 1)]
        if line[0] != 0:
            raise ValueError("only unfiltered rows are read here")
        rows.append(line[1:])
    dt = np.dtype(">u2") if depth == 16 else np.uint8
    return np.frombuffer(b"".join(rows), dt).reshape(h, w, ch).astype(np.uint16 if depth == 16 else np.uint8)


# -------------------------------------------------------------------- EXR
_EXR_MAGIC = 20000630


def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def _zip_pack(raw):
    """OpenEXR ZIP: interleave even/odd bytes, delta predictor, zlib."""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]])
    d = t.astype(np.int16)
    d[1:] = (t[1:].astype(np.int16) - t[:-1].astype(np.int16) + 128) & 0xFF
    return zlib.compress(d.astype(np.uint8).tobytes())


def _zip_unpack(data, size):
    t = np.frombuffer(zlib.decompress(data), np.uint8).astype(np.int32)
    if t.size != size:
        raise ValueError("EXR block size mismatch")
    v = np.cumsum(np.concatenate([t[:1], t[1:] - 128])) & 0xFF
    v = v.astype(np.uint8)
    half = (size + 1) // 2
    out = np.empty(size, np.uint8)
    out[0::2], out[1::2] = v[:half], v[half:]
    return out.tobytes()


def write_exr(path, rgba, half=False, compression="zip"):
    """(H, W, 4) scene-linear RGBA -> scanline OpenEXR (channels A, B, G, R)."""
    rgba = np.asarray(rgba, np.float32)
    h, w = rgba.shape[:2]
    names = ["A", "B", "G", "R"]
    ptype, dt, size = (1, np.dtype("<f2"), 2) if half else (2, np.dtype("<f4"), 4)
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iBBBBii", ptype, 0, 0, 0, 0, 1, 1) for n in names) + b"\0"
    comp = {"none": 0, "zip": 3}[compression]
    lines = 16 if comp == 3 else 1
    header = struct.pack("<ii", _EXR_MAGIC, 2)
    header += _attr("channels", "chlist", chlist)
    header += _attr("compression", "compression", bytes([comp]))
    header += _attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    header += _attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, w - 1, h - 1))
    header += _attr("lineOrder", "lineOrder", b"\0")
    header += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    order = [3, 2, 1, 0]  # A, B, G, R from RGBA
    blocks = []
    for y0 in range(0, h, lines):
        rows = rgba[y0:y0 + lines]
        raw = b"".join(rows[k][:, c].astype(dt).tobytes() for k in range(rows.shape[0]) for c in order)
        payload = raw
        if comp == 3:
            z = _zip_pack(raw)
            payload = z if len(z) < len(raw) else raw
        blocks.append(struct.pack("<ii", y0, len(payload)) + payload)
    table_at = len(header)
    offsets, pos = [], table_at + 8 * len(blocks)
    for b in blocks:
        offsets.append(pos)
        pos += len(b)
    with open(path, "wb") as f:
        f.write(header + struct.pack(f"<{len(offsets)}Q", *offsets) + b"".join(blocks))


def read_exr(path):
    """Scanline OpenEXR (no / ZIPS / ZIP compression, half or float) -> {channel: (H, W) float32}."""
    with open(path, "rb") as f:
        data = f.read()
    magic, version = struct.unpack("<ii", data[:8])
    if magic != _EXR_MAGIC or version & 0x200:
        raise ValueError("not a scanline OpenEXR file")
    pos, attrs = 8, {}
    while data[pos] != 0:
        e = data.index(b"\0", pos)
        name = data[pos:e].decode()
        e2 = data.index(b"\0", e + 1)
        n, = struct.unpack("<i", data[e2 + 1:e2 + 5])
        attrs[name] = data[e2 + 5:e2 + 5 + n]
        pos = e2 + 5 + n
    pos += 1
    chans, cl = [], attrs["channels"]
    p = 0
    while cl[p] != 0:
        e = cl.index(b"\0", p)
        chans.append((cl[p:e].decode(), struct.unpack("<i", cl[e + 1:e + 5])[0]))
        p = e + 17
    comp = attrs["compression"][0]
    if comp not in (0, 2, 3):
        raise ValueError(f"EXR compression {comp} is not read here")
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    lines = {0: 1, 2: 1, 3: 16}[comp]
    nblocks = (h + lines - 1) // lines
    offsets = struct.unpack(f"<{nblocks}Q", data[pos:pos + 8 * nblocks])
    sizes = {1: 2, 2: 4, 0: 4}
    kinds = {1: np.dtype("<f2"), 2: np.dtype("<f4"), 0: np.dtype("<u4")}
    out = {name: np.zeros((h, w), np.float32) for name, _ in chans}
    for k, off in enumerate(offsets):
        y, n = struct.unpack("<ii", data[off:off + 8])
        body = data[off + 8:off + 8 + n]
        rows = min(lines, h - (y - y0))
        expect = rows * w * sum(sizes[t] for _, t in chans)
        raw = body if (comp == 0 or n == expect) else _zip_unpack(body, expect)
        q = 0
        for r in range(rows):
            for name, t in chans:
                nbytes = w * sizes[t]
                out[name][y - y0 + r] = np.frombuffer(raw[q:q + nbytes], kinds[t]).astype(np.float32)
                q += nbytes
    return out


# ------------------------------------------------------------------ write
def write_outputs(path, film, view="agx", exposure=0.0, gamma=1.0, lut=None, depth=8, meta=None):
    """Writes the film as `path` (.png display codes, .exr scene-linear mean)
    and `<path>.meta.json` when `meta` is given."""
    rgb = resolve(film)
    ext = os.path.splitext(path)[1].lower()
    if ext == ".exr":
        a = np.ones(rgb.shape[:2] + (1,))
        write_exr(path, np.concatenate([rgb, a], -1))
    elif ext == ".png":
        codes = view_agx(rgb, exposure, gamma, lut) if view == "agx" else view_standard(rgb, exposure, gamma)
        write_png(path, codes, depth)
    else:
        raise ValueError(f"output {path}: .png or .exr")
    if meta is not None:
        with open(path + ".meta.json" if not path.endswith(".meta.json") else path, "w") as f:
            json.dump(meta, f, indent=1)


def metadata(scene, stats, warnings, samples, wall_ms, camera="camera", profile="std/path", adaptive=True,
             denoised=False):
    """The version-3 sidecar of crates/vork-render/src/shot.rs:1324-1331
    restricted to a still of the path integrator."""
    return {
        "version": 3, "fps": 24, "object_ids": [], "warnings": list(warnings),
        "frames": [{"index": 0, "time": 0.0, "camera": camera,
                    "dropped_path_samples": int(stats.get("dropped_samples", 0)),
                    "invalid_pixel_count": 0, "exposure": None}],
        "profile": {"name": profile, "level": 2},
        "settings": {"integrator": "path", "samples": samples, "seed": scene.seed},
        "render_settings": {"profile": {"name": profile, "level": 2}, "samples": samples, "adaptive": adaptive,
                            "denoise": denoised, "resolution": [scene.width, scene.height], "frames": 1},
        "denoiser": {"name": "vork_amd U-Net (bf16 MFMA)", "guides": ["albedo", "normal", "depth"]} if denoised else None,
        "wall_ms": wall_ms,
    }
