"""Writes tests/golden/agx_samples.json: samples of the reference's baked AgX
display LUT (look None) at LUT grid nodes, read from vork's
libraries/std/agx.exr (OpenEXR half RGB, ZIP; sample (r,g,b) of look L at
x = 65 b + r, y = 65 L + g, per libraries/std/agx-manifest.json).

Run from the repository root in a container that holds the reference:
    python tests/golden/make_agx_samples.py /root/reference/libraries/std/agx.exr
At a grid node the LUT application of libraries/blender/display_lut.slang
returns the stored sample itself, so these are expected outputs for inputs
2^(min_log2 + i * step)."""
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from vork_amd import film  # noqa: E402

src = sys.argv[1]
lut = film.load_agx_lut(src, look=0)   # [b][g][r]
n = film.AGX_N
step = (film.AGX_MAX_LOG2 - film.AGX_MIN_LOG2) / (n - 1)
rng = np.random.default_rng(2024)
idx = [(i, i, i) for i in range(0, n, 2)]                       # the grey axis
idx += [tuple(int(v) for v in rng.integers(1, n, 3)) for _ in range(400)]
idx += [(k, 1, 1) for k in range(1, n, 4)] + [(1, k, 1) for k in range(1, n, 4)] + [(1, 1, k) for k in range(1, n, 4)]
rows = []
for r, g, b in idx:
    x = [0.0 if i == 0 else float(2.0 ** (film.AGX_MIN_LOG2 + i * step)) for i in (r, g, b)]
    rows.append({"index": [r, g, b], "linear": x, "display": [float(v) for v in lut[b, g, r]]})
out = {"source": "libraries/std/agx.exr (look None)", "size": n, "min_log2": film.AGX_MIN_LOG2,
       "max_log2": film.AGX_MAX_LOG2, "samples": rows}
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "agx_samples.json"), "w") as f:
    json.dump(out, f)
print(len(rows), "samples; black", lut[0, 0, 0].tolist(), "white-ish", lut[-1, -1, -1].tolist())
