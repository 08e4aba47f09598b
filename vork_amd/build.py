"""Builds libvork_hip.so, the HIP library of the hot path, in-tree for gfx950.

Each .hip translation unit is compiled with hipcc in parallel into
vork_amd/_build/ and linked into vork_amd/libvork_hip.so.  The knobs of
vork_amd/tuning.json (launch shapes and queue sizes the autotuner chose) are
passed as -D flags.
"""
import concurrent.futures
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvork_hip.so")
ARCH = os.environ.get("VORK_HIP_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
SOURCES = ["vh_api.hip", "vh_build.hip", "vh_trace.hip", "vh_render.hip", "vh_denoise.hip", "vh_train.hip"]
TUNING = os.path.join(HERE, "tuning.json")


def tuned_defines():
    """-D flags of the knobs in tuning.json."""
    with open(TUNING) as f:
        knobs = json.load(f)["knobs"]
    return [f"-D{k}={v}" for k, v in sorted(knobs.items())]


def _compile(src, out_dir, defines):
    out = os.path.join(out_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    warnings = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return out, warnings


def build(verbose=False, out_dir=None, extra=None):
    """Compiles and links the library: in-tree with the tuned knobs, or, with
    `out_dir`, a variant with `extra` -D flags overriding them (the
    autotuner's candidates).  Returns the library's path."""
    obj_dir = out_dir or BUILD
    lib = os.path.join(out_dir, "libvork_hip.so") if out_dir else LIB
    os.makedirs(obj_dir, exist_ok=True)
    overridden = {d.split("=")[0] for d in (extra or [])}
    defines = [d for d in tuned_defines() if d.split("=")[0] not in overridden] + list(extra or [])
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(lambda src: _compile(src, obj_dir, defines), SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for (_, w), src in zip(results, SOURCES):
            if w:
                print(f"--- {src}\n{w}", file=sys.stderr)
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
