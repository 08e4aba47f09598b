"""Autotuner of the library's build-time knobs on the target workload.

The kernels' launch shapes and queue sizes are compile-time constants of the
HIP sources (workgroup size, waves per SIMD each kernel is compiled for, how
many idle lanes a traversal wave waits for before refilling, how many of a
pixel's samples share a wave, the persistent grid's workgroups per CU (its
most, its least, and the paths per thread that pick between them), wavefront
batch length, BVH leaf size and clustering radius, and whether hits are
sorted into per-closure shading queues).  This module searches them
by measurement and records the winners in vork_amd/tuning.json, which
vork_amd/build.py compiles in.

The search is coordinate descent around the current table, in two halves
because compiling and measuring happen on different machines:

    python -m vork_amd.autotune build [--knobs a,b|none] [--from results.json]
        compiles one library per candidate (the current table with one knob
        changed; with --from, also the table of the best values found so far;
        --knobs none builds only that one)
        into vork_amd/_tune/<name>/libvork_hip.so and writes the plan;
    python -m vork_amd.autotune measure --out results.json [--reps 2]
        (on the GPU) runs bench.py on the target workload for every variant
        of the plan, alternating with the baseline, and records ms per step;
    python -m vork_amd.autotune select results.json [--write]
        takes, knob by knob, the value whose variant beat the baseline by more
        than the run-to-run noise, and with --write updates tuning.json.

Each variant is measured with the same bench command the headline number
comes from (config 1: 1920x1080, 64 spp, 7.3M triangles, BVH built every step),
loaded through VORK_HIP_LIB in a child process.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
TABLE = os.path.join(HERE, "tuning.json")
TUNE_DIR = os.path.join(HERE, "_tune")
PLAN = os.path.join(TUNE_DIR, "plan.json")

# knob -> candidate values (the macro each sets in the HIP sources)
KNOBS = {
    "VH_BLOCK": [128, 256, 512],
    "VH_TRACE_WAVES_PER_EU": [3, 4, 5],
    "VH_ANY_WAVES_PER_EU": [4, 5, 6, 7, 8],
    "VH_PACKET_WAVES_PER_EU": [5, 6, 7, 8],
    "VH_PACKET_MIN_GROUP": [4, 16, 32],
    "VH_SHADE_WAVES_PER_EU": [3, 4, 5],
    "VH_TRACE_REFILL_MIN": [8, 16, 24, 32, 48],
    "VH_TRACE_CLAIM_CHUNK": [0, 1],
    "VH_ANY_REFILL_MIN": [16, 32, 48, 56],
    "VH_ANY_CLAIM_CHUNK": [0, 1],
    "VH_TRACE_ITEMS_PER_STEP": [1, 2],
    "VH_WAVE_SAMPLES": [16, 32, 64],
    "VH_GRID_BLOCKS_PER_CU": [4, 6, 8, 12, 16],
    "VH_GRID_MIN_BLOCKS_PER_CU": [2, 4, 6],
    "VH_GRID_PATHS_PER_THREAD": [8, 16, 32],
    "VH_BATCH_PATHS_LOG2": [25, 26, 27],
    "VH_LEAF_MAX": [1, 2, 4],
    "VH_PLOC_RADIUS": [8, 16, 24, 32, 48],
    "VH_SHADE_SORT": [0, 1],
}
# knobs swept when none are named (VH_SHADE_SORT is measured by bench.py --shade instead)
DEFAULT_SWEEP = [k for k in KNOBS if k != "VH_SHADE_SORT"]
# the bench command every variant is timed with
BENCH = ["--steps", "4", "--warmup", "2", "--no-cpu-baseline", "--no-roofline", "--no-traffic"]
NOISE = 0.005   # relative: differences below this are run-to-run noise on this workload


def load_table():
    with open(TABLE) as f:
        return json.load(f)


def defines(values):
    return [f"-D{k}={v}" for k, v in sorted(values.items())]


def variant_name(values):
    return hashlib.sha1(json.dumps(sorted(values.items())).encode()).hexdigest()[:12]


def build_variant(values):
    """Compiles the library with these knob values; returns its path."""
    from . import build
    name = variant_name(values)
    out = os.path.join(TUNE_DIR, name)
    lib = os.path.join(out, "libvork_hip.so")
    if not os.path.exists(lib):
        build.build(out_dir=out, extra=defines(values))
    return lib


def cmd_build(args):
    table = load_table()
    base = dict(table["knobs"])
    knobs = [] if args.knobs == "none" else args.knobs.split(",") if args.knobs else DEFAULT_SWEEP
    plan = {"baseline": {"values": base, "lib": None}, "variants": []}
    plan["baseline"]["lib"] = build_variant(base)
    if args.from_results:
        with open(args.from_results) as f:
            best = select(json.load(f))["values"]
        if best != base:
            plan["variants"].append({"knob": "combined", "value": None, "values": best, "lib": build_variant(best)})
    for k in knobs:
        for v in KNOBS[k]:
            if v == base.get(k):
                continue
            values = dict(base, **{k: v})
            t = time.time()
            lib = build_variant(values)
            plan["variants"].append({"knob": k, "value": v, "values": values, "lib": lib})
            print(f"built {k}={v} in {time.time() - t:.0f}s", flush=True)
    os.makedirs(TUNE_DIR, exist_ok=True)
    with open(PLAN, "w") as f:
        json.dump(plan, f, indent=1)
    print(f"plan: {len(plan['variants'])} variants + baseline -> {PLAN}")


def run_bench(lib, workload_args):
    env = dict(os.environ, VORK_HIP_LIB=lib)
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), *BENCH, *workload_args]
    r = subprocess.run(cmd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"bench failed ({r.returncode}): {r.stderr[-500:]}")
    line = [l for l in r.stdout.splitlines() if l.startswith("{")][-1]
    return json.loads(line)["ms_per_step"]


def cmd_measure(args):
    with open(PLAN) as f:
        plan = json.load(f)
    # libraries travel in-tree: resolve them under this checkout
    def here(path):
        return os.path.join(TUNE_DIR, os.path.basename(os.path.dirname(path)), "libvork_hip.so")
    workload = args.workload_args.split() if args.workload_args else []
    results = {"bench": BENCH + workload, "baseline": plan["baseline"]["values"], "runs": []}
    base_lib = here(plan["baseline"]["lib"])
    for k, var in enumerate(plan["variants"]):
        for rep in range(args.reps):
            b = run_bench(base_lib, workload)
            v = run_bench(here(var["lib"]), workload)
            results["runs"].append({"knob": var["knob"], "value": var["value"], "values": var["values"],
                                    "baseline_ms": b, "variant_ms": v})
            print(f"[{k + 1}/{len(plan['variants'])}] {var['knob']}={var['value']}: {v:.2f} ms "
                  f"(baseline {b:.2f} ms)", flush=True)
        with open(args.out, "w") as f:
            json.dump(results, f, indent=1)
    print(f"results -> {args.out}")


def select(results):
    """Per knob, the value whose variant ran faster than the baseline it was
    paired with, on every repetition, by more than NOISE on average."""
    base = dict(results["baseline"])
    gains = {}
    for r in results["runs"]:
        gains.setdefault((r["knob"], json.dumps(r["value"])), []).append(r["variant_ms"] / r["baseline_ms"] - 1.0)
    chosen, report = dict(base), {}
    for (knob, value), g in sorted(gains.items()):
        mean = sum(g) / len(g)
        report[f"{knob}={value}"] = round(mean * 100, 2)
        if knob == "combined":
            continue
        if all(x < 0 for x in g) and mean < -NOISE:
            prev = report.get(f"_best_{knob}")
            if prev is None or mean < prev:
                chosen[knob] = json.loads(value)
                report[f"_best_{knob}"] = mean
    return {"values": chosen, "percent_change_vs_baseline": {k: v for k, v in report.items()
                                                              if not k.startswith("_")}}


def cmd_select(args):
    with open(args.results) as f:
        results = json.load(f)
    sel = select(results)
    print(json.dumps(sel, indent=1))
    if args.write:
        table = load_table()
        table["knobs"] = sel["values"]
        table["measured"] = {"bench": results["bench"], "percent_change_vs_baseline":
                             sel["percent_change_vs_baseline"], "date": time.strftime("%Y-%m-%d")}
        with open(TABLE, "w") as f:
            json.dump(table, f, indent=1)
            f.write("\n")
        print(f"wrote {TABLE}")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m vork_amd.autotune")
    sub = ap.add_subparsers(dest="cmd", required=True)
    b = sub.add_parser("build")
    b.add_argument("--knobs", default="")
    b.add_argument("--from", dest="from_results", default="")
    m = sub.add_parser("measure")
    m.add_argument("--out", required=True)
    m.add_argument("--reps", type=int, default=2)
    m.add_argument("--workload-args", default="")
    s = sub.add_parser("select")
    s.add_argument("results")
    s.add_argument("--write", action="store_true")
    args = ap.parse_args(argv)
    {"build": cmd_build, "measure": cmd_measure, "select": cmd_select}[args.cmd](args)


if __name__ == "__main__":
    main()
