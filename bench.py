#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Workload (the single-GPU headline configuration): a 5,242,880-triangle
noise-displaced icosphere plus a 2,097,152-triangle heightfield (7,340,032
triangles), GGX materials, two rectangle lights and a procedural sky,
rendered at 1920x1080, 64 spp, depth 8.  A step is one whole render: both
BLASes and the TLAS built on the device from vertex data already resident in
HBM, then all 64 exposure samples of every pixel accumulated into an
RGBA32F film in HBM.

`--workload config2` runs the instanced workload instead: a million copies
of a 2,000-triangle bird, each step rebuilding the TLAS from that frame's
device-resident instance table and rendering 4 spp.

`--workload config4` measures the learned denoiser on the config 1 scene at
1920x1080: 4 spp with guides + the bf16 MFMA U-Net against a plain 64-spp
render, both scored by relMSE against a 4096-spp render of the same scene
(independent seeds), their latencies, and the U-Net's MFMA rate.

`--workload config3` runs the serving workload: a Poisson stream of render
requests over a pool of resident seeded scenes (13k to 5.2M triangles, the
config 0 and config 1 generators, drawn log-uniformly by triangle count) at
resolutions {256^2, 512^2, 1280x720, 1920x1080} and {1, 4, 16, 64} spp,
through the continuous-batching server (vork_amd.serve); it reports the
throughput and p50/p99 latency per arrival rate, and the same requests
served one at a time.

Prints one JSON line: Msamples/s (W*H*spp per second of a step), render
latency percentiles, the roofline of the closest-hit traversal kernel
(algorithmic bytes from counted node/triangle visits over HIP-event-timed
launches) and a CPU baseline (the CPU oracle restating the same estimator,
on a bounded sample of the same workload).
"""
import argparse
import concurrent.futures
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

S_NODE, S_TRI, S_RAYIN, S_HITOUT = 80, 36, 32, 16   # bytes: wide node, triangle record, ray in, hit out
HBM_PEAK_GBPS = 8000.0
# HBM-side traffic of the traced kernels, measured by this script's own
# rocprofv3 children (FETCH_SIZE, then WRITE_SIZE, each over one render step
# and over a plain device copy of known size that scales the counter to bytes),
# and their L2 hit rates (TCC_HIT_sum and TCC_MISS_sum over one render step)
TRAFFIC_KERNELS = {"k_extend_packet": "k_extend_packet", "k_extend": "k_extend<false", "k_shade": "k_shade<",
                   "k_shadow": "k_shadow<false"}
COPY_MIB = 1024


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["config1", "config2", "config3", "config4"], default="config1",
                    help="config1: 7.3M-triangle GGX scene, 64 spp, BVH built every step (the headline); "
                         "config2: 1M instanced birds, 4 spp, TLAS rebuilt for every frame; "
                         "config3: Poisson stream of requests through the render server; "
                         "config4: 4 spp + learned denoiser vs plain 64 spp at 1080p")
    ap.add_argument("--ref-spp", type=int, default=4096, help="config4: samples of the reference render")
    ap.add_argument("--weights", default=None, help="config4: denoiser weight blob (default: the shipped one)")
    ap.add_argument("--train-batch", type=int, default=8, help="config4: crops per timed training step")
    ap.add_argument("--train-crop", type=int, default=256, help="config4: crop size of the timed training step")
    ap.add_argument("--requests", type=int, default=40, help="config3: requests per arrival rate")
    ap.add_argument("--rates", default="4,8,inf", help="config3: arrival rates (requests/s), inf = all at once")
    ap.add_argument("--slots", type=int, default=4, help="config3: requests rendered at once")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=0, help="samples per pixel (0 = the workload's)")
    ap.add_argument("--sphere-level", type=int, default=9)
    ap.add_argument("--terrain", type=int, default=1024)
    ap.add_argument("--birds", type=int, default=1_000_000)
    ap.add_argument("--frames", type=int, default=4, help="config2: distinct animation frames cycled by the steps")
    ap.add_argument("--batch", type=int, default=0, help="exposure samples per wavefront batch (0 = auto)")
    ap.add_argument("--shade", choices=["default", "sorted", "arrival"], default="default",
                    help="shading order: hits sorted into per-closure queues, or as traced (default: the tuned one)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes that measure HBM traffic")
    ap.add_argument("--traffic-pass", action="store_true",
                    help="(run under rocprofv3 by this script) one untimed render step, no report")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write) three times, for scaling "
                         "rocprof FETCH_SIZE / WRITE_SIZE to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the path shards by image tiles or sample ranges with one film reduction, not built here)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from vork_amd import abi, native, scene as S

    if args.workload == "config3":
        print(json.dumps(serve_bench(args)))
        return
    if args.workload == "config4":
        print(json.dumps(denoise_bench(args)))
        return

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return

    t0 = time.time()
    if args.workload == "config1":
        sc = S.config1(width=args.width, height=args.height, spp=args.spp or 64, sphere_level=args.sphere_level,
                       terrain=args.terrain)
        frames = None
    else:
        sc = S.config2(width=args.width, height=args.height, spp=args.spp or 4, birds=args.birds)
        frames = S.flock_frames(sc, args.frames, progress=lambda k: k % 32 == 0 and log(f"{k} frames made"))
    log(f"scene: {sc.triangle_count} triangles generated in {time.time() - t0:.1f}s")

    dev = native.Device(0)
    torch.cuda.init()
    # inputs resident in HBM before anything is timed
    gpu_meshes = []
    for m in sc.meshes:
        mats = m.triangle_materials if m.triangle_materials is not None else np.zeros(m.triangle_count, np.uint32)
        gpu_meshes.append(dict(
            positions=torch.from_numpy(np.ascontiguousarray(m.positions)).cuda(),
            indices=torch.from_numpy(np.ascontiguousarray(m.indices).view(np.int32)).cuda(),
            normals=torch.from_numpy(np.ascontiguousarray(m.normals)).cuda(),
            materials=torch.from_numpy(np.ascontiguousarray(mats).view(np.int32)).cuda()))
    instances = sc.instance_array()
    gpu_frames = [torch.from_numpy(f.view(np.uint8)).cuda() for f in frames] if frames is not None else None
    film = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()

    # The meshes' BLASes build concurrently, each from its own host thread on
    # its own stream (the library keeps a build scratch per stream); each
    # build returns with its BLAS complete.
    build_streams = [torch.cuda.Stream() for _ in gpu_meshes]
    build_pool = concurrent.futures.ThreadPoolExecutor(max(len(gpu_meshes), 1))

    def build_one(k):
        g = gpu_meshes[k]
        return native.Blas(dev, g["positions"], g["indices"], g["normals"], g["materials"], device_pointers=True,
                           stream=build_streams[k].cuda_stream)

    def build_blases():
        return list(build_pool.map(build_one, range(len(gpu_meshes))))

    resident = {}
    if gpu_frames is not None:
        # config2: static BLASes built once; each step rebuilds the TLAS of its frame
        resident["blases"] = build_blases()
        resident["tlas"] = native.Tlas(dev, resident["blases"], gpu_frames[0], device_pointers=True,
                                       count=len(instances))
        resident["scene"] = native.InstalledScene(dev, sc, resident["blases"], resident["tlas"])
        dev.synchronize()
    frame_no = [0]
    tlas_ms = []

    shade_flag = {"default": 0, "sorted": abi.VH_RENDER_SHADE_SORTED, "arrival": abi.VH_RENDER_SHADE_ARRIVAL}[args.shade]

    def step(flags=0):
        p = sc.render_params(adaptive=False)
        p.batch_samples = args.batch
        p.flags = flags | shade_flag
        if gpu_frames is None:
            blases = build_blases()
            tlas = native.Tlas(dev, blases, instances)
            inst = native.InstalledScene(dev, sc, blases, tlas)
        else:
            blases, inst = resident["blases"], resident["scene"]
            f = gpu_frames[frame_no[0] % len(gpu_frames)]
            frame_no[0] += 1
            t = time.perf_counter()
            resident["tlas"].build(f, refit=True, device_pointers=True, count=len(instances))  # returns synchronized
            tlas_ms.append((time.perf_counter() - t) * 1e3)
        inst.render(p, device_film=film)
        dev.synchronize()
        return inst, blases

    if args.traffic_pass:
        film.zero_()
        torch.cuda.synchronize()
        step()
        log("traffic pass: one render step done")
        return

    for w in range(args.warmup):
        film.zero_()
        torch.cuda.synchronize()
        step()
        log(f"warmup {w + 1}/{args.warmup} done")

    times = []
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    for k in range(args.steps):
        film.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        inst, blases = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"step {k + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
    torch.cuda.synchronize()
    total = time.perf_counter() - t_all
    ms_per_step = 1e3 * sum(times) / len(times)
    samples = sc.width * sc.height * sc.spp
    value = samples / (ms_per_step / 1e3) / 1e6
    img = film[..., :3].cpu().numpy() / sc.spp
    finite = bool(np.isfinite(img).all())
    stats = inst.stats()
    binfo = [b.info() for b in blases]

    roofline = None
    if not args.no_roofline:
        film.zero_()
        torch.cuda.synchronize()
        ip, _ = step(abi.VH_RENDER_PROFILE)
        prof = ip.stats()
        film.zero_()
        torch.cuda.synchronize()
        ic, _ = step(abi.VH_RENDER_COUNT_VISITS)
        cnt = ic.stats()
        launches = max(prof["extend_launches"], 1)
        algorithmic = (cnt["extend_node_visits"] * S_NODE + cnt["extend_triangle_tests"] * S_TRI
                       + cnt["extension_rays"] * (S_RAYIN + S_HITOUT))
        per_launch = algorithmic / launches
        avg_ms = prof["extend_ms"] / launches
        achieved = per_launch / (avg_ms * 1e-3) / 1e9
        # the camera bounce and the later bounces (k_extend: incoherent rays,
        # per lane) priced apart, both by the per-ray visits of the per-lane
        # traversal; the camera bounce runs as wave packets (k_extend_packet)
        # when a wave's 64 rays are the samples of at most 4 pixels (16 or more
        # samples of a pixel per batch), else per lane in k_extend too
        group = 1
        while group < 64 and sc.spp % (2 * group) == 0:
            group *= 2
        camera_kernel = "k_extend_packet" if group >= 16 else "k_extend"
        cam_bytes = (cnt["extend_camera_node_visits"] * S_NODE + cnt["extend_camera_triangle_tests"] * S_TRI
                     + cnt["extend_camera_rays"] * (S_RAYIN + S_HITOUT))
        parts = {}
        for name, nbytes, ms, n, rays, nodes, tris in (
                ("camera bounce", cam_bytes, prof["extend_camera_ms"], prof["extend_camera_launches"],
                 cnt["extend_camera_rays"], cnt["extend_camera_node_visits"], cnt["extend_camera_triangle_tests"]),
                ("later bounces", algorithmic - cam_bytes, prof["extend_ms"] - prof["extend_camera_ms"],
                 launches - prof["extend_camera_launches"], cnt["extension_rays"] - cnt["extend_camera_rays"],
                 cnt["extend_node_visits"] - cnt["extend_camera_node_visits"],
                 cnt["extend_triangle_tests"] - cnt["extend_camera_triangle_tests"])):
            n = max(n, 1)
            gbps = nbytes / n / (ms / n * 1e-3) / 1e9 if ms > 0 else None
            parts[name] = {"kernel": camera_kernel if name == "camera bounce" else "k_extend", "launches": n, "avg_launch_ms": round(ms / n, 4), "bytes_per_launch": int(nbytes / n),
                           "rays": rays, "node_visits_per_ray": round(nodes / max(rays, 1), 2),
                           "triangle_tests_per_ray": round(tris / max(rays, 1), 2),
                           "achieved": round(gbps, 1) if gbps else None,
                           "frac": round(gbps / HBM_PEAK_GBPS, 4) if gbps else None}
        roofline = {"kernel": f"extend stage: {camera_kernel} (camera rays) + k_extend (later bounces), closest-hit "
                              "traversal", "bound": "hbm",
                    "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                    "bytes_per_launch": int(per_launch), "avg_launch_ms": round(avg_ms, 4), "launches": launches,
                    "node_visits_per_ray": round(cnt["extend_node_visits"] / max(cnt["extension_rays"], 1), 2),
                    "triangle_tests_per_ray": round(cnt["extend_triangle_tests"] / max(cnt["extension_rays"], 1), 2),
                    "per_launch": parts,
                    "stage_ms": {"extend": round(prof["extend_ms"], 1), "shade": round(prof["shade_ms"], 1),
                                 "shadow": round(prof["shadow_ms"], 1), "other": round(prof["other_ms"], 1)}}
        if not args.no_traffic:
            tr = measure_traffic(args)
            roofline["traffic_measured"] = tr
            ext = tr.get("kernels", {})
            if "k_extend" in ext and (camera_kernel == "k_extend" or "k_extend_packet" in ext):
                launches_all = sum(ext[k]["launches"] for k in {camera_kernel, "k_extend"})
                stage_bytes = sum(ext[k]["bytes_per_launch"] * ext[k]["launches"] for k in {camera_kernel, "k_extend"})
                roofline["traffic"] = int(stage_bytes / launches_all)
                roofline["traffic_source"] = ("rocprofv3 FETCH_SIZE + WRITE_SIZE of this run's own children, "
                                              "per launch of the extend stage, copy-calibrated")
                if camera_kernel == "k_extend_packet":   # the two kernels' traffic apart
                    for part in parts.values():
                        t = ext[part["kernel"]]["bytes_per_launch"]
                        part["traffic_bytes_per_launch"] = t
                        part["traffic_gbps"] = round(t / (part["avg_launch_ms"] * 1e-3) / 1e9, 1)
                        if "l2_hit_rate" in ext[part["kernel"]]:
                            part["l2_hit_rate"] = ext[part["kernel"]]["l2_hit_rate"]
                kl = [ext[k] for k in {camera_kernel, "k_extend"} if "l2_hits" in ext[k]]
                if kl:
                    h = sum(k["l2_hits"] for k in kl)
                    roofline["l2_hit_rate"] = round(h / max(h + sum(k["l2_misses"] for k in kl), 1), 4)

    cpu = None
    if not args.no_cpu_baseline and args.workload == "config1":
        try:
            cpu = cpu_baseline(sc)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "Msamples/s", "error": str(e)}

    if args.workload == "config1":
        workload = ("icosphere 5.2M tris + 1024^2 heightfield 2.1M tris, GGX 50% metallic, 2 rect lights + "
                    "procedural sky; BVH built on the GPU every step")
    else:
        workload = (f"{args.birds} instances of a 2,000-tri bird (Gaussian ribbon, curl-noise advection, "
                    f"{args.frames} frames cycled) + ground, 1 rect light + procedural sky; TLAS rebuilt from the "
                    "frame's device-resident instance table every step, BLASes built once")
    lat = sorted(t * 1e3 for t in times)
    line = {
        "metric": "Msamples/s @1080p + p50/p99 render latency at equal relMSE vs 4096-spp oracle",
        "value": round(value, 3), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded procedural scene, generated on the host, uploaded before timing)",
        "config": {"workload": workload,
                   "width": sc.width, "height": sc.height, "spp": sc.spp, "max_depth": sc.max_depth,
                   "triangles": sc.triangle_count, "batch_samples": args.batch or "auto",
                   "sampling": "fixed spp (adaptive stopping off: every pixel traces every sample)"},
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1), "p99": round(float(np.percentile(lat, 99)), 1)},
        "blas_build_ms": [round(b["build_ms"], 2) for b in binfo],
        "blas_nodes": [b["node_count"] for b in binfo],
        "blas_sah_cost": [round(b["sah_cost"], 1) for b in binfo],
        "rays": {"extension": stats["extension_rays"], "shadow": stats["shadow_rays"]},
        "film_finite": finite, "wall_s": round(total, 2),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    if gpu_frames is not None:
        line["tlas_build_ms_p50"] = round(float(np.percentile(tlas_ms, 50)), 2)
        line["tlas_build_ms_p99"] = round(float(np.percentile(tlas_ms, 99)), 2)
        line["frames"] = args.frames
    print(json.dumps(line), flush=True)


def denoise_bench(args):
    """Config 4: 4 spp + denoise vs plain 64 spp at equal relMSE."""
    import numpy as np
    import torch
    from vork_amd import denoise as D, native, scene as S
    t0 = time.time()
    sc = S.config1(width=args.width, height=args.height, spp=64, sphere_level=args.sphere_level, terrain=args.terrain)
    log(f"scene: {sc.triangle_count} triangles generated in {time.time() - t0:.1f}s")
    r = native.Renderer(0)
    inst = r.install(sc)
    den = D.Denoiser(r.device, open(args.weights, "rb").read() if args.weights else None)
    H, W = sc.height, sc.width
    film = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")
    guides = torch.zeros((H, W, 2, 4), dtype=torch.float32, device="cuda")
    out = torch.zeros((H, W, 4), dtype=torch.float32, device="cuda")

    def render(spp, seed, with_guides):
        film.zero_()
        p = sc.render_params(samples=spp, seed=seed, adaptive=False)
        if with_guides:
            guides.zero_()
            inst.render_guides(p, film, guides, device=True)
        else:
            inst.render(p, device_film=film)

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t) * 1e3)
        return ts

    t = time.time()
    render(args.ref_spp, 90_001, False)
    ref = film[..., :3].cpu().numpy() / args.ref_spp
    log(f"reference {args.ref_spp} spp in {time.time() - t:.1f}s")

    def rel(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))

    lat64 = timed(lambda: render(64, 7, False), 3)
    img64 = film[..., :3].cpu().numpy() / 64
    lat4 = timed(lambda: (render(4, 5, True), den(film, guides, 4, out=out)), 5)
    img4 = out[..., :3].cpu().numpy()
    raw4 = film[..., :3].cpu().numpy() / 4
    render(4, 5, True)
    dn_ms = timed(lambda: den(film, guides, 4, out=out), 10)
    flops = D.flops(W, H)
    dn = float(np.median(dn_ms))
    achieved = flops / (dn * 1e-3) / 1e12
    e4, e64, eraw = rel(img4), rel(img64), rel(raw4)

    # one step of the hand-written trainer (forward, loss, backward, Adam)
    # over crops of the 4-spp render and its reference, all resident
    B, Cr = args.train_batch, args.train_crop
    corners = [((k * 389) % (H - Cr), (k * 677) % (W - Cr)) for k in range(B)]
    ref_d = torch.from_numpy(np.ascontiguousarray(ref)).cuda()
    tf = torch.stack([film[y:y + Cr, x:x + Cr] for y, x in corners]).contiguous()
    tg = torch.stack([guides[y:y + Cr, x:x + Cr] for y, x in corners]).contiguous()
    tt = torch.stack([ref_d[y:y + Cr, x:x + Cr] for y, x in corners]).contiguous()
    tr = D.Trainer(r.device, None, seed=1)
    for _ in range(2):
        tr.step(tf, tg, tt, 4, 1e-4)
    train_ms = float(np.median(timed(lambda: tr.step(tf, tg, tt, 4, 1e-4), 10)))
    tr.close()
    # forward + weight gradients + input gradients (all layers but the first)
    train_flops = B * (3 * D.flops(Cr, Cr) - 2 * Cr * Cr * 9 * 32 * D.LEVELS[0])
    train_tflops = train_flops / (train_ms * 1e-3) / 1e12
    p50_4, p50_64 = float(np.median(lat4)), float(np.median(lat64))
    equal = e4 <= e64
    return {
        "metric": "Msamples/s @1080p + p50/p99 render latency at equal relMSE vs 4096-spp oracle",
        "value": round(W * H * 64 / (p50_4 * 1e-3) / 1e6, 1) if equal else None, "unit": "Msamples/s",
        "n_gpus": 1, "steps": len(lat4), "warmup": 0, "ms_per_step": round(p50_4, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 render, bf16 U-Net",
        "data": "synthetic (seeded procedural scene); U-Net weights trained on path-traced pairs of seeded scenes",
        "config": {"workload": "config4: config-1 scene at 1920x1080, 4 spp + guides + 5-level bf16 U-Net vs plain "
                               "64 spp, relMSE against a 4096-spp render", "width": W, "height": H,
                   "ref_spp": args.ref_spp, "value_is": "64-spp samples per second of the denoised path (equal relMSE)"},
        "relmse": {"denoised_4spp": round(e4, 6), "plain_64spp": round(e64, 6), "raw_4spp": round(eraw, 6)},
        "equal_relmse": equal,
        "latency_ms": {"denoised_4spp_p50": round(p50_4, 2), "plain_64spp_p50": round(p50_64, 2),
                       "speedup": round(p50_64 / p50_4, 2)},
        "roofline": {"kernel": "U-Net (19 bf16 MFMA convolution launches; the first computes its input features, "
                                "the last writes the radiance, encoder convolutions write their 2x2 pools)", "bound": "mfma",
                     "achieved": round(achieved, 1), "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": round(achieved / 2500.0, 4), "traffic": None, "flops_per_launch": flops,
                     "denoise_ms_p50": round(dn, 3),
                     "speed_of_light_ms": round(D.speed_of_light_ms(W, H), 4),
                     "frac_of_speed_of_light": round(D.speed_of_light_ms(W, H) / dn, 4),
                     "speed_of_light": "sum over the 19 launches of max(flops / 2.5 PFLOP/s, least HBM bytes / "
                                       "8 TB/s) (vork_amd.denoise.layer_costs)"},
        "training": {"step": f"vh_trainer_step over {B} crops of {Cr}x{Cr} (forward, loss, backward, Adam)",
                     "ms_per_step": round(train_ms, 3), "flops_per_step": train_flops,
                     "achieved_tflops": round(train_tflops, 1), "frac_bf16_peak": round(train_tflops / 2500.0, 4)},
        "weights": args.weights or "shipped"}


def serve_bench(args):
    """Config 3: a Poisson stream of requests through vork_amd.serve."""
    import numpy as np
    from vork_amd import scene as S, serve
    t0 = time.time()
    specs = [("c0-1", lambda: S.config0(seed=1)), ("c0-2", lambda: S.config0(seed=2)),
             ("g4", lambda: S.config1(seed=4, sphere_level=4, terrain=64, materials=8)),
             ("g6", lambda: S.config1(seed=6, sphere_level=6, terrain=128, materials=16)),
             ("g7", lambda: S.config1(seed=7, sphere_level=7, terrain=256, materials=16)),
             ("g8", lambda: S.config1(seed=8, sphere_level=8, terrain=512)),
             ("g9", lambda: S.config1(seed=9, sphere_level=9, terrain=32))]
    pool = {k: f() for k, f in specs}
    tris = {k: sc.triangle_count for k, sc in pool.items()}
    log(f"scene pool: {tris} generated in {time.time() - t0:.1f}s")
    keys = sorted(pool, key=lambda k: tris[k])
    logt = np.log([tris[k] for k in keys])
    rng = np.random.default_rng(3)
    resolutions = [(256, 256), (512, 512), (1280, 720), (1920, 1080)]
    spps = [1, 4, 16, 64]
    mix = []
    for k in range(args.requests):
        u = rng.uniform(np.log(10_000), np.log(5_300_000))
        key = keys[int(np.argmin(np.abs(logt - u)))]
        w, h = resolutions[rng.integers(len(resolutions))]
        mix.append((key, w, h, spps[rng.integers(len(spps))], 1 + k))
    backend = serve.GpuBackend(0, slots=args.slots)

    def stream(server, rate):
        server.finished = []
        gaps = rng.exponential(1.0 / rate, len(mix)) if rate != float("inf") else np.zeros(len(mix))
        t = time.perf_counter()
        for (key, w, h, spp, seed), gap in zip(mix, gaps):
            t += gap
            delay = t - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            server.submit(key, pool[key], w, h, spp, seed)
        server.wait_all(600)
        return server.report()

    def warm(server):
        for key in keys:   # every scene resident before timing
            server.submit(key, pool[key], 32, 32, 1, 0).result(600)

    def sweep(server, name):
        res = {}
        warm(server)
        for rate in [float(r) for r in args.rates.split(",")]:
            rep = stream(server, rate)
            res["inf" if rate == float("inf") else f"{rate:g}/s"] = {
                "msamples_per_s": round(rep["msamples_per_s"], 2), "wall_s": round(rep["wall_s"], 3),
                "latency_ms": {k: round(v, 1) for k, v in rep["latency_ms"].items()}}
            log(f"{name} rate {rate}: {res[list(res)[-1]]}")
        server.close()
        return res

    # two designs of the same loop: one HIP stream per slot (kernels of the
    # admitted requests overlap on the device), and one shared wavefront
    # (the admitted requests' paths in the same queues, vh_render_shared)
    server = serve.Server(backend, slots=args.slots)
    streams_res = sweep(server, "streams")
    shared_server = serve.Server(backend, slots=16, shared=True)
    shared_res = sweep(shared_server, "shared")
    best_shared = (shared_res.get("inf") or {}).get("msamples_per_s", 0) > \
        (streams_res.get("inf") or {}).get("msamples_per_s", 0)
    out = shared_res if best_shared else streams_res
    serial = serve.Server(backend, slots=1)
    warm(serial)
    srep = stream(serial, float("inf"))
    serial.close()
    sat = out.get("inf") or out[list(out)[-1]]
    samples = sum(w * h * spp for _, w, h, spp, _ in mix)
    return {
        "metric": "Msamples/s @1080p + p50/p99 render latency at equal relMSE vs 4096-spp oracle",
        "value": sat["msamples_per_s"], "unit": "Msamples/s", "n_gpus": 1, "steps": len(mix), "warmup": len(keys),
        "ms_per_step": round(sat["wall_s"] * 1e3 / len(mix), 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded procedural scenes, generated on the host, installed before timing)",
        "config": {"workload": "config3 serving: Poisson request stream over resident scenes of "
                               f"{min(tris.values())}-{max(tris.values())} tris, resolutions 256^2..1920x1080, "
                               "spp 1..64, fixed spp", "requests": len(mix), "samples": samples,
                   "slots": args.slots, "scenes": tris},
        "latency_ms": sat["latency_ms"], "rates": out,
        "design": "shared wavefront queues (16 requests per call)" if best_shared else
                  f"one HIP stream per slot ({args.slots} slots)",
        "designs": {"streams": streams_res, "shared": shared_res},
        "serial": {"msamples_per_s": round(srep["msamples_per_s"], 2),
                   "latency_ms": {k: round(v, 1) for k, v in srep["latency_ms"].items()}},
        "evictions": server.evictions}


CPU_SHARE = 16   # host cores of one GPU's share of a multi-GPU host


def cpu_baseline(sc, samples=8):
    """The CPU oracle (a restatement of the estimator) on the host cores, over a
    bounded sample of the same workload: exposure samples [0, samples) of
    every pixel, about 10 s of CPU work."""
    import oracle
    from vork_amd import native  # noqa: F401  (the oracle shares the ABI structs)
    visible = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = min(visible, CPU_SHARE)
    t = time.time()
    o = oracle.OracleScene(sc)
    build_s = time.time() - t
    p = sc.render_params(adaptive=False)
    p.sample_begin, p.sample_end = 0, samples
    t = time.perf_counter()
    o.render(p, threads=threads, film32=False)
    dt = time.perf_counter() - t
    n = sc.width * sc.height * samples
    return {"value": round(n / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
            "host_cpus_visible": visible,
            "sample": f"exposure samples 0-{samples - 1} of {sc.spp} for all {sc.width}x{sc.height} pixels "
                      f"({n} samples) in {dt:.1f}s on {threads} threads (all host cores of the run, at most the "
                      f"one GPU's share of {CPU_SHARE}); oracle BVH build {build_s:.1f}s untimed"}


def _rocprof_counter(counter, argv, outdir, timeout=300):
    """Runs `argv` (a python command) under rocprofv3 --pmc <counter> and
    returns {kernel name: [value per dispatch]} from its CSV."""
    return _rocprof_counters([counter], argv, outdir, timeout)[counter]


def _rocprof_counters(counters, argv, outdir, timeout=300):
    """Runs `argv` under one rocprofv3 --pmc pass over `counters` and returns
    {counter: {kernel name: [value per dispatch]}}."""
    import csv
    import glob
    import shutil
    import subprocess
    exe = shutil.which("rocprofv3")
    if not exe:
        raise RuntimeError("rocprofv3 not found")
    if os.path.isdir(outdir):
        shutil.rmtree(outdir)
    cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", outdir, "-o", "run", "--", *argv]
    env = dict(os.environ, TMPDIR=os.environ.get("TMPDIR", "/tmp"))
    r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=timeout)
    files = glob.glob(os.path.join(outdir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise RuntimeError(f"rocprofv3 --pmc {' '.join(counters)} wrote no counter CSV (rc {r.returncode}): "
                           + r.stdout[-400:].replace("\n", " | "))
    values = {c: {} for c in counters}
    for f in files:
        with open(f) as fh:
            rows = list(csv.DictReader(fh))
        per = {}
        for row in rows:
            c = row.get("Counter_Name")
            if c not in values:
                continue
            key = (c, row.get("Dispatch_Id"), row.get("Kernel_Name", ""))
            per[key] = per.get(key, 0.0) + float(row["Counter_Value"])
        for (c, _, name), v in per.items():
            values[c].setdefault(name, []).append(v)
    return values


def measure_traffic(args):
    """HBM bytes per launch of the traced kernels, from rocprofv3 FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch) over one render step of the same workload,
    each scaled by what the counter reads for a plain device copy of known
    size.  Runs in child processes; returns the figures or the error."""
    out = {}
    base = os.path.join(ROOT, "gpurun_out", "bench_traffic")
    me = [sys.executable, os.path.abspath(__file__)]
    step = me + ["--traffic-pass", "--workload", args.workload, "--width", str(args.width), "--height",
                 str(args.height), "--spp", str(args.spp), "--sphere-level", str(args.sphere_level), "--terrain",
                 str(args.terrain), "--batch", str(args.batch)]
    copy = me + ["--copy-calibration", str(COPY_MIB)]
    try:
        t = time.time()
        scale = {}
        for counter, truth in (("FETCH_SIZE", COPY_MIB << 20), ("WRITE_SIZE", COPY_MIB << 20)):
            vals = _rocprof_counter(counter, copy, os.path.join(base, "copy_" + counter))
            kib = [v for name, vs in vals.items() if "copy" in name.lower() or "add" in name.lower() for v in vs]
            if not kib:
                raise RuntimeError(f"no copy kernel in the {counter} calibration run: {sorted(vals)[:4]}")
            scale[counter] = truth / (max(kib) * 1024.0)
            log(f"traffic: {counter} copy calibration {max(kib):.0f} KiB for {truth >> 20} MiB")
        kernels = {}
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            vals = _rocprof_counter(counter, step, os.path.join(base, "step_" + counter))
            for key, pattern in TRAFFIC_KERNELS.items():
                got = [v for name, vs in vals.items() if pattern in name for v in vs]
                if got:
                    k = kernels.setdefault(key, {"launches": len(got)})
                    k[counter] = sum(got) / len(got) * 1024.0 * scale[counter]
            log(f"traffic: {counter} pass done")
        for k in kernels.values():
            k["bytes_per_launch"] = int(k.get("FETCH_SIZE", 0.0) + k.get("WRITE_SIZE", 0.0))
            k["read_bytes_per_launch"] = int(k.pop("FETCH_SIZE", 0.0))
            k["written_bytes_per_launch"] = int(k.pop("WRITE_SIZE", 0.0))
        # L2 (TCC) hit rate of each kernel over the same step: hits / (hits + misses)
        # summed over its launches and the L2 instances
        hm = _rocprof_counters(["TCC_HIT_sum", "TCC_MISS_sum"], step, os.path.join(base, "step_TCC_HIT_MISS"))
        for key, pattern in TRAFFIC_KERNELS.items():
            hits = sum(v for name, vs in hm["TCC_HIT_sum"].items() if pattern in name for v in vs)
            miss = sum(v for name, vs in hm["TCC_MISS_sum"].items() if pattern in name for v in vs)
            if key in kernels and hits + miss > 0:
                kernels[key]["l2_hits"] = int(hits)
                kernels[key]["l2_misses"] = int(miss)
                kernels[key]["l2_hit_rate"] = round(hits / (hits + miss), 4)
        log("traffic: L2 hit/miss pass done")
        out = {"kernels": kernels, "fetch_scale": round(scale["FETCH_SIZE"], 4),
               "write_scale": round(scale["WRITE_SIZE"], 4), "seconds": round(time.time() - t, 1)}
    except Exception as e:  # reported, never required
        out = {"error": f"{type(e).__name__}: {e}"}
    return out


if __name__ == "__main__":
    main()
