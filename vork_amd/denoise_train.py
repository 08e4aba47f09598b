"""Training of the denoiser on pairs the path tracer renders itself, with the
hand-written HIP training kernels (vh_trainer_*: forward pass, loss, backward
pass and Adam on the device).

`python -m vork_amd.denoise_train --out weights.vhdn --report report.json`

Scenes come from the config 0, 1 and 2 generators with random seeds and
orbiting cameras, generated on the host by a pool of worker processes while
the GPU renders.  Each scene gives an input (4 samples with guides) and a
target (1024 samples of an independent seed), both rendered by
vork_amd.native at fixed spp and kept on the device.  Training steps draw
random crops and mirror images from the rendered pairs and call
vork_amd.denoise.Trainer.  `--init` continues from a blob (fresh Adam
moments), so a long run can be split over several invocations with
`--first-scene` naming the scenes each one renders.  Held-out scenes then
score the denoised 4-spp images, the raw 4-spp and a plain 64-spp render
against a high-spp reference (relMSE).
"""
import argparse
import json
import math
import multiprocessing
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def random_scene(rng, size):
    from vork_amd import scene as S
    seed = int(rng.integers(1, 1 << 30))
    kind = int(rng.integers(3))
    w, h = size
    if kind == 0:
        sc = S.config0(seed=seed, width=w, height=h)
    elif kind == 1:
        sc = S.config1(seed=seed, width=w, height=h, sphere_level=int(rng.integers(4, 7)),
                       terrain=int(rng.choice([32, 64, 128])), materials=16)
    else:
        sc = S.config2(seed=seed, width=w, height=h, birds=int(rng.integers(500, 4000)),
                       frame=int(rng.integers(0, 48)))
        return sc
    t = np.array(sc.camera["target"], np.float64)
    d = np.array(sc.camera["eye"], np.float64) - t
    r = np.linalg.norm(d) * rng.uniform(0.55, 1.3)
    az = math.atan2(d[0], d[2]) + rng.uniform(-math.pi, math.pi)
    el = float(np.clip(math.asin(d[1] / np.linalg.norm(d)) + rng.uniform(-0.25, 0.35), 0.05, 1.2))
    eye = t + r * np.array([math.sin(az) * math.cos(el), math.sin(el), math.cos(az) * math.cos(el)])
    sc.camera = dict(sc.camera, eye=tuple(float(v) for v in eye))
    return sc


def _make_scene(job):
    """Worker: scene `index` of the run's stream (seeded by the run seed and
    the index, so any split of the run makes the same scenes)."""
    seed, index, size = job
    return index, random_scene(np.random.default_rng([seed, index]), (size, size))


def rel_mse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default=None)
    ap.add_argument("--init", default=None, help="weight blob to start from (else a seeded initialisation)")
    ap.add_argument("--scenes", type=int, default=20000)
    ap.add_argument("--first-scene", type=int, default=0)
    ap.add_argument("--size", type=int, default=96)
    ap.add_argument("--spp-ref", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1.5e-3)
    ap.add_argument("--workers", type=int, default=14)
    ap.add_argument("--eval-scenes", type=int, default=8)
    ap.add_argument("--eval-size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=11)
    args = ap.parse_args(argv)
    import torch
    from vork_amd import denoise as D, native
    renderer = native.Renderer(0)
    dev = torch.device("cuda")
    n, s = args.scenes, args.size
    films = torch.empty((n, s, s, 4), dtype=torch.float32, device=dev)
    guides = torch.empty((n, s, s, 2, 4), dtype=torch.float32, device=dev)
    targets = torch.empty((n, s, s, 3), dtype=torch.float32, device=dev)
    t0 = time.time()
    jobs = [(args.seed, args.first_scene + k, s) for k in range(n)]
    samples_rendered = 0
    with multiprocessing.get_context("spawn").Pool(args.workers) as pool:
        for k, (index, sc) in enumerate(pool.imap(_make_scene, jobs, chunksize=4)):
            inst = renderer.install(sc)
            p = sc.render_params(samples=4, seed=1 + index, adaptive=False)
            f = films[k]
            f.zero_()
            g = guides[k]
            g.zero_()
            inst.render_guides(p, f, g, device=True)
            ref = torch.zeros((s, s, 4), dtype=torch.float32, device=dev)
            inst.render(sc.render_params(samples=args.spp_ref, seed=1 + index + 7919, adaptive=False),
                        device_film=ref)
            targets[k] = ref[..., :3] / args.spp_ref
            samples_rendered += s * s * (4 + args.spp_ref)
            inst.close()
            if (k + 1) % 500 == 0:
                torch.cuda.synchronize()
                log(f"rendered {k + 1}/{n} training pairs, {time.time() - t0:.0f}s")
    torch.cuda.synchronize()
    t1 = time.time()
    blob = open(args.init, "rb").read() if args.init else None
    tr = D.Trainer(renderer.device, blob, seed=args.seed)
    rng = np.random.default_rng(args.seed + 17)
    c = args.crop
    losses = []
    for step in range(args.steps):
        idx = torch.from_numpy(rng.integers(0, n, args.batch)).to(dev)
        y0, x0 = (int(v) for v in rng.integers(0, s - c + 1, 2))
        f = films[idx, y0:y0 + c, x0:x0 + c]
        g = guides[idx, y0:y0 + c, x0:x0 + c]
        t = targets[idx, y0:y0 + c, x0:x0 + c]
        if rng.random() < 0.5:   # a mirrored image mirrors the normals' x
            f, g, t = f.flip(2), g.flip(2).clone(), t.flip(2)
            g[..., 1, 0] = -g[..., 1, 0]
        # cosine decay from the base rate, after a short warm-up
        lr = args.lr * min(1.0, (step + 1) / 200) * 0.5 * (1 + math.cos(math.pi * step / args.steps))
        losses.append(tr.step(f.contiguous(), g.contiguous(), t.contiguous(), 4, lr))
        if (step + 1) % 500 == 0:
            log(f"step {step + 1}/{args.steps}: loss {np.mean(losses[-500:]):.4f}, {time.time() - t1:.0f}s")
    blob = tr.export()
    tr.close()
    with open(args.out, "wb") as fh:
        fh.write(blob)
    t2 = time.time()
    # held-out scenes (indices past every training scene): denoised 4 spp vs
    # raw 4 spp vs plain 64 spp against a high-spp reference
    den = D.Denoiser(renderer.device, blob)
    scores = []
    for k in range(args.eval_scenes):
        sc = random_scene(np.random.default_rng([args.seed, 10_000_000 + k]), (args.eval_size, args.eval_size))
        inst = renderer.install(sc)
        f4, g4 = inst.render_guides(sc.render_params(samples=4, seed=50_000 + k, adaptive=False))
        f64 = inst.render(sc.render_params(samples=64, seed=60_000 + k, adaptive=False))
        ref = inst.render(sc.render_params(samples=1024, seed=70_000 + k, adaptive=False))[..., :3] / 1024
        inst.close()
        out = den(f4, g4, 4)[..., :3]
        scores.append({"denoised_4spp": rel_mse(out, ref), "raw_4spp": rel_mse(f4[..., :3] / 4, ref),
                       "plain_64spp": rel_mse(f64[..., :3] / 64, ref)})
    den.close()
    report = {"scenes": n, "first_scene": args.first_scene, "size": s, "spp_in": 4, "spp_ref": args.spp_ref,
              "steps": args.steps, "batch": args.batch, "crop": c, "init": args.init,
              "data_s": round(t1 - t0, 1), "render_msamples_per_s": round(samples_rendered / (t1 - t0) / 1e6, 1),
              "train_s": round(t2 - t1, 1), "ms_per_step": round((t2 - t1) * 1e3 / max(args.steps, 1), 2),
              "loss_first_500": float(np.mean(losses[:500])) if losses else None,
              "loss_last_500": float(np.mean(losses[-500:])) if losses else None,
              "trainer": "hand-written HIP kernels (vh_trainer_step: forward, loss, backward, Adam)",
              "eval": scores, "mean": {k: float(np.mean([sc_[k] for sc_ in scores])) for k in scores[0]}}
    log(json.dumps(report["mean"]))
    if args.report:
        with open(args.report, "w") as fh:
            json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
