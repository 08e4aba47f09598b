"""Training of the denoiser on pairs the path tracer renders itself.

`python -m vork_amd.denoise_train --out weights.vhdn --report report.json`

Scenes come from the config 0, 1 and 2 generators with random seeds, sizes
and orbiting cameras.  Each gives an input (4 samples with guides) and a
target (many samples, an independent seed) rendered by vork_amd.native at
fixed spp.  The U-Net of vork_amd.denoise is trained with torch on the GPU
(Adam, a relative-MSE plus log-L1 loss, random crops and flips) and written
as the blob vh_denoiser_create reads.  Held-out scenes then score the
denoised 4-spp images, the raw 4-spp and a plain 64-spp render against a
high-spp reference (relMSE).
"""
import argparse
import json
import math
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


def render_pair(renderer, sc, spp_in, spp_ref, seed):
    inst = renderer.install(sc)
    p = sc.render_params(samples=spp_in, seed=seed, adaptive=False)
    film, guides = inst.render_guides(p)
    q = sc.render_params(samples=spp_ref, seed=seed + 7919, adaptive=False)
    ref = inst.render(q)[..., :3] / spp_ref
    inst.close()
    return film, guides, ref


def rel_mse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", default=None)
    ap.add_argument("--scenes", type=int, default=240)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--spp-ref", type=int, default=512)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=96)
    ap.add_argument("--eval-scenes", type=int, default=8)
    ap.add_argument("--eval-size", type=int, default=256)
    ap.add_argument("--seed", type=int, default=11)
    args = ap.parse_args(argv)
    import torch
    from vork_amd import denoise as D, native
    rng = np.random.default_rng(args.seed)
    renderer = native.Renderer(0)
    t0 = time.time()
    films, guides, refs = [], [], []
    for k in range(args.scenes):
        sc = random_scene(rng, (args.size, args.size))
        f, g, r = render_pair(renderer, sc, 4, args.spp_ref, 1 + k)
        films.append(f)
        guides.append(g)
        refs.append(r)
        if (k + 1) % 40 == 0:
            log(f"rendered {k + 1}/{args.scenes} training pairs, {time.time() - t0:.0f}s")
    dev = torch.device("cuda")
    F_ = torch.from_numpy(np.stack(films)).to(dev)
    G_ = torch.from_numpy(np.stack(guides)).to(dev)
    R_ = torch.from_numpy(np.stack(refs)).permute(0, 3, 1, 2).contiguous().to(dev)
    X_ = D.features(F_, G_, 4.0)
    model = D.make_unet().to(dev)
    opt = torch.optim.Adam(model.parameters(), lr=1e-3)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=1.5e-3, total_steps=args.steps, pct_start=0.1)
    n, c = X_.shape[0], args.crop
    t1 = time.time()
    for step in range(args.steps):
        idx = torch.randint(0, n, (args.batch,), device=dev)
        y0, x0 = (int(v) for v in rng.integers(0, args.size - c + 1, 2))
        x = X_[idx][..., y0:y0 + c, x0:x0 + c]
        t = R_[idx][..., y0:y0 + c, x0:x0 + c]
        if rng.random() < 0.5:
            x, t = x.flip(-1), t.flip(-1)
            x = x.clone()
            x[:, 6] = -x[:, 6]   # a mirrored image mirrors the normals' x
        pred = model(x)
        d = D.radiance(pred)
        loss = ((d - t) ** 2 / (t.detach() ** 2 + 1e-2)).mean() + (pred - torch.log1p(t)).abs().mean()
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if (step + 1) % 250 == 0:
            log(f"step {step + 1}/{args.steps}: loss {loss.item():.4f}, {time.time() - t1:.0f}s")
    blob = D.export(model.cpu())
    with open(args.out, "wb") as f:
        f.write(blob)
    # held-out scenes: denoised 4 spp vs raw 4 spp vs plain 64 spp against a high-spp reference
    model = D.load_model(args.out).to(dev).eval()
    den = D.Denoiser(renderer.device, blob)
    scores = []
    erng = np.random.default_rng(args.seed + 1000)
    for k in range(args.eval_scenes):
        sc = random_scene(erng, (args.eval_size, args.eval_size))
        inst = renderer.install(sc)
        f4, g4 = inst.render_guides(sc.render_params(samples=4, seed=50_000 + k, adaptive=False))
        f64 = inst.render(sc.render_params(samples=64, seed=60_000 + k, adaptive=False))
        ref = inst.render(sc.render_params(samples=1024, seed=70_000 + k, adaptive=False))[..., :3] / 1024
        inst.close()
        out = den(f4, g4, 4)[..., :3]
        scores.append({"denoised_4spp": rel_mse(out, ref), "raw_4spp": rel_mse(f4[..., :3] / 4, ref),
                       "plain_64spp": rel_mse(f64[..., :3] / 64, ref)})
    report = {"scenes": args.scenes, "size": args.size, "spp_ref": args.spp_ref, "steps": args.steps,
              "train_s": round(time.time() - t1, 1), "data_s": round(t1 - t0, 1), "final_loss": float(loss.item()),
              "eval": scores,
              "mean": {k: float(np.mean([s[k] for s in scores])) for k in scores[0]}}
    log(json.dumps(report["mean"]))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
    den.close()


if __name__ == "__main__":
    main()
