"""`python -m vork_amd.render SCENE --out FILE [--set KEY=VALUE]...`: the
render call of the std/path hot path on one GPU (the subset of
`vork render SCENE --out FILE --set ...`, crates/vork-cli/src/main.rs:103-171).

Loads scene.json (vork_amd.scene_json), installs it on the device (BLASes,
TLAS, materials, lights), renders samples [0, samples) into an RGBA32F film
and writes FILE (.png display codes through the agx or standard view, .exr
scene-linear mean) plus FILE.meta.json.  Exit codes follow the reference's
failure kinds: 2 usage, 3 scene, 4 device, 5 io.
"""
import argparse
import os
import sys
import time


def render_image(device, inst, scene, params, denoiser=None, stream=None):
    """The image of a render call: the film sums, or with the scene's
    `denoise` setting the denoised mean of samples with guides
    (crates/vork-render/src/denoise.rs).  Returns (image (H, W, 4), stats,
    denoised)."""
    if scene.notes.get("denoise"):
        from .denoise import Denoiser
        film, guides = inst.render_guides(params, stream=stream)
        stats = inst.stats()
        den = denoiser if denoiser is not None else Denoiser(device)
        return den(film, guides, params.samples), stats, True
    img = inst.render(params, stream=stream)
    return img, inst.stats(), False


def write_render(scene, outs, img, stats, samples, wall_ms, view="agx", agx_lut=None, warnings=(),
                 adaptive=True, denoised=False):
    """Writes the film to each path of `outs` (+ .meta.json).  Returns the
    sidecar (the same for every output)."""
    from . import film as F
    warnings = list(warnings)
    lut = F.load_agx_lut(agx_lut) if agx_lut else None
    if view == "agx" and lut is None and any(o.lower().endswith(".png") for o in outs):
        warnings.append("view agx: analytic AgX (no AgX LUT given); codes differ from the baked view")
    meta = F.metadata(scene, stats, warnings, samples, wall_ms, adaptive=adaptive, denoised=denoised)
    for out in outs:
        F.write_outputs(out, img, view=view, lut=lut, meta=meta)
    return meta


def render_scene(scene, out, view="agx", agx_lut=None, warnings=()):
    """Renders a loaded Scene on device 0 and writes `out` (+ .meta.json).
    Returns the film (H, W, 4) and the render stats."""
    from . import native
    t0 = time.perf_counter()
    r = native.Renderer(0)
    inst = r.install(scene)
    p = scene.render_params()
    img, stats, denoised = render_image(r.device, inst, scene, p)
    wall = (time.perf_counter() - t0) * 1e3
    write_render(scene, [out], img, stats, p.samples, wall, view, agx_lut, warnings, denoised=denoised)
    return img, stats


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m vork_amd.render")
    ap.add_argument("scene")
    ap.add_argument("--out", required=True)
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE")
    ap.add_argument("--view", choices=["agx", "standard"], default="agx")
    ap.add_argument("--agx-lut", default=os.environ.get("VORK_AGX_LUT"),
                    help="vork's libraries/std/agx.exr for the exact AgX view")
    args = ap.parse_args(argv)
    from . import scene_json as J
    from .native import VorkError
    try:
        scene, warnings = J.load(args.scene, J.parse_set(args.set))
    except J.SceneError as e:
        print(f"vork_amd.render: scene: {e}", file=sys.stderr)
        return 2 if str(e).startswith("--set") else 3
    except OSError as e:
        print(f"vork_amd.render: io: {e}", file=sys.stderr)
        return 5
    try:
        render_scene(scene, args.out, args.view, args.agx_lut, warnings)
    except VorkError as e:
        print(f"vork_amd.render: device: {e}", file=sys.stderr)
        return e.status if e.status in (2, 3, 4, 5) else 4
    except OSError as e:
        print(f"vork_amd.render: io: {e}", file=sys.stderr)
        return 5
    return 0


if __name__ == "__main__":
    sys.exit(main())
