"""Train the learned-strength network on the GPU and write the package's weights.

    python -m climage_amd.train_strength [--frames 122] [--epochs 30] [--out climage_amd/weights]

Training runs inside the library (clim_strength_train): seeded synthetic
4032x3024 RGGB frames and their clean scenes go through the plain pipeline;
every pyramid level is denoised at seven constant strengths and compared with
the clean scene's pyramid, giving the error curves of 32x32 crops of the
level-3 grid (256x256 frame pixels; 165 crops a frame, 20130 for 122 frames);
the network (BF16 MFMA kernels) is fitted to them with Adam.  The script then
measures what the trained network gains: PSNR against the clean truth (the
clean scene tone-mapped by the pipeline's own tone map) of the plain, gated
and forced-learned pipelines on held-out frames at the six noise levels.
Writes strength.f32 (the parameters, little-endian float32, canonical
layout) and strength.json (configuration, losses, gain) into --out.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import climage_amd as ca  # noqa: E402

NOISE_A = [0.0005, 0.001, 0.002, 0.004, 0.008, 0.016]


def psnr(a, b):
    mse = float(np.mean((a[..., :3].astype(np.float64) - b[..., :3]) ** 2))
    return 10.0 * np.log10(1.0 / max(mse, 1e-20))


def evaluate(ctx, w, h, seeds, algorithm=ca.BILATERAL):
    """PSNR against the clean truth of the plain, gated and forced-learned
    pipelines on frames (scene seed, noise seed, a); returns per-frame rows."""
    raw = ctx.image(w, h, ca.FMT_R16U)
    clean = ctx.image(w, h, ca.FMT_RGBA32F)
    o8 = ctx.image(w, h, ca.FMT_RGBA8)
    of = ctx.image(w, h, ca.FMT_RGBA32F)
    truth = ctx.image(w, h, ca.FMT_RGBA32F)
    rows = []
    for scene, noise, a in seeds:
        ctx.generate_bayer(raw, ca.RGGB, scene, noise, a, a / 200.0, 64.0, 4095.0, clean=clean)
        p = ca.default_params(ca.RGGB, 64.0, 4095.0, levels=5, algorithm=algorithm)
        ident = np.eye(3, dtype=np.float32)
        ctx.convertTosRGB(clean, o8, ident, p.rgbConversionParameters, out_f32=truth)
        t = truth.download()
        row = {"scene": scene, "a": a}
        for name, mode in (("plain", ca.STRENGTH_OFF), ("gated", ca.STRENGTH_GATED),
                           ("learned", ca.STRENGTH_FORCE_LEARNED)):
            ctx.strength_set_mode(mode)
            ctx.raw_to_rgb(raw, o8, p, out_f32=of, writeback=False)
            row[name] = round(psnr(of.download(), t), 4)
            if mode == ca.STRENGTH_GATED:
                row["gate_open"] = ctx.strength_gate_read()[0].tolist()
        ctx.strength_set_mode(ca.STRENGTH_OFF)
        rows.append(row)
        print(f"  eval scene {scene} a {a}: plain {row['plain']:.3f} gated {row['gated']:.3f} "
              f"learned {row['learned']:.3f} dB, gate {row['gate_open']}", flush=True)
    for im in (raw, clean, o8, of, truth):
        im.free()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=122)
    ap.add_argument("--width", type=int, default=4032)
    ap.add_argument("--height", type=int, default=3024)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=10000)
    ap.add_argument("--eval-frames", type=int, default=6)
    ap.add_argument("--out", default=os.path.join(ROOT, "climage_amd", "weights"))
    args = ap.parse_args()
    with ca.Context(0) as ctx:
        t0 = time.perf_counter()
        print(f"training: {args.frames} frames {args.width}x{args.height}, {args.epochs} epochs", flush=True)
        log = ctx.strength_train(frames=args.frames, width=args.width, height=args.height, crop=32,
                                 epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, verbose=True)
        dt = time.perf_counter() - t0
        for e, (ls, lr) in enumerate(log):
            print(f"  epoch {e}: strength loss {ls:.5f} (plain = 1), prediction loss {lr:.5f}", flush=True)
        params = ctx.strength_get_weights()
        os.makedirs(args.out, exist_ok=True)
        params.astype("<f4").tofile(os.path.join(args.out, "strength.f32"))
        # held-out frames: the benchmark's scene seeds (1000 + i), noise seeds 2000 + i
        seeds = [(1000 + i, 2000 + i, NOISE_A[i % 6]) for i in range(args.eval_frames)]
        rows = evaluate(ctx, args.width, args.height, seeds)
        gain = {k: round(float(np.mean([r[k] - r["plain"] for r in rows])), 4) for k in ("gated", "learned")}
        print(f"mean PSNR gain over plain: gated {gain['gated']} dB, learned {gain['learned']} dB", flush=True)
        meta = {"frames": args.frames, "size": [args.width, args.height], "crop": 32, "grid_level": 3,
                "crops": args.frames * (args.width // 8 // 32) * (args.height // 8 // 32), "epochs": args.epochs,
                "batch": args.batch, "lr": args.lr, "seed": args.seed, "train_seconds": round(dt, 2),
                "loss_log": log.round(6).tolist(), "eval": rows, "mean_gain_db": gain}
        with open(os.path.join(args.out, "strength.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print(json.dumps({"train_seconds": round(dt, 2), "mean_gain_db": gain}))


if __name__ == "__main__":
    main()
