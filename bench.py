"""RAW->RGB throughput on one MI355X.

Workload (default): 12 MP (4032x3024) RGGB uint16 12-bit mosaics, black 64,
white 4095, six noise levels a in {0.0005, 0.001, 0.002, 0.004, 0.008, 0.016},
b = a/200, seeded synthetic scenes generated on the device, 12 scenes per
noise level (72 frames, --scenes-per-level).  One step runs the full path
(scale, raw noise fit, demosaic, highlights, YCbCr, despeckle, 5-level
bilateral pyramid with per-level noise fits, tone map to RGBA8) on all the
frames with one clim_raw_to_rgb_batch call (the frames alternate between two
streams; --sequential: one clim_raw_to_rgb call per frame), about 50 ms of
work, so that even a short run times a second; inputs are resident in HBM
before timing starts.

Prints one JSON line:
* value: Mpix/s over the timed steps (at least about a second of work);
* stages: a few untimed steps with every stage bracketed by HIP events;
* roofline: the dominant single-kernel stage; frac from its launches timed
  alone (the profiled steps, one frame at a time), the same kernel's
  launches inside the batched timed steps under "batched";
* path_roofline: the whole path's algorithmic bytes (each reference stage's
  inputs read and outputs written once, FP32 texels) and its measured HBM
  bytes (the committed counter profile) per pixel, times the pixel rate;
* fp64: the first frame against the FP64 CPU restatement (max-abs and PSNR
  of the RGB output before quantisation, relative errors of the fitted noise
  models) and whether it equals the FP32 restatement bit for bit;
* learned: the same frames with the learned denoise strength (the committed
  weights, gated), its rate, the strength stage's cost, the gates and the
  PSNR gain against the clean truth on one frame;
* mixed: the mixed-precision form (FP16 level storage, packed-FP16
  bilateral weights with FP32 sums) on the same frames, its error against
  the FP64 and FP16-storage restatements, and config 2's 48 MP 14-bit frame
  at both precisions;
* burst: the uneven 16-frame burst recorded once as a HIP graph and
  replayed: Mpix/s, the frames' completion times within a replay (p50/p99)
  and the autotuner's choice for its cross-frame launches;
* cpu_baseline: the FP32 CPU restatement built -O3 -march=native, all host
  threads, on a sample of the same frames.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import climage_amd as ca  # noqa: E402

NOISE_A = [0.0005, 0.001, 0.002, 0.004, 0.008, 0.016]
ALGORITHMS = {"bilateral": ca.BILATERAL, "guided_fast": ca.GUIDED_FAST, "guided_precise": ca.GUIDED_PRECISE}
ALGORITHM_NAMES = {ca.BILATERAL: "bilateral", ca.GUIDED_FAST: "GuidedFast", ca.GUIDED_PRECISE: "GuidedPrecise"}
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0  # dense MFMA
# flops of the strength network per texel of its grid (pyramid level 3): 2 x 9 x
# (8 x 32 + 32 x 32 + 32 x 32 + 32 x 20) multiply-adds
STRENGTH_FLOP_PER_TEXEL = 2 * 9 * (8 * 32 + 32 * 32 + 32 * 32 + 32 * 20)


def path_bytes_per_px(levels):
    """Algorithmic bytes per output pixel of the whole unfused path: every
    reference stage reads each input once and writes each output once at its
    own resolution, FP32 texels, each host fit one read of its statistics
    image: 206 + 60 S(L-2) + 80 S(L-1), S(k) = sum of 4^-l for l <= k."""
    S = lambda k: sum(4.0 ** -l for l in range(k + 1))
    return 206 + 60 * S(levels - 2) + 80 * S(levels - 1)


def stage_bytes(stage, texels, preview=False, half=False):
    """Algorithmic bytes of one stage: each input read once, each output written
    once, FP32 texels (luma 4 B, RGBA 16 B; the level images' RGBA texels 8 B
    with half, the mixed-precision form's FP16 storage), raw 2 B, RGBA8 4 B."""
    L = 8 if half else 16
    per = {
        "scale_raw": 2 + 4,             # u16 in, f32 out (per pixel; stand-alone operator only)
        "raw_stats": 8 + 32,            # 4 u16 texels in, mean+var RGBA out (per quad)
        "raw_fit": 32,                  # the refit pass over mean+var (per quad)
        "green": 2 + 4,                 # u16 mosaic in (scaled on the fly), f32 green out
        "red_blue_ycc": 2 + 4 + L,      # + highlights, YCbCr and despeckle in the same kernel
        "despeckle": 16 + 16,
        "downsample": 4 * L + L,        # per output texel: 4 input texels + 1 output
        "level_stats": L + 16,          # the level in, the statistics (FP32 RGBA) out
        "level_fit": 16,                # the refit pass over the statistics image
        "denoise": L + L,
        # per texel: the level in (the noisy level, which the FP32 bilateral
        # path denoises in the same launch, or the denoised one), the next
        # level's noisy and reassembled texels at a quarter each, the result out
        "reassemble": L + L + L / 4 + L / 4,
        "tone": L + L / 4 + L / 4 + 4,  # level-0 reassembly fused with the tone map: RGBA8 out
        "raw_denoise": 8 + 16,          # 4 u16 texels in, 4 FP32 texels out (per quad)
    }
    if preview:
        per["tone"] = 8 + 4  # fastDemosaicImage: one u16 quad in, one RGBA8 texel out (per output texel)
    return per[stage] * texels


# stages that are one kernel (launched once per frame, or once per pyramid
# level), so that their event timing is that kernel's launch duration; the
# names are the kernel instances this workload runs (as rocprofv3 lists them)
SINGLE_KERNEL = {"scale_raw": "k_scale_raw", "raw_stats": "k_raw_stats", "green": "k_green",
                 "red_blue_ycc": "k_rb_despeckle<true, false>", "despeckle": "k_despeckle",
                 "downsample": "k_downsample<true, false>",
                 "denoise": "k_bilateral_levels<true>", "reassemble": "k_bilateral_reassemble<true>",
                 "tone": "k_reassemble<true, true, false, false>", "raw_denoise": "k_raw_rgba_filter<2, true>"}
# the same stages' kernels in the mixed-precision form
SINGLE_KERNEL_MIXED = dict(SINGLE_KERNEL, **{
    "red_blue_ycc": "k_rb_despeckle<true, true>", "downsample": "k_downsample<true, true>",
    "denoise": "k_bilateral_levels_mixed<true>", "reassemble": "k_reassemble<true, false, false, true>",
    "tone": "k_reassemble<true, true, false, true>"})


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, h, bayer, ltm=False, noise_level=0.0, algorithm=0, seconds_min=10.0):
    """The FP32 CPU restatement built for this host (-O3 -march=native, OpenMP
    row bands over every host thread) on the bench frames (same sizes and
    noise levels), as many as fit in about `seconds_min` seconds (at least one)."""
    from oracle import oracle as O
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "fast"], check=True)
    o = O.Oracle(double=False, lib_path=os.path.join(ROOT, "oracle", "_build", "liboracle_fast.so"))
    p = O.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    p.localToneMapping = int(ltm)
    p.noiseLevel = noise_level
    raws = [O.generate_bayer(w, h, bayer, 1000 + i, 2000 + i, a, a / 200.0, 64.0, 4095.0)
            for i, a in enumerate(NOISE_A)]
    o.raw_to_rgb(raws[0], p)  # first touch of the buffers, thread start-up
    done, dt = 0, 0.0
    while dt < seconds_min and done < 8 * len(raws):
        raw = raws[done % len(raws)]
        t0 = time.perf_counter()
        o.raw_to_rgb(raw, p)
        dt += time.perf_counter() - t0
        done += 1
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(done * w * h / dt / 1e6, 2), "unit": "Mpix/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(), "nproc": os.cpu_count(), "build": "gcc -O3 -march=native -fopenmp",
            "sample": f"{done} frames ({w}x{h} RGGB, the {len(NOISE_A)} bench noise levels in turn) through the "
                      f"FP32 CPU restatement in {dt:.1f} s, OpenMP row bands on {cores} threads"}


def fp64_parity(raw_host, bayer, gpu_f32, gpu_nlf, levels=5, algorithm=0):
    """The GPU frame against the FP64 and FP32 CPU restatements."""
    from oracle import oracle as O
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    p = O.default_params(bayer, 64.0, 4095.0, levels=levels, algorithm=algorithm)
    r64 = O.Oracle(double=True).raw_to_rgb(raw_host, p)
    r32 = O.Oracle(double=False).raw_to_rgb(raw_host, p)
    d = gpu_f32[..., :3].astype(np.float64) - r64["rgbf"][..., :3]
    mse = float(np.mean(d * d))
    raw_nlf, pyr_nlf = gpu_nlf

    def rel(a, b):
        a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
        return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))

    return {"frame": "first bench frame (a = 0.0005)",
            "rgb": {"max_abs": float(np.abs(d).max()), "psnr_db": round(10 * np.log10(1.0 / max(mse, 1e-30)), 2),
                    "pixels_above_1e-3": int(np.sum(np.abs(d) > 1e-3))},
            "raw_nlf_max_rel": rel(raw_nlf, r64["rawNlf"]),
            "pyramid_nlf_max_rel": rel(pyr_nlf, r64["pyramidNlf"]),
            "bit_exact_vs_fp32": bool(np.array_equal(gpu_f32.view(np.uint32), r32["rgbf"].view(np.uint32)))}


def measured_traffic(kernel, w, h):
    """HBM bytes per dispatch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by the
    copy calibration) from the committed counter summary of the default
    workload, or None for another workload or an unlisted kernel."""
    path = os.path.join(ROOT, "profiles", "traffic_raw_to_rgb_12mp.json")
    if (w, h) != (4032, 3024) or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)["per_dispatch"].get(kernel)


def measured_frame_bytes(w, h):
    """HBM bytes of one frame (every dispatch) from the committed counter summary."""
    path = os.path.join(ROOT, "profiles", "traffic_raw_to_rgb_12mp.json")
    if (w, h) != (4032, 3024) or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("frame_bytes")


def copy_calibration(nbytes):
    """hipMemset then hipMemcpy (device to device) of nbytes: the runtime's fill
    kernel writes nbytes, its copy kernel reads and writes nbytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(nbytes)) != 0 or \
                hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("copy calibration failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


BURST_SIZES = [(4032, 3024), (4000, 3000), (3264, 2448), (1920, 1080)]


def burst_leg(ctx, replays=20):
    """The uneven burst (16 frames: 4 each of 4032x3024, 4000x3000, 3264x2448
    and 1920x1080, RGGB and BGGR alternating, noise levels drawn from the
    headline set) recorded once as a HIP graph (clim_burst_create: the
    cross-frame persistent green and red/blue launches, autotuned) and
    replayed: Mpix/s over the replays (host clock, inputs resident), and the
    frames' completion times within a replay (HIP events in the graph)."""
    draws = np.random.default_rng(3000).integers(0, len(NOISE_A), 16)
    frames = [(*BURST_SIZES[k // 4], ca.RGGB if k % 2 == 0 else ca.BGGR, NOISE_A[draws[k]]) for k in range(16)]
    raws, outs, params = [], [], []
    for k, (fw, fh, fb, a) in enumerate(frames):
        r = ctx.image(fw, fh, ca.FMT_R16U)
        ctx.generate_bayer(r, fb, 1000 + k, 2000 + k, a, a / 200.0, 64.0, 4095.0)
        raws.append(r)
        outs.append(ctx.image(fw, fh, ca.FMT_RGBA8))
        params.append(ca.default_params(fb, 64.0, 4095.0, levels=5, algorithm=ca.BILATERAL))
    px = sum(fw * fh for fw, fh, _, _ in frames)
    ctx.sync()
    t0 = time.perf_counter()
    burst = ca.Burst(ctx, raws, outs, params, frame_times=True)
    create_s = time.perf_counter() - t0
    try:
        lat = []
        for _ in range(3):  # untimed replays; their frame completion times
            burst.launch()
            lat.extend(burst.frame_times().tolist())
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(replays):
            burst.launch()
        ctx.sync()
        dt = time.perf_counter() - t0
        tuning = burst.tuning()
    finally:
        burst.close()
        for im in raws + outs:
            im.free()
    return {"value": round(replays * px / dt / 1e6, 1), "unit": "Mpix/s", "frames": len(frames), "pixels": px,
            "replays": replays, "ms_per_burst": round(dt / replays * 1e3, 3),
            "frame_completion_ms": {"p50": round(float(np.percentile(lat, 50)), 3),
                                    "p99": round(float(np.percentile(lat, 99)), 3), "samples": len(lat),
                                    "note": "each frame's last kernel, from the replay's start (events in the graph)"},
            "create_s": round(create_s, 3), "autotune": tuning,
            "form": "one hipGraphLaunch per burst: per-frame raw fits and pyramids on three lanes, the 16 frames' "
                    + ("green and red/blue + despeckle stages as two cross-frame persistent launches"
                       if tuning["form"] == "cross-frame launches" else
                       "green and red/blue + despeckle stages as per-frame kernels on the lanes (the faster "
                       "recorded form at creation)")}


# the path's algorithmic bytes per pixel with every float texel stored as FP16
# (the reference's ARM build), five levels
PATH_BYTES_FP16 = 199.1


def mixed_leg(ctx, raws, outs, fparams, steps, mpix_fp32):
    """The mixed-precision form (CLIM_PRECISION_MIXED) on the bench frames:
    Mpix/s batched, its first frame against the FP64 and the FP16-storage
    FP32 restatements, and config 2's geometry (8064x6048 RGGB, 14-bit, one
    frame) at both precisions."""
    from oracle import oracle as O
    ctx.set_precision(ca.PRECISION_MIXED)
    try:
        for _ in range(2):
            ctx.raw_to_rgb_batch(raws, outs, fparams)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ctx.raw_to_rgb_batch(raws, outs, fparams)
        ctx.sync()
        dt = time.perf_counter() - t0
        px = sum(r.width * r.height for r in raws)
        mpix = steps * px / dt / 1e6
        # the first frame's error (outside the timing)
        w, h = raws[0].width, raws[0].height
        of = ctx.image(w, h, ca.FMT_RGBA32F)
        raw_nlf, pyr_nlf = ctx.raw_to_rgb(raws[0], outs[0], fparams[0], out_f32=of)
        got = of.download()
        of.free()
        host = raws[0].download()
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
        p = O.default_params(ca.RGGB, 64.0, 4095.0, levels=5)
        r64 = O.Oracle(double=True).raw_to_rgb(host, p)
        p.fp16_storage = 1
        r16 = O.Oracle(double=False).raw_to_rgb(host, p)

        def err(ref):
            d = got[..., :3].astype(np.float64) - ref["rgbf"][..., :3]
            mse = float(np.mean(d * d))
            return {"max_abs": round(float(np.abs(d).max()), 5), "psnr_db": round(10 * np.log10(1.0 / max(mse, 1e-30)), 2),
                    "raw_nlf_max_rel": float(np.max(np.abs(raw_nlf - ref["rawNlf"]) / np.abs(ref["rawNlf"]))),
                    "pyramid_nlf_max_rel": float(np.max(np.abs(pyr_nlf - ref["pyramidNlf"]) /
                                                        np.abs(ref["pyramidNlf"])))}
        res = {"value": round(mpix, 1), "unit": "Mpix/s", "steps": steps, "vs_fp32": round(mpix / mpix_fp32, 4),
               "storage": "FP16 level images (despeckled level 0, downsampled and denoised levels)",
               "arithmetic": "bilateral range test and weights in packed FP16, sums in FP32 (v_dot2_f32_f16); "
                             "FP64 fits and colour matrices",
               "error_first_frame": {"vs_fp64": err(r64), "vs_fp16_storage_fp32": err(r16)},
               "path_roofline_fp16_accounting": {"bytes_per_px": PATH_BYTES_FP16,
                                                 "achieved": round(PATH_BYTES_FP16 * mpix * 1e6 / 1e9, 1),
                                                 "frac": round(PATH_BYTES_FP16 * mpix * 1e6 / 1e9 / HBM_PEAK_GBS, 4)}}
    finally:
        ctx.set_precision(ca.PRECISION_FP32)
    # config 2: one 48 MP 14-bit frame per call, both precisions
    w2, h2 = 8064, 6048
    r2 = ctx.image(w2, h2, ca.FMT_R16U)
    ctx.generate_bayer(r2, ca.RGGB, 1000, 2000, 0.002, 1e-5, 64.0, 16383.0)
    o2 = ctx.image(w2, h2, ca.FMT_RGBA8)
    p2 = ca.default_params(ca.RGGB, 64.0, 16383.0, levels=5)
    c2 = {"workload": "8064x6048 RGGB uint16 14-bit (white 16383), one frame per call"}
    for name, prec in (("fp32", ca.PRECISION_FP32), ("mixed", ca.PRECISION_MIXED)):
        ctx.set_precision(prec)
        for _ in range(2):
            ctx.raw_to_rgb(r2, o2, p2, writeback=False)
        ctx.sync()
        n2 = max(steps // 2, 4)
        t0 = time.perf_counter()
        for _ in range(n2):
            ctx.raw_to_rgb(r2, o2, p2, writeback=False)
        ctx.sync()
        c2[name] = round(n2 * w2 * h2 / (time.perf_counter() - t0) / 1e6, 1)
    ctx.set_precision(ca.PRECISION_FP32)
    for im in (r2, o2):
        im.free()
    res["config2_48mp"] = c2
    return res


def psnr(a, b):
    d = a[..., :3].astype(np.float64) - b[..., :3]
    return 10.0 * np.log10(1.0 / max(float(np.mean(d * d)), 1e-30))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=4032)
    ap.add_argument("--height", type=int, default=3024)
    ap.add_argument("--scenes-per-level", type=int, default=12,
                    help="frames per noise level in one step (frame i: noise level i mod 6, scene seed 1000 + i)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the FP64 / FP32 CPU comparison of one frame")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-strength leg")
    ap.add_argument("--no-burst", action="store_true", help="skip the recorded-burst leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--learned", action="store_true",
                    help="the learned denoise strength (committed weights, gated) is the headline path")
    ap.add_argument("--ltm", action="store_true",
                    help="local tone mapping on (its mask is computed in the level-0 reassembly + tone kernel)")
    ap.add_argument("--preview", action="store_true",
                    help="the half-resolution preview (RawConverter::fastDemosaicImage) instead of the full path")
    ap.add_argument("--noise-level", type=float, default=0.0,
                    help="DemosaicParameters::noiseLevel; above 0.6 the raw-domain pre-denoise runs")
    ap.add_argument("--burst", action="store_true",
                    help="the uneven burst: 16 frames, 4 each of 4032x3024, 4000x3000, 3264x2448 and 1920x1080, "
                         "RGGB and BGGR alternating, noise levels drawn from the headline set, one context")
    ap.add_argument("--graph", action="store_true",
                    help="with --burst: the burst recorded once as a HIP graph (clim_burst_create) and replayed, "
                         "one graph launch per step")
    ap.add_argument("--mixed", action="store_true",
                    help="the mixed-precision form (FP16 level storage, packed-FP16 bilateral weights) as the "
                         "timed path")
    ap.add_argument("--sequential", action="store_true",
                    help="one clim_raw_to_rgb call per frame instead of clim_raw_to_rgb_batch over the step's frames")
    ap.add_argument("--algorithm", choices=list(ALGORITHMS), default="bilateral",
                    help="the pyramid's denoiser (PyramidalDenoise::DenoiseAlgorithm)")
    ap.add_argument("--no-profile", action="store_true", help="skip the per-stage event timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a plain 256 MiB device fill and device-to-device copy (reference kernels "
                         "of known traffic for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(frames are independent: a burst shards whole frames across GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.burst and args.preview:
        print("bench.py: --burst runs the full path", file=sys.stderr)
        sys.exit(2)
    if args.graph and not args.burst:
        print("bench.py: --graph records the --burst frames", file=sys.stderr)
        sys.exit(2)

    w, h, bayer = args.width, args.height, ca.RGGB
    algorithm = ALGORITHMS[args.algorithm]
    lib = ca.lib()
    ctx = ca.Context(0)
    if args.mixed:
        ctx.set_precision(ca.PRECISION_MIXED)
    # frames: (width, height, CFA, noise a, frame id)
    if args.burst:
        sizes = [(4032, 3024), (4000, 3000), (3264, 2448), (1920, 1080)]
        draws = np.random.default_rng(3000).integers(0, len(NOISE_A), 16)
        frames = [(*sizes[k // 4], ca.RGGB if k % 2 == 0 else ca.BGGR, NOISE_A[draws[k]], k) for k in range(16)]
    else:
        nf = len(NOISE_A) * max(args.scenes_per_level, 1)
        frames = [(w, h, bayer, NOISE_A[i % len(NOISE_A)], i) for i in range(nf)]
    raws, outs, fparams = [], [], []
    for fw, fh, fb, a, i in frames:
        r = ctx.image(fw, fh, ca.FMT_R16U)
        ctx.generate_bayer(r, fb, 1000 + i, 2000 + i, a, a / 200.0, 64.0, 4095.0)
        raws.append(r)
        outs.append(ctx.image(fw // 2, fh // 2, ca.FMT_RGBA8) if args.preview else ctx.image(fw, fh, ca.FMT_RGBA8))
        fp = ca.default_params(fb, 64.0, 4095.0, levels=5, algorithm=algorithm)
        fp.rgbConversionParameters.localToneMapping = int(args.ltm)
        fp.noiseLevel = args.noise_level
        fparams.append(fp)
    px_per_step = sum(fw * fh for fw, fh, _, _, _ in frames)
    single_kernel = dict(SINGLE_KERNEL_MIXED if args.mixed else SINGLE_KERNEL)
    if args.ltm:
        single_kernel["tone"] = f"k_reassemble<true, true, true, {'true' if args.mixed else 'false'}>"
    if args.preview:
        single_kernel["tone"] = "k_fast_preview<true>"
    if algorithm != ca.BILATERAL and not args.mixed:
        # levels 1-3 reassembled on their own (the bilateral path denoises
        # them in the same launch)
        single_kernel["reassemble"] = "k_reassemble<true, false, false, false>"
    if algorithm == ca.GUIDED_FAST:
        single_kernel["denoise"] = f"k_guided_fast<{'true' if args.mixed else 'false'}>"
    elif algorithm == ca.GUIDED_PRECISE:
        del single_kernel["denoise"]  # seven kernels per level
    learned_on = not args.preview and not args.no_learned
    if learned_on or args.learned:
        ctx.strength_set_weights(ca.load_strength_weights())
    if args.learned:
        ctx.strength_set_mode(ca.STRENGTH_GATED)
    ctx.sync()
    graph = ca.Burst(ctx, raws, outs, fparams, frame_times=True) if args.graph else None

    def step(batched=not args.sequential):
        if batched and graph is not None:
            graph.launch()
            return
        if batched and not args.preview:
            ctx.raw_to_rgb_batch(raws, outs, fparams)
            return
        for r, o, fp in zip(raws, outs, fparams):
            if args.preview:
                ctx.fast_demosaic(r, o, fp)
            else:
                ctx.raw_to_rgb(r, o, fp, writeback=False)

    import ctypes as C

    class Prof(C.Structure):
        _fields_ = [("ms", C.c_double * ca.PROFILE_STAGES), ("launches", C.c_int64 * ca.PROFILE_STAGES),
                    ("texels", C.c_int64 * ca.PROFILE_STAGES)]

    lib.clim_profile_stage_name.restype = C.c_char_p
    names = [lib.clim_profile_stage_name(i).decode() for i in range(ca.PROFILE_STAGES)]

    def read_profile():
        pr = Prof()
        lib.clim_profile_read(ctx.h, C.byref(pr))
        return {names[i]: {"ms_total": pr.ms[i], "launches": int(pr.launches[i]), "texels": int(pr.texels[i])}
                for i in range(ca.PROFILE_STAGES) if pr.launches[i]}

    def profile_steps(n):
        """n untimed steps, one frame at a time on the profiled stream, every
        stage bracketed by events: per-stage totals and us per frame."""
        lib.clim_profile_select(ctx.h, C.c_uint32(0xFFFFFFFF))
        lib.clim_profile_enable(ctx.h, 1)
        read_profile()
        for _ in range(n):
            step(batched=False)
        st = read_profile()
        lib.clim_profile_enable(ctx.h, 0)
        for v in st.values():
            v["us_per_frame"] = round(v["ms_total"] * 1e3 / (n * len(raws)), 2)
        return st

    def timed(nsteps):
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            step()
        ctx.sync()
        return time.perf_counter() - t0

    for _ in range(args.warmup):
        step()
    ctx.sync()

    # 1. stage breakdown (every stage evented, one frame at a time)
    stages, dom = {}, None
    prof_on = not args.no_profile
    if prof_on:
        stages = profile_steps(min(args.steps, 3))
        dom = max((k for k in stages if k in single_kernel), key=lambda k: stages[k]["ms_total"])
        # the timed steps record events around the dominant kernel only
        lib.clim_profile_select(ctx.h, C.c_uint32(1 << names.index(dom)))
        lib.clim_profile_enable(ctx.h, 1)
        read_profile()

    # 2. the timed steps
    dt = timed(args.steps)
    mpix = args.steps * px_per_step / dt / 1e6
    batched = not args.sequential and not args.preview

    if args.copy_calibration:
        copy_calibration(256 << 20)

    latency = None
    if graph is not None:
        # each frame's completion within a replay (events in the graph)
        lat = []
        for _ in range(5):
            graph.launch()
            lat.extend(graph.frame_times().tolist())
        latency = {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                   "frames": len(lat), "note": "frame completion from the replay's start",
                   "autotune": graph.tuning()}
    elif args.burst:
        # per-frame latency: each frame launched on an idle device and waited for
        lat = []
        for _ in range(3):
            for r, o, fp in zip(raws, outs, fparams):
                t1 = time.perf_counter()
                ctx.raw_to_rgb(r, o, fp, writeback=False)
                ctx.sync()
                lat.append((time.perf_counter() - t1) * 1e3)
        latency = {"p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                   "frames": len(lat)}
    roofline = None
    if prof_on:
        # (a recorded burst's replays carry no stage events)
        s = read_profile().get(dom)
        lib.clim_profile_enable(ctx.h, 0)
        # primary: the kernel alone (the profiled steps, one frame at a time)
        iso = stages[dom]
        iso_s = iso["ms_total"] / iso["launches"] / 1e3
        bytes_per_launch = stage_bytes(dom, iso["texels"] / iso["launches"], preview=args.preview, half=args.mixed)
        ach = bytes_per_launch / iso_s / 1e9
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4),
                    "traffic": None if (args.burst or args.mixed) else measured_traffic(single_kernel[dom], w, h),
                    "kernel": single_kernel[dom], "stage": dom, "avg_launch_us": round(iso_s * 1e6, 2),
                    "launches_timed": iso["launches"], "algorithmic_bytes_per_launch": int(bytes_per_launch),
                    "timing": "kernel alone (one frame at a time), HIP events on the library's stream"}
        if batched and s:
            # the same kernel's launches in the timed batched steps, sharing the
            # device with the other lanes' frames
            b_s = s["ms_total"] / s["launches"] / 1e3
            b_ach = stage_bytes(dom, s["texels"] / s["launches"], half=args.mixed) / b_s / 1e9
            roofline["batched"] = {"avg_launch_us": round(b_s * 1e6, 2), "achieved": round(b_ach, 1),
                                   "frac": round(b_ach / HBM_PEAK_GBS, 4), "lanes": 3,
                                   "launches_timed": s["launches"]}

    path = None
    if not args.preview:
        # the FP16-storage accounting for the mixed form (every float texel halved)
        bpp = PATH_BYTES_FP16 if args.mixed else path_bytes_per_px(5)
        path = {"algorithmic": {"bytes_per_px": bpp, "achieved": round(bpp * mpix * 1e6 / 1e9, 1),
                                "frac": round(bpp * mpix * 1e6 / 1e9 / HBM_PEAK_GBS, 4)},
                "peak": HBM_PEAK_GBS, "unit": "GB/s"}
        default_workload = not (args.burst or args.learned or args.ltm or args.noise_level > 0.6 or args.mixed
                                or algorithm != ca.BILATERAL)
        fb = measured_frame_bytes(w, h) if default_workload else None
        if fb:
            mbpp = fb / (w * h)
            path["measured"] = {"bytes_per_px": round(mbpp, 2), "achieved": round(mbpp * mpix * 1e6 / 1e9, 1),
                                "frac": round(mbpp * mpix * 1e6 / 1e9 / HBM_PEAK_GBS, 4),
                                "source": "profiles/traffic_raw_to_rgb_12mp.json (FETCH_SIZE + WRITE_SIZE)"}

    # 3. the learned denoise strength on the same frames
    learned = None
    # (a recorded burst keeps the strength mode it was recorded with)
    if learned_on and not args.learned and graph is None:
        ctx.strength_set_mode(ca.STRENGTH_GATED)
        for _ in range(2):
            step()
        lst = profile_steps(1) if prof_on else {}
        lsteps = max(args.steps // 2, 1)
        ldt = timed(lsteps)
        lmpix = lsteps * px_per_step / ldt / 1e6
        opened, pred = ctx.strength_gate_read()
        learned = {"value": round(lmpix, 1), "unit": "Mpix/s", "steps": lsteps,
                   "ms_per_step": round(ldt / lsteps * 1e3, 3), "vs_plain": round(lmpix / mpix, 4),
                   "weights": "climage_amd/weights/strength.f32 (gated)",
                   "gate_last_frame": opened.tolist(), "predicted_log2_error_ratio": np.round(pred, 4).tolist()}
        if "strength" in lst:
            st_us = lst["strength"]["us_per_frame"]
            texels = lst["strength"]["texels"] / lst["strength"]["launches"]
            flops = STRENGTH_FLOP_PER_TEXEL * texels
            learned["strength_stage"] = {"us_per_frame": st_us, "grid_texels": int(texels),
                                         "flop_per_frame": int(flops),
                                         "achieved_tflops": round(flops / (st_us * 1e-6) / 1e12, 2),
                                         "mfma_peak_tflops": BF16_PEAK_TFLOPS,
                                         "frac": round(flops / (st_us * 1e-6) / 1e12 / BF16_PEAK_TFLOPS, 4),
                                         "note": "the fused network launch (features and four layers) and the gate, events around both"}
            learned["denoise_us_per_frame"] = lst.get("denoise", {}).get("us_per_frame")
        if not args.burst:
            # PSNR against the clean truth (the clean scene through the pipeline's
            # tone map) on the first frame, plain and learned
            fw, fh, fb_, a, i = frames[2]
            clean = ctx.image(fw, fh, ca.FMT_RGBA32F)
            raw = ctx.image(fw, fh, ca.FMT_R16U)
            ctx.generate_bayer(raw, fb_, 1000 + i, 2000 + i, a, a / 200.0, 64.0, 4095.0, clean=clean)
            truth = ctx.image(fw, fh, ca.FMT_RGBA32F)
            of = ctx.image(fw, fh, ca.FMT_RGBA32F)
            o8 = ctx.image(fw, fh, ca.FMT_RGBA8)
            ctx.convertTosRGB(clean, o8, np.eye(3, dtype=np.float32), fparams[2].rgbConversionParameters,
                              out_f32=truth)
            t = truth.download()
            res = {}
            for name, mode in (("plain", ca.STRENGTH_OFF), ("learned", ca.STRENGTH_GATED)):
                ctx.strength_set_mode(mode)
                ctx.raw_to_rgb(raw, o8, fparams[2], out_f32=of, writeback=False)
                res[name] = round(psnr(of.download(), t), 3)
            learned["psnr_vs_clean_truth_db"] = {"frame": f"bench frame {i} (a = {a})", **res,
                                                 "gain": round(res["learned"] - res["plain"], 3)}
            for im in (clean, raw, truth, of, o8):
                im.free()
        ctx.strength_set_mode(ca.STRENGTH_OFF)

    # 4. the uneven burst recorded as one graph (its own frames and timing)
    burst = None
    if not (args.no_burst or args.preview or args.burst or args.learned or args.mixed):
        burst = burst_leg(ctx)

    # 5. the mixed-precision form on the same frames, and config 2's geometry
    mixed = None
    if not (args.no_mixed or args.mixed or args.preview or args.burst or args.learned or args.ltm
            or args.noise_level > 0.6 or algorithm != ca.BILATERAL):
        mixed = mixed_leg(ctx, raws, outs, fparams, max(args.steps // 4, 10), mpix)

    # 6. the first frame against the CPU restatements (outside any timing)
    parity = None
    if not (args.no_parity or args.preview or args.burst or args.learned or args.mixed or args.ltm
            or args.noise_level > 0.6):
        of = ctx.image(w, h, ca.FMT_RGBA32F)
        nlf = ctx.raw_to_rgb(raws[0], outs[0], fparams[0], out_f32=of)
        parity = fp64_parity(raws[0].download(), bayer, of.download(), nlf, algorithm=algorithm)
        of.free()

    cpu = None
    if not args.no_cpu_baseline and not args.preview:
        # the burst's CPU sample: its smallest frame size
        cw, ch = (1920, 1080) if args.burst else (w, h)
        cpu = cpu_baseline(cw, ch, bayer, ltm=args.ltm, noise_level=args.noise_level, algorithm=algorithm)

    if args.burst:
        metric = "Mpix/s RAW->RGB (demosaic+pyramid denoise), burst of 16 uneven frames"
        workload = ("burst: 4 each of 4032x3024, 4000x3000, 3264x2448, 1920x1080 RGGB/BGGR uint16 12-bit, "
                    "per-frame noise level drawn from the headline set, 5-level bilateral pyramid, one context"
                    + (", recorded as one HIP graph, one graph launch per step" if graph is not None else ""))
    else:
        metric = ("Mpix/s RAW->RGB half-resolution preview (mosaic pixels) at 12 MP" if args.preview
                  else "Mpix/s RAW->RGB (demosaic+learned-strength pyramid denoise) at 12 MP" if args.learned
                  else "Mpix/s RAW\u2192RGB (demosaic+pyramid denoise) at 12 MP; PSNR vs FP64 oracle")
        workload = (f"{w}x{h} RGGB uint16 12-bit, {len(raws)} frames per step (the 6 noise levels, "
                    f"{len(raws) // len(NOISE_A)} seeded scenes each), "
                    + ("half-resolution preview (fastDemosaicImage)" if args.preview
                       else f"5-level {ALGORITHM_NAMES[algorithm]} pyramid")
                    + (", learned strength (gated)" if args.learned else "")
                    + (", mixed precision (FP16 level storage, packed-FP16 bilateral weights)" if args.mixed else "")
                    + (", local tone mapping" if args.ltm else "")
                    + (f", noiseLevel {args.noise_level} (raw pre-denoise)" if args.noise_level > 0.6 else ""))
    res = {
        "metric": metric,
        "value": round(mpix, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f16 storage + f32" if args.mixed else "f32",
        "data": "synthetic (seeded scenes generated on the device)",
        "config": {"workload": workload, "frames_per_step": len(raws), "pixels_per_step": px_per_step,
                   "timed_seconds": round(dt, 3), "batched": batched},
        "roofline": roofline,
        "path_roofline": path,
        "fp64": parity,
        "learned": learned,
        "burst": burst,
        "mixed": mixed,
        "cpu_baseline": cpu,
        "stages": stages,
    }
    if latency:
        res["latency"] = latency
    print(json.dumps(res))
    if graph is not None:
        graph.close()
    ctx.close()


if __name__ == "__main__":
    main()
