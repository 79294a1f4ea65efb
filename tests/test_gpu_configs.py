"""The measurement configurations beyond the headline one, as parity cases.

* Config 2's geometry: 8064x6048 RGGB, 14-bit (white 16383), 5-level
  bilateral pyramid, through the FP32 path.
* Config 3's uneven burst: frames of 1920x1080, 4032x3024, 3264x2448 and
  4000x3000 with RGGB and BGGR mixed, each at a noise level drawn from the
  headline set, through one context, whose work buffers grow once and are
  then reused as views of smaller size (RawConverter keeps its textures
  across frames, cpp/raw_converter.cpp:27).  4000x3000 and 1920x1080 reach
  the general (non-2x) sampler path at their deepest levels (375 -> 187,
  135 -> 67).

Every frame equals the FP32 oracle bit for bit (float and RGBA8 outputs and
the fitted noise models); config 2 also reports its error against the FP64
oracle (PSNR >= 60 dB).
"""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NOISE_A = [0.0005, 0.001, 0.002, 0.004, 0.008, 0.016]


def bitexact(gpu, ref, name):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        first = tuple(idx[0])
        pytest.fail(f"{name}: {len(idx)} of {same.size} values differ, first at {first}: "
                    f"gpu {gpu[first]!r} oracle {ref[first]!r}")


def run(ctx, raw, p):
    h, w = raw.shape
    r = ctx.upload(raw, ca.FMT_R16U)
    o8, of = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
    raw_nlf, pyr_nlf = ctx.raw_to_rgb(r, o8, p, out_f32=of)
    return o8.download(), of.download(), raw_nlf, pyr_nlf


def check_frame(ctx, oracle32, raw, bayer, white, name):
    print(f"{name}: {raw.shape[1]}x{raw.shape[0]} bayer {bayer}", flush=True)
    p = ca.default_params(bayer, 64.0, white, levels=5)
    o8, of, raw_nlf, pyr_nlf = run(ctx, raw, p)
    ref = oracle32.raw_to_rgb(raw, O.default_params(bayer, 64.0, white, levels=5))
    bitexact(raw_nlf, ref["rawNlf"], f"{name} rawNlf")
    bitexact(pyr_nlf, ref["pyramidNlf"], f"{name} pyramidNlf")
    bitexact(of, ref["rgbf"], f"{name} float")
    bitexact(o8, ref["rgba8"], f"{name} rgba8")
    return of


def test_config3_uneven_burst_through_one_context(oracle32):
    ctx = ca.Context(0)
    try:
        frames = [(1920, 1080, ca.RGGB), (4032, 3024, ca.BGGR), (3264, 2448, ca.RGGB), (4000, 3000, ca.BGGR),
                  (1920, 1080, ca.BGGR)]
        draws = np.random.default_rng(3000).integers(0, len(NOISE_A), len(frames))
        for k, ((w, h, bayer), d) in enumerate(zip(frames, draws)):
            a = NOISE_A[d]
            raw = O.generate_bayer(w, h, bayer, 1000 + k, 2000 + k, a, a / 200.0, 64.0, 4095.0)
            check_frame(ctx, oracle32, raw, bayer, 4095.0, f"burst frame {k} (a={a})")
    finally:
        ctx.close()


def test_config2_geometry_48mp_14bit(ctx, oracle32, oracle64):
    w, h, white = 8064, 6048, 16383.0
    raw = O.generate_bayer(w, h, ca.RGGB, 1000, 2000, 0.002, 1e-5, 64.0, white)
    of = check_frame(ctx, oracle32, raw, ca.RGGB, white, "config 2")
    r64 = oracle64.raw_to_rgb(raw, O.default_params(ca.RGGB, 64.0, white, levels=5))["rgbf"][..., :3]
    d = of[..., :3].astype(np.float64) - r64
    psnr = 10 * np.log10(1.0 / np.mean(d ** 2))
    print(f"config 2: vs FP64 max {np.abs(d).max():.3g} psnr {psnr:.1f} dB", flush=True)
    assert psnr >= 60.0
