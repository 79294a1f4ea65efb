"""The mixed-precision form (CLIM_PRECISION_MIXED): FP16 level images, the
bilateral's range test and weights in packed FP16 with FP32 sums, FP64 fits
and colour matrices.  Its results are not bit-exact with any restatement:
each frame states its error against the FP64 oracle for every output (RGB
max-abs and PSNR, rawNlf and pyramidNlf relative error) and against the
FP16-storage oracle (FP16 level images, FP32 arithmetic: the difference left
is the FP16 arithmetic of the bilateral), with bounds written here.  The
batch and burst paths at this precision equal the single-frame calls bit for
bit."""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def errors(name, of, raw_nlf, pyr_nlf, ref, flush=True):
    d = of[..., :3].astype(np.float64) - ref["rgbf"][..., :3]
    ok = np.isfinite(d)
    d = d[ok]
    mse = float(np.mean(d * d))
    e = {"max_abs": float(np.abs(d).max()), "psnr": 10 * np.log10(1.0 / mse) if mse > 0 else np.inf,
         "over_1e-2": float(np.mean(np.abs(d) > 1e-2)), "nonfinite": int((~ok).sum()),
         "rawNlf_rel": float(np.max(np.abs(raw_nlf - ref["rawNlf"]) / np.abs(ref["rawNlf"]))),
         "pyramidNlf_rel": float(np.max(np.abs(pyr_nlf - ref["pyramidNlf"]) / np.abs(ref["pyramidNlf"])))}
    print(f"{name}: RGB max-abs {e['max_abs']:.3g} PSNR {e['psnr']:.1f} dB (>1e-2: {100 * e['over_1e-2']:.3f} %, "
          f"non-finite {e['nonfinite']}), rawNlf rel {e['rawNlf_rel']:.2g}, pyramidNlf rel {e['pyramidNlf_rel']:.2g}",
          flush=flush)
    return e


def run_mixed(ctx, raw, p):
    h, w = raw.shape
    r = ctx.upload(raw, ca.FMT_R16U)
    o8, of = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
    raw_nlf, pyr_nlf = ctx.raw_to_rgb(r, o8, p, out_f32=of)
    out = (o8.download(), of.download(), raw_nlf, pyr_nlf)
    for im in (r, o8, of):
        im.free()
    return out


@pytest.fixture(scope="module")
def mctx():
    c = ca.Context(0)
    c.set_precision(ca.PRECISION_MIXED)
    yield c
    c.close()


@pytest.mark.parametrize("bayer,algorithm,a,gb", [
    (ca.RGGB, ca.BILATERAL, 0.004, 1.0), (ca.BGGR, ca.BILATERAL, 0.0005, 1.0), (ca.GRBG, ca.BILATERAL, 0.016, 2.0),
    (ca.GBRG, ca.GUIDED_FAST, 0.004, 1.0), (ca.RGGB, ca.GUIDED_PRECISE, 0.002, 1.0)])
def test_mixed_against_fp64_and_fp16_storage(mctx, oracle32, oracle64, bayer, algorithm, a, gb):
    w, h = 512, 384
    raw = O.generate_bayer(w, h, bayer, 70 + bayer, 80 + bayer, a, a / 200.0, 64.0, 4095.0)
    p = ca.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    op = O.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    for lv in range(5):
        p.denoiseParameters[lv].gradientBoost = gb
        op.dp[lv].gradientBoost = gb
    o8, of, raw_nlf, pyr_nlf = run_mixed(mctx, raw, p)
    name = f"mixed {w}x{h} bayer {bayer} alg {algorithm} a={a} gB={gb}"
    e64 = errors(name + " vs FP64", of, raw_nlf, pyr_nlf, oracle64.raw_to_rgb(raw, op))
    op.fp16_storage = 1
    e16 = errors(name + " vs FP16-storage FP32", of, raw_nlf, pyr_nlf, oracle32.raw_to_rgb(raw, op))
    # the raw fit precedes every FP16 store: exactly the FP32 one
    assert e16["rawNlf_rel"] == 0.0
    assert e64["nonfinite"] == 0
    if algorithm == ca.GUIDED_PRECISE:
        # the FP32 form itself parts from FP64 here (its 3x3 colour covariance inverse)
        assert e64["psnr"] > 25.0
    else:
        assert e64["psnr"] > 50.0 and e16["psnr"] > 50.0, (e64, e16)
        # FP16 storage quantises the levels: the smallest fitted terms move most
        assert e64["pyramidNlf_rel"] < 0.1


def test_config2_mixed_48mp_14bit(mctx, oracle64):
    """Config 2: 8064x6048 RGGB, 14-bit (white 16383), the mixed-precision
    form, error against the FP64 oracle for each output."""
    w, h, white = 8064, 6048, 16383.0
    raw = O.generate_bayer(w, h, ca.RGGB, 1000, 2000, 0.002, 1e-5, 64.0, white)
    p = ca.default_params(ca.RGGB, 64.0, white, levels=5)
    print("config 2 mixed: 8064x6048", flush=True)
    o8, of, raw_nlf, pyr_nlf = run_mixed(mctx, raw, p)
    e = errors("config 2 mixed vs FP64", of, raw_nlf, pyr_nlf,
               oracle64.raw_to_rgb(raw, O.default_params(ca.RGGB, 64.0, white, levels=5)))
    assert e["nonfinite"] == 0 and e["psnr"] > 50.0 and e["rawNlf_rel"] < 1e-4 and e["pyramidNlf_rel"] < 0.1


def test_mixed_batch_and_burst_equal_single_frames(mctx):
    frames = [(640, 480, ca.RGGB, 0.004), (512, 384, ca.BGGR, 0.016), (320, 256, ca.GRBG, 0.0005)]
    raws, outs, outsb, params, singles = [], [], [], [], []
    for k, (w, h, bayer, a) in enumerate(frames):
        raw = O.generate_bayer(w, h, bayer, 5 + k, 50 + k, a, a / 200.0, 64.0, 4095.0)
        raws.append(mctx.upload(raw, ca.FMT_R16U))
        outs.append(mctx.image(w, h, ca.FMT_RGBA8))
        outsb.append(mctx.image(w, h, ca.FMT_RGBA8))
        params.append(ca.default_params(bayer, 64.0, 4095.0, levels=5))
        o = mctx.image(w, h, ca.FMT_RGBA8)
        mctx.raw_to_rgb(raws[-1], o, params[-1], writeback=False)
        singles.append(o.download())
    mctx.raw_to_rgb_batch(raws, outs, params)
    with ca.Burst(mctx, raws, outsb, params) as b:
        b.launch()
        b.launch()
        mctx.sync()
        for k in range(len(frames)):
            assert np.array_equal(outs[k].download(), singles[k]), f"batch frame {k}"
            assert np.array_equal(outsb[k].download(), singles[k]), f"burst frame {k}"
    # and the mixed form differs from the FP32 one
    with ca.Context(0) as c32:
        r = c32.upload(raws[0].download(), ca.FMT_R16U)
        o = c32.image(frames[0][0], frames[0][1], ca.FMT_RGBA8)
        c32.raw_to_rgb(r, o, params[0], writeback=False)
        assert not np.array_equal(o.download(), singles[0])
