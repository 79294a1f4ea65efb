"""GPU kernels vs the FP32 oracle, stage by stage and end to end.

Every stage gets the oracle's output of the previous stage as its input, so
a mismatch is attributed to one kernel.  The kernels perform the same FP32
operations as the oracle (both built with -ffp-contract=off, shared scalar
primitives from include/clim_math.h), so the expected difference is zero:
the tests assert bit equality.  The end-to-end tests additionally state the
tolerance of the task (max-abs <= 1e-3 on the normalised FP32 output) and
report the error against the FP64 oracle.
"""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

W, H = 256, 192
BAYERS = [ca.GRBG, ca.GBRG, ca.RGGB, ca.BGGR]
A_NOISE, B_NOISE = 0.002, 1e-5


def bitexact(gpu, ref, name):
    gpu = np.asarray(gpu)
    ref = np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        d = np.abs(gpu.astype(np.float64) - ref.astype(np.float64))
        d = np.where(np.isfinite(d), d, 0)
        first = tuple(idx[0])
        pytest.fail(f"{name}: {len(idx)} of {same.size} values differ, max |d| = {d.max():.3g}, "
                    f"first at {first}: gpu {gpu[first]!r} oracle {ref[first]!r}")


def up(ctx, arr, fmt):
    return ctx.upload(arr, fmt)


def scene(bayer, w=W, h=H, seed=0):
    return O.generate_bayer(w, h, bayer, 1000 + seed, 2000 + seed, A_NOISE, B_NOISE, 64.0, 4095.0, clean=True)


@pytest.fixture(scope="module")
def chain():
    """Oracle outputs of every stage for each CFA pattern."""
    o = O.Oracle(False)
    out = {}
    for bayer in BAYERS:
        raw, clean = scene(bayer)
        sm = [65535.0 / (4095.0 - 64.0)] * 4
        black = np.float32(64.0) / np.float32(65535.0)
        scaled = o.scale_raw(raw, bayer, sm, black)
        mean, var = o.raw_noise_statistics(scaled, bayer)
        rc, fit = o.nlf_fit_raw(mean, var)
        green = o.interpolate_green(scaled, bayer, fit[5])
        chroma = (fit[4] + fit[6]) / np.float32(2.0)
        rgb = o.interpolate_red_blue(scaled, green, bayer, chroma)
        c2y = o.cam_ycbcr(np.eye(3))
        ycc = o.transform(o.blend_highlights(rgb), c2y)
        desp = o.despeckle(ycc, O.DEFAULT_PYRAMID_NLF[0][:3], O.DEFAULT_PYRAMID_NLF[0][3:])
        out[bayer] = dict(raw=raw, clean=clean, sm=sm, black=black, scaled=scaled, mean=mean, var=var, fit=fit,
                          green=green, chroma=chroma, rgb=rgb, c2y=c2y, ycc=ycc, desp=desp)
    return out


@pytest.mark.parametrize("bayer", BAYERS)
def test_scene_generator(ctx, bayer):
    raw_ref, clean_ref = scene(bayer)
    raw = ctx.image(W, H, ca.FMT_R16U)
    clean = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.generate_bayer(raw, bayer, 1000, 2000, A_NOISE, B_NOISE, 64.0, 4095.0, clean=clean)
    bitexact(raw.download(), raw_ref, "raw mosaic")
    bitexact(clean.download(), clean_ref, "clean scene")


@pytest.mark.parametrize("bayer", BAYERS)
def test_scale_and_raw_stats(ctx, chain, bayer):
    c = chain[bayer]
    raw = up(ctx, c["raw"], ca.FMT_R16U)
    scaled = ctx.image(W, H, ca.FMT_R32F)
    ctx.scaleRawData(raw, scaled, bayer, c["sm"], float(c["black"]))
    bitexact(scaled.download(), c["scaled"], "scaleRawData")
    s_in = up(ctx, c["scaled"], ca.FMT_R32F)
    mean = ctx.image(W // 2, H // 2, ca.FMT_RGBA32F)
    var = ctx.image(W // 2, H // 2, ca.FMT_RGBA32F)
    ctx.rawNoiseStatistics(s_in, bayer, mean, var)
    bitexact(mean.download(), c["mean"], "rawNoiseStatistics mean")
    bitexact(var.download(), c["var"], "rawNoiseStatistics var")
    fit = ctx.computeRawNoiseStatistics(s_in, bayer)
    bitexact(fit, c["fit"], "computeRawNoiseStatistics")


@pytest.mark.parametrize("w,h", [(202, 142), (66, 34), (32, 32)])
def test_fits_on_partial_tiles(ctx, oracle32, w, h):
    """computeRawNoiseStatistics and computeNoiseStatistics at sizes whose
    statistics tiles (16x16 quads, 32x16 texels) are cut by the image edge:
    the samples the statistics kernels list and the refit over those lists
    give the oracle's fits (a 32x32 mosaic leaves a single raw tile)."""
    rng = np.random.default_rng(w + h)
    yy, xx = np.mgrid[0:h, 0:w]
    field = 64.0 + 4031.0 * (0.2 + 0.1 * np.sin(xx / 11.0) * np.cos(yy / 7.0))
    raw = np.clip(np.rint(field + rng.normal(0.0, 20.0, (h, w))), 0, 4095).astype(np.uint16)
    sm = [65535.0 / (4095.0 - 64.0)] * 4
    black = np.float32(64.0) / np.float32(65535.0)
    scaled = oracle32.scale_raw(raw, ca.GBRG, sm, black)
    mean, var = oracle32.raw_noise_statistics(scaled, ca.GBRG)
    rc, fit = oracle32.nlf_fit_raw(mean, var)
    assert rc in (0, 2)
    bitexact(ctx.computeRawNoiseStatistics(up(ctx, scaled, ca.FMT_R32F), ca.GBRG), fit, f"raw fit {w}x{h}")
    img = np.stack([0.2 + 0.1 * np.sin(xx / 9.0 + k) for k in range(4)], -1)
    img = (img + rng.normal(0.0, 0.003, (h, w, 4))).astype(np.float32)
    rc, lfit = oracle32.nlf_fit_level(oracle32.noise_statistics(img))
    assert rc in (0, 2)
    bitexact(ctx.computeNoiseStatistics(up(ctx, img, ca.FMT_RGBA32F)), lfit, f"level fit {w}x{h}")


@pytest.mark.parametrize("bayer", BAYERS)
def test_interpolate_green(ctx, chain, bayer):
    c = chain[bayer]
    s_in = up(ctx, c["scaled"], ca.FMT_R32F)
    green = ctx.image(W, H, ca.FMT_R32F)
    ctx.interpolateGreen(s_in, green, bayer, float(c["fit"][5]))
    bitexact(green.download(), c["green"], "interpolateGreen")


@pytest.mark.parametrize("bayer", BAYERS)
@pytest.mark.parametrize("rotate", [False, True])
def test_interpolate_red_blue(ctx, chain, oracle32, bayer, rotate):
    c = chain[bayer]
    s_in = up(ctx, c["scaled"], ca.FMT_R32F)
    g_in = up(ctx, c["green"], ca.FMT_R32F)
    rgb = ctx.image(W, H, ca.FMT_RGBA32F)
    if rotate:
        rgb.upload(np.zeros((H, W, 4), np.float32))
    ctx.interpolateRedBlue(s_in, g_in, rgb, bayer, float(c["chroma"]), rotate)
    ref = oracle32.interpolate_red_blue(c["scaled"], c["green"], bayer, c["chroma"], rotate) if rotate else c["rgb"]
    bitexact(rgb.download(), ref, "interpolateRedBlue")


def test_highlights_transform_despeckle(ctx, chain, oracle32):
    c = chain[ca.RGGB]
    rng = np.random.default_rng(5)
    hot = c["rgb"].copy()
    hot[..., :3] *= rng.uniform(0.5, 2.5, size=hot[..., :3].shape).astype(np.float32)
    src = up(ctx, hot, ca.FMT_RGBA32F)
    out = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.blendHighlightsImage(src, 1.0, out)
    ref_hl = oracle32.blend_highlights(hot, 1.0)
    bitexact(out.download(), ref_hl, "blendHighlightsImage")
    ctx.transformImage(out, out, c["c2y"])
    bitexact(out.download(), oracle32.transform(ref_hl, c["c2y"]), "transformImage")
    ycc = up(ctx, c["ycc"], ca.FMT_RGBA32F)
    d = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.despeckleImage(ycc, O.DEFAULT_PYRAMID_NLF[0][:3], O.DEFAULT_PYRAMID_NLF[0][3:], d)
    bitexact(d.download(), c["desp"], "despeckleImage")


@pytest.mark.parametrize("size", [(256, 192), (375, 187), (135, 67)])
def test_downsample(ctx, chain, oracle32, size):
    src = chain[ca.RGGB]["desp"]
    iw, ih = size
    src = np.ascontiguousarray(src[:ih, :iw])
    ow, oh = iw // 2, ih // 2
    a = up(ctx, src, ca.FMT_RGBA32F)
    b = ctx.image(ow, oh, ca.FMT_RGBA32F)
    ctx.downsampleImage(a, b)
    bitexact(b.download(), oracle32.downsample(src, ow, oh), f"downsampleImage {size}")


@pytest.mark.parametrize("w,h,levels", [(320, 200, 5), (200, 136, 4), (203, 141, 4)])
def test_pyramid_levels_of_random_rgba(ctx, oracle32, w, h, levels):
    """The image pyramid pyramidalDenoiseInto builds from a random RGBA image
    (all four channels vary): level 1 is written by level 0's statistics
    kernel when the reduction is exactly 2x (320x200, 200x136 with partial
    tiles) and by the downsample kernel otherwise (203x141); every level, the
    per-level noise model and the denoised level 0 equal the oracle's."""
    rng = np.random.default_rng(w * h)
    yy, xx = np.mgrid[0:h, 0:w]
    smooth = np.stack([0.2 + 0.15 * np.sin(xx / 17.0 + k) * np.cos(yy / 23.0 - k) for k in range(4)], -1)
    img = (smooth + rng.normal(0.0, 0.004, (h, w, 4))).astype(np.float32)
    dp = [O.DenoiseParams(1.0, 1.0, 1.0, 1.0, 1.0)] * 5
    rc, den0, nlf = oracle32.pyramidal_denoise(img, levels, 0, dp, O.DEFAULT_PYRAMID_NLF)
    lv = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(1, levels)]
    dn = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(levels)]
    cp = [ca.DenoiseParameters(1.0, 1.0, 1.0, 1.0, 1.0)] * 5
    got_nlf = ctx.pyramidalDenoiseInto(up(ctx, img, ca.FMT_RGBA32F), levels, ca.BILATERAL, cp, lv, dn,
                                       nlf=np.asarray(O.DEFAULT_PYRAMID_NLF[:levels]))
    prev = img
    for i in range(1, levels):
        prev = oracle32.downsample(prev, w >> i, h >> i)
        bitexact(lv[i - 1].download(), prev, f"level {i} of {w}x{h}")
    bitexact(got_nlf, nlf[:levels], "per-level noise model")
    bitexact(dn[0].download(), den0, "denoised level 0")


def test_noise_statistics_and_fit(ctx, chain, oracle32):
    src = chain[ca.RGGB]["desp"]
    a = up(ctx, src, ca.FMT_RGBA32F)
    st = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.noiseStatistics(a, st)
    ref = oracle32.noise_statistics(src)
    bitexact(st.download(), ref, "noiseStatistics")
    rc, fit = oracle32.nlf_fit_level(ref)
    assert rc in (0, 2)
    bitexact(ctx.computeNoiseStatistics(a), fit, "computeNoiseStatistics")


@pytest.mark.parametrize("cb,gb", [(1.0, 1.0), (1.5, 2.0)])
def test_denoise_bilateral(ctx, chain, oracle32, cb, gb):
    src = chain[ca.RGGB]["desp"]
    va, vb = [1e-5, 2e-5, 3e-5], [8e-4, 1e-4, 2e-4]
    a = up(ctx, src, ca.FMT_RGBA32F)
    b = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.denoiseImage(a, va, vb, cb, gb, b)
    bitexact(b.download(), oracle32.denoise_bilateral(src, va, vb, cb, gb), "denoiseImage")


def test_denoise_guided(ctx, chain, oracle32):
    src = chain[ca.RGGB]["desp"]
    va, vb = [1e-5, 2e-5, 3e-5], [8e-4, 1e-4, 2e-4]
    a = up(ctx, src, ca.FMT_RGBA32F)
    b = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.denoiseImageGuided(a, va, vb, b)
    bitexact(b.download(), oracle32.denoise_guided_fast(src, va, vb), "denoiseImageGuided")
    ctx.guidedFilter(a, 5, vb, b)
    bitexact(b.download(), oracle32.guided_filter(src, 5, vb), "GuidedFilter (precise)")


@pytest.mark.parametrize("sharpening", [1.0, 1.6])
def test_reassemble(ctx, chain, oracle32, sharpening):
    d0 = chain[ca.RGGB]["desp"]
    n1 = oracle32.downsample(d0, W // 2, H // 2)
    d1 = oracle32.denoise_bilateral(n1, [1e-5] * 3, [5e-4] * 3)
    nlf = [1e-5, 5e-4]
    ref = oracle32.reassemble(d0, n1, d1, sharpening, nlf)
    a, b, c = up(ctx, d0, ca.FMT_RGBA32F), up(ctx, n1, ca.FMT_RGBA32F), up(ctx, d1, ca.FMT_RGBA32F)
    ctx.reassembleImage(a, b, c, sharpening, nlf, a)
    bitexact(a.download(), ref, "reassembleImage")


@pytest.mark.parametrize("contrast,saturation,ebias,blacks", [(1.05, 1.0, 0.0, 0.0), (1.2, 1.3, 0.5, 0.02)])
def test_convert_to_srgb(ctx, chain, oracle32, contrast, saturation, ebias, blacks):
    lin = chain[ca.RGGB]["rgb"]
    cam = np.array([[1.6, -0.4, -0.2], [-0.2, 1.4, -0.2], [0.0, -0.5, 1.5]], np.float32)
    ref8, reff = oracle32.convert_to_srgb(lin, cam, contrast, saturation, 3.5, ebias, blacks)
    p = ca.RGBConversionParameters(contrast, saturation, 3.5, ebias, blacks, 0)
    a = up(ctx, lin, ca.FMT_RGBA32F)
    o8 = ctx.image(W, H, ca.FMT_RGBA8)
    of = ctx.image(W, H, ca.FMT_RGBA32F)
    ctx.convertTosRGB(a, o8, cam, p, of)
    bitexact(of.download(), reff, "convertTosRGB float")
    bitexact(o8.download(), ref8, "convertTosRGB rgba8")


def test_host_matrices(oracle32):
    cam = np.array([[1.6, -0.4, -0.2], [-0.2, 1.4, -0.2], [0.0, -0.5, 1.5]], np.float32)
    bitexact(ca.cam_ycbcr(cam), oracle32.cam_ycbcr(cam), "cam_ycbcr")
    rc, inv = oracle32.inverse3x3(ca.cam_ycbcr(cam))
    bitexact(ca.inverse3x3(ca.cam_ycbcr(cam)), inv, "inverse3x3")


def run_pipeline(ctx, raw_np, params):
    h, w = raw_np.shape
    raw = up(ctx, raw_np, ca.FMT_R16U)
    o8 = ctx.image(w, h, ca.FMT_RGBA8)
    of = ctx.image(w, h, ca.FMT_RGBA32F)
    raw_nlf, pyr_nlf = ctx.raw_to_rgb(raw, o8, params, out_f32=of)
    return o8.download(), of.download(), raw_nlf, pyr_nlf


@pytest.mark.parametrize("bayer", BAYERS)
@pytest.mark.parametrize("algorithm", [ca.BILATERAL, ca.GUIDED_FAST, ca.GUIDED_PRECISE])
def test_raw_to_rgb(ctx, oracle32, bayer, algorithm):
    raw, _ = scene(bayer, seed=bayer)
    p = ca.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    o8, of, raw_nlf, pyr_nlf = run_pipeline(ctx, raw, p)
    r = oracle32.raw_to_rgb(raw, O.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm))
    err = np.abs(of[..., :3].astype(np.float64) - r["rgbf"][..., :3]).max()
    assert err <= 1e-3, f"max-abs {err}"
    bitexact(raw_nlf, r["rawNlf"], "rawNlf")
    bitexact(pyr_nlf, r["pyramidNlf"], "pyramidNlf")
    bitexact(of, r["rgbf"], "raw_to_rgb float")
    bitexact(o8, r["rgba8"], "raw_to_rgb rgba8")


@pytest.mark.parametrize("cb,gb,sharp", [(1.5, 1.0, 1.0), (1.0, 2.0, 1.6), (1.3, 1.5, 1.3)])
def test_raw_to_rgb_denoise_parameters(ctx, oracle32, cb, gb, sharp):
    """The bilateral pyramid with per-level chroma boost, gradient boost and
    sharpening: levels 1-3 are denoised in the launch that reassembles them,
    in the general-edge form when the boosts differ, and sharpen there."""
    raw, _ = scene(ca.RGGB, seed=7)
    p = ca.default_params(ca.RGGB, 64.0, 4095.0, levels=5)
    op = O.default_params(ca.RGGB, 64.0, 4095.0, levels=5)
    for lv in range(5):
        for d in (p.denoiseParameters[lv], op.dp[lv]):
            d.chromaBoost = cb
            d.gradientBoost = gb
            d.sharpening = sharp
    o8, of, raw_nlf, pyr_nlf = run_pipeline(ctx, raw, p)
    r = oracle32.raw_to_rgb(raw, op)
    bitexact(pyr_nlf, r["pyramidNlf"], "pyramidNlf")
    bitexact(of, r["rgbf"], f"raw_to_rgb float cb={cb} gb={gb} sharpening={sharp}")
    bitexact(o8, r["rgba8"], "raw_to_rgb rgba8")


def test_config0_1024x768(ctx, oracle32, oracle64):
    """1024x768 RGGB, 12-bit, black 64, a = 0.002, b = 1e-5, 4-level bilateral."""
    raw = O.generate_bayer(1024, 768, ca.RGGB, 1000, 2000, 0.002, 1e-5, 64.0, 4095.0)
    p = ca.default_params(ca.RGGB, 64.0, 4095.0, levels=4)
    o8, of, raw_nlf, pyr_nlf = run_pipeline(ctx, raw, p)
    op = O.default_params(ca.RGGB, 64.0, 4095.0, levels=4)
    r32 = oracle32.raw_to_rgb(raw, op)
    r64 = oracle64.raw_to_rgb(raw, op)
    d32 = np.abs(of[..., :3].astype(np.float64) - r32["rgbf"][..., :3])
    d64 = np.abs(of[..., :3].astype(np.float64) - r64["rgbf"][..., :3])
    psnr64 = 10 * np.log10(1.0 / np.mean(d64 ** 2))
    print(f"config0: max|gpu-fp32 oracle| {d32.max():.3g}, vs fp64: max {d64.max():.3g} psnr {psnr64:.1f} dB")
    assert d32.max() <= 1e-3
    assert psnr64 >= 60.0
    bitexact(of, r32["rgbf"], "config0 float")
