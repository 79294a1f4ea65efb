"""Local tone mapping on the GPU vs the FP32 oracle (bit equality) and the
FP64 oracle (tolerance): localToneMappingMask as a stand-alone operator
(guide footprint staged in LDS, and read directly where it does not fit),
convertTosRGB with a mask, and the pipeline, where the mask is computed
inside the kernel that reassembles pyramid level 0 (cpp/raw_converter.cpp:164-174)."""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LTM = (0.01, 1.25, 1.0, 1.1)


def bitexact(gpu, ref, name):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        first = tuple(idx[0])
        pytest.fail(f"{name}: {len(idx)} of {same.size} values differ, first at {first}: "
                    f"gpu {gpu[first]!r} oracle {ref[first]!r}")


def ltm_params():
    return ca.LTMParameters(*LTM)


def ycc_image(w, h, seed):
    """Denoised-YCbCr-like input: luma 0.01..0.7 with texture, small chroma."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = 0.01 + 0.6 * (xx / w) * (0.6 + 0.4 * np.cos(yy / 13)) + rng.normal(0, 0.01, (h, w))
    img[..., 1] = 0.03 * np.sin(xx / 17 + yy / 29)
    img[..., 2] = 0.025 * np.cos(xx / 23 - yy / 11)
    return img


def guide_image(gw, gh, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:gh, 0:gw].astype(np.float64)
    g = np.zeros((gh, gw, 4), np.float32)
    g[..., 0] = 0.03 + 0.5 * (xx / max(gw - 1, 1)) + 0.05 * np.sin(yy / 3) + rng.normal(0, 0.005, (gh, gw))
    g[..., 1:3] = rng.normal(0, 0.02, (gh, gw, 2))
    return g


def y2c(o):
    return o.inverse3x3(o.cam_ycbcr(np.eye(3)))[1]


@pytest.mark.parametrize("w,h,gw,gh", [
    (256, 192, 32, 24),    # pipeline ratio 1/8: LDS footprint
    (203, 117, 25, 14),    # ragged edges, ratio ~1/8
    (256, 192, 128, 96),   # ratio 1/2: largest LDS footprint
    (160, 100, 53, 33),    # ratio ~1/3
    (256, 192, 256, 192),  # ratio 1: footprint too wide for LDS, direct reads
    (64, 8, 200, 40),      # guide larger than the mask: direct reads
])
def test_ltm_mask_operator(ctx, oracle32, oracle64, w, h, gw, gh):
    img, guide = ycc_image(w, h, w + h), guide_image(gw, gh, gw)
    T = y2c(oracle32)
    ref = oracle32.ltm_mask(img, guide, LTM, T)
    a, g = ctx.upload(img, ca.FMT_RGBA32F), ctx.upload(guide, ca.FMT_RGBA32F)
    out = ctx.image(w, h, ca.FMT_R32F)
    ctx.localToneMappingMask(a, g, ltm_params(), T, out)
    m = out.download()
    bitexact(m, ref, f"ltm mask {w}x{h} guide {gw}x{gh}")
    # against the FP64 restatement where the reference's expression is conditioned
    m64, ve = oracle64.ltm_mask(img.astype(np.float64), guide.astype(np.float64), LTM, T, var_eps=True)
    good = (np.abs(ve) >= 1e-3) & np.isfinite(m64)
    assert good.mean() > 0.8
    rel = np.abs(m[good] / m64[good] - 1)
    assert rel.max() < 1e-4, rel.max()


def test_convert_to_srgb_with_mask(ctx, oracle32):
    w, h = 256, 192
    lin = np.abs(ycc_image(w, h, 5)) * np.float32(1.4)
    cam = np.array([[1.6, -0.4, -0.2], [-0.2, 1.4, -0.2], [0.0, -0.5, 1.5]], np.float32)
    rng = np.random.default_rng(9)
    mask = rng.uniform(0.6, 1.8, (h, w)).astype(np.float32)
    mask[::7, ::5] = np.nan
    mask[1::9, ::3] = 1.0
    for blacks in (0.0, 0.05):
        ref8, reff = oracle32.convert_to_srgb(lin, cam, 1.05, 0.9, 3.5, -0.3, blacks, ltm_mask=mask)
        p = ca.RGBConversionParameters(1.05, 0.9, 3.5, -0.3, blacks, 1)
        a, mk = ctx.upload(lin, ca.FMT_RGBA32F), ctx.upload(mask, ca.FMT_R32F)
        o8, of = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
        ctx.convertTosRGB(a, o8, cam, p, of, ltm_mask=mk)
        bitexact(of.download(), reff, "convertTosRGB+LTM float")
        bitexact(o8.download(), ref8, "convertTosRGB+LTM rgba8")
    # localToneMapping set without a mask is an error, not a silent skip
    with pytest.raises(ca.ClimError):
        ctx.convertTosRGB(a, o8, cam, ca.RGBConversionParameters(1.05, 0.9, 3.5, -0.3, 0.0, 1), of)


@pytest.mark.parametrize("bayer,w,h,levels", [(ca.RGGB, 256, 192, 5), (ca.GBRG, 200, 136, 5),
                                              (ca.BGGR, 256, 192, 4)])
def test_raw_to_rgb_with_ltm(ctx, oracle32, oracle64, bayer, w, h, levels):
    raw = O.generate_bayer(w, h, bayer, 1000 + bayer, 2000 + bayer, 0.002, 1e-5, 64.0, 4095.0)
    p = ca.default_params(bayer, 64.0, 4095.0, levels=levels)
    p.rgbConversionParameters.localToneMapping = 1
    r = ctx.upload(raw, ca.FMT_R16U)
    o8, of = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
    ctx.raw_to_rgb(r, o8, p, out_f32=of)
    op = O.default_params(bayer, 64.0, 4095.0, levels=levels)
    op.localToneMapping = 1
    r32, r64 = oracle32.raw_to_rgb(raw, op), oracle64.raw_to_rgb(raw, op)
    f = of.download()
    bitexact(f, r32["rgbf"], "raw_to_rgb+LTM float")
    bitexact(o8.download(), r32["rgba8"], "raw_to_rgb+LTM rgba8")
    d = f[..., :3].astype(np.float64) - r64["rgbf"][..., :3]
    psnr = 10 * np.log10(1.0 / np.mean(d ** 2))
    print(f"LTM {w}x{h} levels {levels}: vs fp64 max {np.abs(d).max():.3g} psnr {psnr:.1f} dB")
    assert psnr >= 60.0


def test_raw_to_rgb_ltm_needs_level3(ctx):
    raw = O.generate_bayer(64, 48, ca.RGGB, 1, 2, 0.002, 1e-5, 64.0, 4095.0)
    p = ca.default_params(ca.RGGB, 64.0, 4095.0, levels=3)
    p.rgbConversionParameters.localToneMapping = 1
    r = ctx.upload(raw, ca.FMT_R16U)
    with pytest.raises(ca.ClimError):
        ctx.raw_to_rgb(r, ctx.image(64, 48, ca.FMT_RGBA8), p, writeback=False)


@pytest.mark.parametrize("w,h,levels", [(256, 192, 5), (200, 136, 4)])
def test_pyramidal_denoise_into_fills_levels(ctx, oracle32, w, h, levels):
    """The fused pyramid run in the caller's level images: the downsampled
    levels, the denoised level 0 and the reassembled level 3 (the local tone
    mapping guide) equal the oracle's."""
    raw = O.generate_bayer(w, h, ca.RGGB, 7, 8, 0.002, 1e-5, 64.0, 4095.0)
    op = O.default_params(ca.RGGB, 64.0, 4095.0, levels=levels)
    ycc = oracle32.raw_to_rgb(raw, op, dumps=True)["ycc"]
    dp = [O.DenoiseParams(1.0, 1.0, 1.0, 1.0, 1.0)] * 5
    rc, den0, nlf, g3 = oracle32.pyramidal_denoise(ycc, levels, 0, dp, O.DEFAULT_PYRAMID_NLF, level3=True)
    img = ctx.upload(ycc, ca.FMT_RGBA32F)
    lv = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(1, levels)]
    dn = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(levels)]
    cp = [ca.DenoiseParameters(1.0, 1.0, 1.0, 1.0, 1.0)] * 5
    got_nlf = ctx.pyramidalDenoiseInto(img, levels, ca.BILATERAL, cp, lv, dn,
                                       nlf=np.asarray(O.DEFAULT_PYRAMID_NLF[:levels]))
    bitexact(dn[0].download(), den0, "denoised level 0")
    bitexact(dn[3].download(), g3, "denoised level 3")
    bitexact(got_nlf, nlf[:levels], "per-level noise model")
    prev = ycc
    for i in range(1, levels):
        prev = oracle32.downsample(prev, w >> i, h >> i)
        bitexact(lv[i - 1].download(), prev, f"downsampled level {i}")
    # a level of the wrong size is rejected
    bad = [ctx.image(w >> i, h >> i, ca.FMT_RGBA32F) for i in range(1, levels)]
    bad[1] = ctx.image((w >> 2) + 1, h >> 2, ca.FMT_RGBA32F)
    with pytest.raises(ca.ClimError):
        ctx.pyramidalDenoiseInto(img, levels, ca.BILATERAL, cp, bad, dn)
