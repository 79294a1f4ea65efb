"""The high-noise raw branch (noiseLevel > 0.6, cpp/raw_converter.cpp:111-123)
on the GPU vs the FP32 oracle (bit equality) and the FP64 oracle (tolerance):
the four stand-alone operators, and the pipeline, where they run as one
kernel between the raw noise fit and the demosaic."""
import numpy as np
import pytest

import climage_amd as ca
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def bitexact(gpu, ref, name):
    gpu, ref = np.asarray(gpu), np.asarray(ref)
    assert gpu.shape == ref.shape, (name, gpu.shape, ref.shape)
    same = (gpu == ref) | (np.isnan(gpu) & np.isnan(ref))
    if not same.all():
        idx = np.argwhere(~same)
        first = tuple(idx[0])
        pytest.fail(f"{name}: {len(idx)} of {same.size} values differ, first at {first}: "
                    f"gpu {gpu[first]!r} oracle {ref[first]!r}")


def noisy_quads(qw, qh, seed, var):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:qh, 0:qw].astype(np.float64)
    base = 0.05 + 0.5 * (xx / qw) * (0.7 + 0.3 * np.sin(yy / 5))
    q = base[..., None] * np.array([0.9, 1.0, 1.1, 1.0]) + np.sqrt(np.asarray(var) * base[..., None]) * \
        rng.normal(size=(qh, qw, 4))
    q = np.maximum(q, 0.0)
    q[rng.uniform(size=(qh, qw, 4)) < 0.002] = 1.5  # specks
    return q.astype(np.float32)


@pytest.mark.parametrize("bayer", [ca.GRBG, ca.GBRG, ca.RGGB, ca.BGGR])
def test_relayouts(ctx, oracle32, bayer):
    rng = np.random.default_rng(bayer)
    raw = rng.uniform(0, 1, (46, 70)).astype(np.float32)
    r = ctx.upload(raw, ca.FMT_R32F)
    q = ctx.image(35, 23, ca.FMT_RGBA32F)
    ctx.bayerToRawRGBA(r, q, bayer)
    bitexact(q.download(), oracle32.bayer_to_raw_rgba(raw, bayer), "bayerToRawRGBA")
    back = ctx.image(70, 46, ca.FMT_R32F)
    ctx.rawRGBAToBayer(q, back, bayer)
    bitexact(back.download(), raw, "rawRGBAToBayer")
    with pytest.raises(ca.ClimError):
        ctx.bayerToRawRGBA(r, ctx.image(34, 23, ca.FMT_RGBA32F), bayer)


@pytest.mark.parametrize("qw,qh", [(64, 16), (37, 23), (1, 1), (300, 9)])
def test_despeckle_raw_rgba(ctx, oracle32, qw, qh):
    img = noisy_quads(qw, qh, qw, [3e-4] * 4)
    a, o = ctx.upload(img, ca.FMT_RGBA32F), ctx.image(qw, qh, ca.FMT_RGBA32F)
    ctx.despeckleRawRGBAImage(a, o)
    bitexact(o.download(), oracle32.despeckle_raw_rgba(img), f"despeckleRawRGBA {qw}x{qh}")


@pytest.mark.parametrize("qw,qh", [(64, 16), (37, 23), (2, 3)])
def test_denoise_raw_rgba(ctx, oracle32, oracle64, qw, qh):
    var = [2e-4, 1e-4, 3e-4, 1.5e-4]
    img = noisy_quads(qw, qh, 7 + qw, var)
    img[min(1, qh - 1), min(1, qw - 1), 2] = 0.0  # a zero channel: that quad is NaN
    a, o = ctx.upload(img, ca.FMT_RGBA32F), ctx.image(qw, qh, ca.FMT_RGBA32F)
    ctx.denoiseRawRGBAImage(a, var, o)
    got = o.download()
    bitexact(got, oracle32.denoise_raw_rgba(img, var), f"denoiseRawRGBA {qw}x{qh}")
    assert np.isnan(got[min(1, qh - 1), min(1, qw - 1)]).all()
    r64 = oracle64.denoise_raw_rgba(img.astype(np.float64), var)
    ok = np.isfinite(r64)
    assert np.array_equal(ok, np.isfinite(got))
    # rare tap decisions at distance ~1 may differ from FP64: compare in bulk
    assert np.mean(np.abs(got[ok] - r64[ok]) > 1e-5) < 0.01


@pytest.mark.parametrize("bayer,a,ltm", [(ca.RGGB, 0.002, 0), (ca.GRBG, 0.02, 0), (ca.BGGR, 0.02, 1)])
def test_raw_to_rgb_high_noise(ctx, oracle32, oracle64, bayer, a, ltm):
    w, h = 256, 192
    raw = O.generate_bayer(w, h, bayer, 1000 + bayer, 2000 + bayer, a, a / 200, 64.0, 4095.0)
    p = ca.default_params(bayer, 64.0, 4095.0, levels=5)
    p.noiseLevel = 0.7
    p.rgbConversionParameters.localToneMapping = ltm
    r = ctx.upload(raw, ca.FMT_R16U)
    o8, of = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
    raw_nlf, pyr_nlf = ctx.raw_to_rgb(r, o8, p, out_f32=of)
    op = O.default_params(bayer, 64.0, 4095.0, levels=5)
    op.noiseLevel = 0.7
    op.localToneMapping = ltm
    r32, r64 = oracle32.raw_to_rgb(raw, op), oracle64.raw_to_rgb(raw, op)
    f = of.download()
    bitexact(raw_nlf, r32["rawNlf"], "rawNlf")
    bitexact(pyr_nlf, r32["pyramidNlf"], "pyramidNlf")
    bitexact(f, r32["rgbf"], "raw_to_rgb high-noise float")
    bitexact(o8.download(), r32["rgba8"], "raw_to_rgb high-noise rgba8")
    ok = np.isfinite(r64["rgbf"][..., :3])
    d = f[..., :3].astype(np.float64)[ok] - r64["rgbf"][..., :3][ok]
    mse = np.mean(d ** 2)
    psnr = 10 * np.log10(1.0 / mse) if mse > 0 else np.inf
    print(f"high-noise a={a}: vs fp64 max {np.abs(d).max():.3g} psnr {psnr:.1f} dB")
    assert psnr >= 60.0
    # the branch changes the result
    p.noiseLevel = 0.0
    o8b, ofb = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
    ctx.raw_to_rgb(r, o8b, p, out_f32=ofb, writeback=False)
    assert np.nanmax(np.abs(ofb.download() - f)) > 0.05
