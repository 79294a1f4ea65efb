"""rotate_180 of RawConverter::demosaicImage (clim_raw_to_rgb_ex) on the GPU
against the FP32 oracle (bit equality) and the FP64 oracle (PSNR), for every
CFA pattern, both denoisers' families and the high-noise branch."""
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


@pytest.mark.parametrize("bayer,w,h,algorithm,noise", [
    (ca.RGGB, 256, 192, 0, 0.0), (ca.GRBG, 258, 130, 1, 0.0), (ca.GBRG, 320, 96, 2, 0.0),
    (ca.BGGR, 192, 160, 0, 0.7)])
def test_raw_to_rgb_rotate_180(ctx, oracle32, oracle64, bayer, w, h, algorithm, noise):
    raw = O.generate_bayer(w, h, bayer, 50 + bayer, 60 + bayer, 0.004, 2e-5, 64.0, 4095.0)
    p = ca.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    p.noiseLevel = noise
    r = ctx.upload(raw, ca.FMT_R16U)
    o8, of = ctx.image(w, h, ca.FMT_RGBA8), ctx.image(w, h, ca.FMT_RGBA32F)
    # the output buffers hold a previous unrotated frame: row 0 / column 0 must not leak it
    ctx.raw_to_rgb(r, o8, p, out_f32=of, writeback=False)
    raw_nlf, pyr_nlf = ctx.raw_to_rgb(r, o8, p, out_f32=of, rotate_180=True)
    op = O.default_params(bayer, 64.0, 4095.0, levels=5, algorithm=algorithm)
    op.noiseLevel = noise
    op.rotate_180 = 1
    r32, r64 = oracle32.raw_to_rgb(raw, op), oracle64.raw_to_rgb(raw, op)
    f = of.download()
    bitexact(raw_nlf, r32["rawNlf"], "rawNlf")
    bitexact(pyr_nlf, r32["pyramidNlf"], "pyramidNlf")
    bitexact(f, r32["rgbf"], "rotated float")
    bitexact(o8.download(), r32["rgba8"], "rotated rgba8")
    ok = np.isfinite(r64["rgbf"][..., :3])
    d = f[..., :3].astype(np.float64)[ok] - r64["rgbf"][..., :3][ok]
    mse = np.mean(d ** 2)
    psnr = 10 * np.log10(1.0 / mse) if mse > 0 else np.inf
    print(f"rotate_180 bayer {bayer} {w}x{h} alg {algorithm}: vs fp64 max {np.abs(d).max():.3g} psnr {psnr:.1f} dB")
    assert psnr >= 60.0
