"""The half-resolution preview (RawConverter::fastDemosaicImage) on the GPU
vs the FP32 oracle (bit equality): fasteDebayer alone, and the whole preview
as one kernel from the uint16 mosaic to RGBA8."""
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


@pytest.mark.parametrize("bayer", [ca.GRBG, ca.GBRG, ca.RGGB, ca.BGGR])
def test_fast_debayer(ctx, oracle32, bayer):
    raw = np.random.default_rng(bayer).uniform(0, 1.3, (46, 70)).astype(np.float32)
    r, q = ctx.upload(raw, ca.FMT_R32F), ctx.image(35, 23, ca.FMT_RGBA32F)
    ctx.fasteDebayer(r, q, bayer)
    bitexact(q.download(), oracle32.fast_debayer(raw, bayer), "fasteDebayer")


@pytest.mark.parametrize("bayer,w,h,variant", [(ca.RGGB, 256, 192, 0), (ca.GBRG, 202, 138, 1),
                                               (ca.BGGR, 4032, 3024, 0)])
def test_fast_demosaic(ctx, oracle32, oracle64, bayer, w, h, variant):
    raw = O.generate_bayer(w, h, bayer, 21 + bayer, 22, 0.002, 1e-5, 64.0, 4095.0)
    p = ca.default_params(bayer, 64.0, 4095.0)
    op = O.default_params(bayer, 64.0, 4095.0)
    if variant:
        cam = [1.6, -0.4, -0.2, -0.2, 1.4, -0.2, 0.0, -0.5, 1.5]
        for k in range(9):
            p.rgb_cam[k // 3][k % 3] = cam[k]
            op.rgb_cam[k] = cam[k]
        p.rgbConversionParameters.saturation = op.saturation = 1.2
        p.rgbConversionParameters.exposureBias = op.exposureBias = 0.3
        p.rgbConversionParameters.blacks = op.blacks = 0.02
    r = ctx.upload(raw, ca.FMT_R16U)
    o8, of = ctx.image(w // 2, h // 2, ca.FMT_RGBA8), ctx.image(w // 2, h // 2, ca.FMT_RGBA32F)
    ctx.fast_demosaic(r, o8, p, out_f32=of)
    e8, ef = oracle32.fast_demosaic(raw, op)
    bitexact(of.download(), ef, f"fastDemosaicImage float {w}x{h}")
    bitexact(o8.download(), e8, f"fastDemosaicImage rgba8 {w}x{h}")
    if w <= 512:
        _, f64 = oracle64.fast_demosaic(raw, op)
        assert np.abs(of.download()[..., :3].astype(np.float64) - f64[..., :3]).max() <= 1e-3
    with pytest.raises(ca.ClimError):
        ctx.fast_demosaic(r, ctx.image(w // 2 + 1, h // 2, ca.FMT_RGBA8), p)
