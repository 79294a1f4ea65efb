"""The FP16-storage form of the CPU oracle (the mixed-precision pyramid's
level images, and the reference's ARM build, src/gls_image.hpp:32,189-210):
its rounding against numpy's IEEE half conversion, and where the pipeline
applies it."""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("double", [False, True])
def test_half_round_is_ieee_half(double):
    o = O.Oracle(double=double)
    rng = np.random.default_rng(5)
    vals = np.concatenate([
        rng.normal(size=20000) * 10.0 ** rng.integers(-9, 5, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, -65520.0, 1e9, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26,
                  2.0 ** -14, 2.0 ** -14 - 2.0 ** -25, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, np.inf, -np.inf]),
        # ties between FP16 neighbours (round to even)
        (np.arange(1, 2000, dtype=np.float64) + 0.5) * 2.0 ** -10,
    ])
    dt = np.float64 if double else np.float32
    vals = vals.astype(dt)
    got = o.half_round(vals)
    want = vals.astype(np.float16).astype(dt)
    assert np.array_equal(got, want, equal_nan=True), np.flatnonzero(got != want)[:10]
    assert np.isnan(o.half_round(np.array([np.nan], dt)))[0]


def test_pipeline_stores_levels_as_fp16(oracle32):
    w, h = 192, 128
    raw = O.generate_bayer(w, h, 2, 5, 6, 0.004, 2e-5, 64.0, 4095.0)
    p = O.default_params(2, 64.0, 4095.0, levels=5)
    plain = oracle32.raw_to_rgb(raw, p, dumps=True)
    p.fp16_storage = 1
    half = oracle32.raw_to_rgb(raw, p, dumps=True)
    # the despeckled level 0 is stored as FP16; level 0's reassembly (the
    # den0 dump) goes to the tone map unstored
    v = half["ycc"][..., :3]
    assert np.array_equal(v, v.astype(np.float16).astype(np.float32))
    v = half["den0"][..., :3]
    assert not np.array_equal(v, v.astype(np.float16).astype(np.float32))
    # the level-0 YCbCr image is the plain one rounded (the demosaic is not affected)
    assert np.array_equal(half["ycc"], plain["ycc"].astype(np.float16).astype(np.float32))
    # and the result stays near the FP32 one
    d = half["rgbf"][..., :3].astype(np.float64) - plain["rgbf"][..., :3]
    psnr = 10 * np.log10(1.0 / np.mean(d * d))
    assert psnr > 55.0, psnr
    assert not np.array_equal(half["rgbf"], plain["rgbf"])
