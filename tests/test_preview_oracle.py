"""Pins of the half-resolution preview in the CPU oracle
(RawConverter::fastDemosaicImage, cpp/raw_converter.cpp:188-224): fasteDebayer
against numpy slicing by the Bayer table, and the preview as the composition
of its stages (each pinned in tests/test_oracle.py)."""
import numpy as np
import pytest

from oracle import oracle as O

# OpenCL Bayer table (cl/demosaic.cl:29-34): (x, y) of R, G, B, G2
OFFS = {0: [(1, 0), (0, 0), (0, 1), (1, 1)], 1: [(0, 1), (0, 0), (1, 0), (1, 1)],
        2: [(0, 0), (0, 1), (1, 1), (1, 0)], 3: [(1, 1), (0, 1), (0, 0), (1, 0)]}


@pytest.fixture(scope="module")
def o64():
    return O.Oracle(double=True)


@pytest.fixture(scope="module")
def o32():
    return O.Oracle(double=False)


@pytest.mark.parametrize("bayer", [0, 1, 2, 3])
def test_fast_debayer_slicing(o64, bayer):
    raw = np.random.default_rng(bayer).uniform(0, 1, (38, 54))
    got = o64.fast_debayer(raw, bayer)
    (rx, ry), (gx, gy), (bx, by), (g2x, g2y) = OFFS[bayer]
    np.testing.assert_array_equal(got[..., 0], raw[ry::2, rx::2])
    np.testing.assert_array_equal(got[..., 1], (raw[gy::2, gx::2] + raw[g2y::2, g2x::2]) / 2)
    np.testing.assert_array_equal(got[..., 2], raw[by::2, bx::2])
    assert (got[..., 3] == 0).all()


@pytest.mark.parametrize("double", [False, True])
def test_fast_demosaic_is_the_composition(o32, o64, double):
    o = o64 if double else o32
    raw = O.generate_bayer(128, 96, 1, 11, 12, 0.004, 2e-5, 64.0, 4095.0)
    p = O.default_params(1, 64.0, 4095.0)
    cam = np.array([[1.6, -0.4, -0.2], [-0.2, 1.4, -0.2], [0.0, -0.5, 1.5]], np.float32)
    for k in range(9):
        p.rgb_cam[k] = float(cam.ravel()[k])
    p.saturation, p.exposureBias, p.blacks = 1.2, 0.3, 0.02
    o8, of = o.fast_demosaic(raw, p)
    sc = o.scale_raw(raw, 1, list(p.scale_mul), np.float32(64.0) / np.float32(65535.0))
    rgb = o.blend_highlights(o.fast_debayer(sc, 1), 1.0)
    e8, ef = o.convert_to_srgb(rgb, cam, 1.05, 1.2, 3.5, 0.3, 0.02)
    np.testing.assert_array_equal(of, ef)
    np.testing.assert_array_equal(o8, e8)


def test_fast_demosaic_fp32_against_fp64(o32, o64):
    raw = O.generate_bayer(256, 192, 2, 3, 4, 0.002, 1e-5, 64.0, 4095.0)
    p = O.default_params(2, 64.0, 4095.0)
    _, f32 = o32.fast_demosaic(raw, p)
    _, f64 = o64.fast_demosaic(raw, p)
    d = f32[..., :3].astype(np.float64) - f64[..., :3]
    assert np.abs(d).max() < 1e-4
