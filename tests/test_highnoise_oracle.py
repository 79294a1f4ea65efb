"""Pins of the high-noise raw branch in the CPU oracle (noiseLevel > 0.6,
cpp/raw_converter.cpp:111-123): bayerToRawRGBA / rawRGBAToBayer, the raw
RGBA despeckle (despeckle_3x3x4) and denoiseRawRGBA.

Checks: numpy slicing for the re-layouts, a sort-based brute force for the
despeckle (a different algorithm from the reference's running trackers),
closed forms for the denoise (constant image; a noise model so large that
every tap counts = the 5x5 box mean of scipy; so small that only equal taps
count = identity) and a brute force of its distance test."""
import numpy as np
import pytest
from scipy import ndimage

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


def quads(w, h, seed, lo=0.05, hi=0.9):
    return np.random.default_rng(seed).uniform(lo, hi, (h, w, 4))


def padded_windows(img, r):
    """(h, w, (2r+1)^2, C) windows with clamp-to-edge, y outer / x inner."""
    h, w = img.shape[:2]
    p = np.pad(img, ((r, r), (r, r), (0, 0)), mode="edge")
    return np.stack([p[r + dy:r + dy + h, r + dx:r + dx + w] for dy in range(-r, r + 1) for dx in range(-r, r + 1)],
                    axis=2)


@pytest.mark.parametrize("bayer", [0, 1, 2, 3])
def test_relayout_slicing_and_round_trip(o64, bayer):
    rng = np.random.default_rng(bayer)
    raw = rng.uniform(0, 1, (48, 64))
    q = o64.bayer_to_raw_rgba(raw, bayer)
    for c, (ox, oy) in enumerate(OFFS[bayer]):
        np.testing.assert_array_equal(q[..., c], raw[oy::2, ox::2])
    np.testing.assert_array_equal(o64.raw_rgba_to_bayer(q, bayer), raw)


@pytest.mark.parametrize("double", [False, True])
def test_despeckle_sort_brute_force(o32, o64, double):
    o = o64 if double else o32
    img = quads(40, 24, 1)
    img[5, 7, 0] = 3.0      # a bright speck
    img[9, 3, 2] = 0.0      # a dark speck
    img[12, 12:15, 1] = 2.0  # a short bright line: two of three survive
    img = img.astype(np.float32 if not double else np.float64)
    got = o.despeckle_raw_rgba(img)
    win = padded_windows(img.astype(np.float64), 1)          # (h, w, 9, 4)
    hi = np.sort(np.concatenate([win, np.zeros(win.shape[:2] + (2, 4))], 2), axis=2)[:, :, -2]
    lo = np.sort(np.concatenate([win, np.full(win.shape[:2] + (2, 4), 65535.0)], 2), axis=2)[:, :, 1]
    exp = np.minimum(np.maximum(img, lo), hi)
    np.testing.assert_array_equal(got, exp)
    assert got[5, 7, 0] < 1 and got[9, 3, 2] > 0 and got[12, 13, 1] == 2.0


@pytest.mark.parametrize("double", [False, True])
def test_denoise_closed_forms(o32, o64, double):
    o = o64 if double else o32
    dt = np.float64 if double else np.float32
    img = quads(32, 20, 2).astype(dt)
    # constant image: identity
    const = np.full((20, 32, 4), 0.3, dt)
    np.testing.assert_allclose(o.denoise_raw_rgba(const, [1e-3] * 4), const, rtol=1e-15 if double else 4e-7)
    # a huge noise model: every tap counts, 5x5 box mean with clamped edges
    box = ndimage.uniform_filter(img.astype(np.float64), size=(5, 5, 1), mode="nearest")
    np.testing.assert_allclose(o.denoise_raw_rgba(img, [1e6] * 4), box, rtol=1e-12 if double else 2e-6)
    # a tiny one: only the centre counts (at the borders its clamped copies
    # too: n * s / n, exact up to rounding)
    got = o.denoise_raw_rgba(img, [1e-20] * 4)
    np.testing.assert_array_equal(got[2:-2, 2:-2], img[2:-2, 2:-2])
    np.testing.assert_allclose(got, img, rtol=1e-15 if double else 2.5e-7)


@pytest.mark.parametrize("double", [False, True])
def test_denoise_brute_force_distance(o32, o64, double):
    """The joint test over the four channels: a tap counts when
    sum_c (s_c - p_c)^2 / (var_c p_c) < 1.  Noise of 0.35 sigma per channel puts
    the distances around 1; pixels whose taps come within 1e-4 of the
    threshold are skipped (FP32 may decide them either way)."""
    o = o64 if double else o32
    rng = np.random.default_rng(4)
    var = np.array([2e-4, 1e-4, 3e-4, 1.5e-4])
    base = 0.4
    img = base + 0.35 * np.sqrt(var * base) * rng.normal(size=(24, 36, 4))
    img = img.astype(np.float64 if double else np.float32).astype(np.float64)
    got = o.denoise_raw_rgba(img, var)
    var32 = np.asarray(var, np.float32).astype(np.float64)
    win = padded_windows(img, 2)
    d = ((win - img[:, :, None, :]) ** 2 / (var32 * img)[:, :, None, :]).sum(-1)
    wgt = (d < 1).astype(np.float64)
    exp = (wgt[..., None] * win).sum(2) / wgt.sum(2)[..., None]
    clear = (np.abs(d - 1) > 1e-4).all(-1)
    assert clear.mean() > 0.8 and 0.2 < wgt.mean() < 0.9
    np.testing.assert_allclose(got[clear], exp[clear], rtol=1e-12 if double else 2e-6)


@pytest.mark.parametrize("double", [False, True])
def test_denoise_zero_channel_is_nan(o32, o64, double):
    """sigma = sqrt(var * 0) = 0: no tap passes the test and the quad is
    0/0 = NaN, as the reference's expression is under IEEE arithmetic."""
    o = o64 if double else o32
    img = quads(16, 12, 3)
    img[6, 8, 2] = 0.0
    got = o.denoise_raw_rgba(img, [1e-3] * 4)
    assert np.isnan(got[6, 8]).all()
    assert np.isfinite(np.delete(got.reshape(-1, 4), 6 * 16 + 8, axis=0)).all()


@pytest.mark.parametrize("a", [0.002, 0.02])
def test_pipeline_high_noise_fp32_against_fp64(o32, o64, a):
    raw = O.generate_bayer(256, 192, 2, 1000, 2000, a, a / 200, 64.0, 4095.0)
    p = O.default_params(2, 64.0, 4095.0, levels=5)
    p.noiseLevel = 0.7
    r32, r64 = o32.raw_to_rgb(raw, p), o64.raw_to_rgb(raw, p)
    f32, f64 = r32["rgbf"][..., :3].astype(np.float64), r64["rgbf"][..., :3]
    np.testing.assert_array_equal(np.isnan(f32), np.isnan(f64))
    ok = np.isfinite(f64)
    psnr = 10 * np.log10(1.0 / np.mean((f32[ok] - f64[ok]) ** 2))
    assert psnr >= 60.0, psnr
    p.noiseLevel = 0.0
    assert np.nanmax(np.abs(o64.raw_to_rgb(raw, p)["rgbf"][..., :3] - f64)) > 0.05  # the branch is active
