"""Pins of the CPU oracle (oracle/clim_oracle.c) on the CPU.

The reference ships no tests, fixtures or golden vectors for this path, so
every oracle stage is pinned here by what the mathematics fixes: closed
forms, invariants, special cases that reduce to textbook routines (numpy /
scipy: median, sort, uniform box filters, least squares, matrix inverse,
tanh), brute force on small inputs, and shift equivariance across the four
CFA layouts.  Each check is chosen so that a plausible transcription
mistake (wrong offset or sign, transposed operand, dropped term) fails it.
"""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

BAYERS = [0, 1, 2, 3]
# OpenCL Bayer table (cl/demosaic.cl:29-34): (x, y) of R, G, B, G2
OFFS = {0: [(1, 0), (0, 0), (0, 1), (1, 1)], 1: [(0, 1), (0, 0), (1, 0), (1, 1)],
        2: [(0, 0), (0, 1), (1, 1), (1, 0)], 3: [(1, 1), (0, 1), (0, 0), (1, 0)]}
COLOR_OF_SITE = [0, 1, 2, 1]


def cfa_color_map(bayer, w, h):
    cm = np.zeros((h, w), np.int64)
    for c, (ox, oy) in enumerate(OFFS[bayer]):
        cm[oy::2, ox::2] = COLOR_OF_SITE[c]
    return cm


def mosaic(rgb, bayer):
    h, w, _ = rgb.shape
    cm = cfa_color_map(bayer, w, h)
    return np.take_along_axis(rgb, cm[..., None], axis=2)[..., 0]


def gather_clamped(img, xs, ys):
    h, w = img.shape[:2]
    return img[np.clip(ys, 0, h - 1), np.clip(xs, 0, w - 1)]


@pytest.fixture(scope="module")
def o64():
    return O.Oracle(double=True)


@pytest.fixture(scope="module")
def o32():
    return O.Oracle(double=False)


def smooth_rgb(w, h, seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 4))
    for c in range(3):
        img[..., c] = 0.3 + 0.1 * c + 0.05 * np.sin(xx / (7 + c)) + 0.04 * np.cos(yy / (5 + 2 * c))
    img[..., :3] += rng.normal(0, 0.01, (h, w, 3))
    return img


# ------------------------------------------------------------------ scale
@pytest.mark.parametrize("bayer", BAYERS)
def test_scale_raw_closed_form(o64, bayer):
    rng = np.random.default_rng(bayer)
    raw = rng.integers(0, 4096, (48, 64)).astype(np.uint16)
    sm = [3.0, 1.5, 2.0, 1.25]
    black = 64.0 / 65535.0
    out = o64.scale_raw(raw, bayer, sm, black)
    ref = np.zeros_like(out)
    for c, (ox, oy) in enumerate(OFFS[bayer]):
        v = raw[oy::2, ox::2].astype(np.float64) / 65535.0
        ref[oy::2, ox::2] = np.maximum(np.float32(sm[c]) * (v - np.float32(black)), 0.0)
    np.testing.assert_array_equal(out, ref)


# --------------------------------------------------------------- raw stats
@pytest.mark.parametrize("bayer", BAYERS)
def test_raw_noise_statistics_brute_force(o64, bayer):
    rng = np.random.default_rng(10 + bayer)
    h, w = 24, 32
    img = rng.uniform(0, 1, (h, w))
    mean, var = o64.raw_noise_statistics(img, bayer)
    qy, qx = np.mgrid[0:h // 2, 0:w // 2]
    for c, (ox, oy) in enumerate(OFFS[bayer]):
        samples = np.stack([gather_clamped(img, 2 * (qx + i) + ox, 2 * (qy + j) + oy)
                            for j in range(-2, 3) for i in range(-2, 3)])
        np.testing.assert_allclose(mean[..., c], samples.mean(0), rtol=0, atol=1e-14)
        np.testing.assert_allclose(var[..., c], samples.var(0), rtol=0, atol=1e-13)
    # interior quads agree with a textbook 5x5 box filter of each CFA plane
    for c, (ox, oy) in enumerate(OFFS[bayer]):
        plane = img[oy::2, ox::2]
        np.testing.assert_allclose(mean[2:-2, 2:-2, c], ndimage.uniform_filter(plane, 5)[2:-2, 2:-2], atol=1e-14)


# ------------------------------------------------------------ noise fit
def fit3_numpy(m, v):
    """The three-pass regression of cpp/pyramidal_denoise.cpp:259-362 with numpy's lstsq."""
    ok = np.all((m >= 0.001) & (m <= 0.5), 1) & np.all((v >= 0) & (v <= 0.001), 1)

    def line(x, y):
        # slope by least squares, floored; the intercept uses the floored slope
        b = max(np.polyfit(x, y, 1)[0], 1e-8)
        return max(y.mean() - b * x.mean(), 1e-8), b

    A, B = np.array([line(m[ok, c], v[ok, c]) for c in range(v.shape[1])]).T
    diff = A + B * m[ok] - v[ok]
    err2 = np.sqrt((diff ** 2).mean(0))
    inl = np.all(diff ** 2 <= err2, 1)
    A2, B2 = np.array([line(m[ok][inl, c], v[ok][inl, c]) for c in range(v.shape[1])]).T
    return A2, B2


def test_fit_recovers_exact_line(o64):
    rng = np.random.default_rng(4)
    n = 5000
    m = rng.uniform(0.01, 0.45, (n, 4))
    A = np.array([1e-5, 2e-5, 3e-5, 4e-5])
    B = np.array([1e-3, 1.2e-3, 0.9e-3, 1.1e-3])
    v = A + B * m
    mean = np.zeros((1, n, 4)); var = np.zeros((1, n, 4))
    mean[0], var[0] = m, v
    rc, fit = o64.nlf_fit_raw(mean, var)
    assert rc == 0
    np.testing.assert_allclose(fit[:4], A, rtol=1e-5)
    np.testing.assert_allclose(fit[4:], B, rtol=1e-5)


def test_fit_matches_numpy_three_pass(o64):
    rng = np.random.default_rng(5)
    n = 20000
    m = rng.uniform(0.0, 0.6, (n, 4))
    v = 2e-5 + 1e-3 * m + rng.normal(0, 3e-5, (n, 4))
    mean = m[None].copy(); var = v[None].copy()
    rc, fit = o64.nlf_fit_raw(mean, var)
    A, B = fit3_numpy(m, v)
    assert rc == 0
    np.testing.assert_allclose(fit[:4], A, rtol=2e-6)
    np.testing.assert_allclose(fit[4:], B, rtol=2e-6)
    # level variant: one shared mean, three variances
    stats = np.zeros((1, n, 4))
    stats[0, :, 0] = m[:, 0]
    stats[0, :, 1:] = v[:, :3]
    rc, fit = o64.nlf_fit_level(stats)
    A, B = fit3_numpy(np.repeat(m[:, :1], 3, 1), v[:, :3])
    np.testing.assert_allclose(fit[:3], A, rtol=2e-6)
    np.testing.assert_allclose(fit[3:], B, rtol=2e-6)


def test_fit_with_flat_half_patches(o32, o64):
    """Flat patches at 0.5 give window means of exactly 0.5 with zero variance:
    in range, and the largest mean term (2^32 at 2^-33 resolution)."""
    rng = np.random.default_rng(6)
    n = 4000
    m = rng.uniform(0.01, 0.45, (n, 4))
    v = 2e-5 + 1e-3 * m + rng.normal(0, 2e-6, (n, 4))
    m[: n // 4] = 0.5
    v[: n // 4] = 0.0
    mean = m[None].astype(np.float32).astype(np.float64)
    var = v[None].astype(np.float32).astype(np.float64)
    rc64, f64 = o64.nlf_fit_raw(mean, var)
    rc32, f32 = o32.nlf_fit_raw(mean.astype(np.float32), var.astype(np.float32))
    A, B = fit3_numpy(mean[0], var[0])
    assert rc64 == 0 and rc32 == 0
    np.testing.assert_allclose(f64[:4], A, rtol=1e-5)
    np.testing.assert_allclose(f64[4:], B, rtol=1e-5)
    np.testing.assert_allclose(f32, f64, rtol=1e-5)


def test_fit_empty_sample_reports_fallback(o64):
    mean = np.full((4, 4, 4), 0.9)  # every mean above 0.5
    var = np.full((4, 4, 4), 1e-4)
    rc, _ = o64.nlf_fit_raw(mean, var)
    assert rc == 1


def test_fit_on_known_poisson_gaussian_noise(o64):
    """Flat CFA patches with noise variance a*x + b (normalised units) give
    B close to a and A close to b through scale + raw stats + fit."""
    rng = np.random.default_rng(6)
    a, b = 0.001, 1e-5
    h, w = 512, 512
    slope = 0.43 / w
    level = np.tile(0.02 + slope * np.arange(w), (h, 1))
    noisy = level + rng.normal(0, 1, (h, w)) * np.sqrt(a * level + b)
    mean, var = o64.raw_noise_statistics(noisy, 2)
    rc, fit = o64.nlf_fit_raw(mean, var)
    assert rc == 0
    # the 5x5-quad variance is the biased (n = 25) estimator; the ramp adds
    # the variance of five samples two pixels apart
    ramp_var = (2 * slope) ** 2 * 2.0
    np.testing.assert_allclose(fit[4:], a * 24 / 25, rtol=0.05)
    np.testing.assert_allclose(fit[:4], b * 24 / 25 + ramp_var, rtol=0.3)


# ------------------------------------------------------------ demosaic
@pytest.mark.parametrize("bayer", BAYERS)
def test_green_constant_and_gray_ramp(o64, bayer):
    h, w = 32, 40
    flat = np.full((h, w), 0.37)
    np.testing.assert_allclose(o64.interpolate_green(flat, bayer, 1e-3), flat, atol=1e-15)
    yy, xx = np.mgrid[0:h, 0:w]
    ramp = 0.1 + 0.004 * xx + 0.003 * yy
    g = o64.interpolate_green(ramp, bayer, 1e-3)
    np.testing.assert_allclose(g[3:-3, 3:-3], ramp[3:-3, 3:-3], atol=1e-12)


@pytest.mark.parametrize("bayer", BAYERS)
def test_red_blue_constant_colour(o64, bayer):
    h, w = 32, 40
    rgb = np.zeros((h, w, 3))
    rgb[..., 0], rgb[..., 1], rgb[..., 2] = 0.6, 0.3, 0.15
    m = mosaic(rgb, bayer)
    g = o64.interpolate_green(m, bayer, 1e-3)
    out = o64.interpolate_red_blue(m, g, bayer, 1e-3)
    np.testing.assert_allclose(out[3:-3, 3:-3, :3], rgb[3:-3, 3:-3], atol=1e-12)
    assert np.all(out[..., 3] == 0)


def test_demosaic_shift_equivariance(o64):
    """The four layouts are one another shifted by a pixel; demosaicking the
    matching shifted mosaic must give the shifted result (interior)."""
    rng = np.random.default_rng(7)
    h, w = 48, 64
    base = ndimage.gaussian_filter(rng.uniform(0.05, 0.9, (h + 2, w + 2, 3)), (2, 2, 0))
    # RGGB at (x, y) == GRBG at (x-1, y) == GBRG at (x, y-1) == BGGR at (x-1, y-1)
    shifts = {2: (0, 0), 0: (1, 0), 1: (0, 1), 3: (1, 1)}
    outs = {}
    for bayer, (sx, sy) in shifts.items():
        scene = base[sy:sy + h, sx:sx + w]
        m = mosaic(scene, bayer)
        g = o64.interpolate_green(m, bayer, 1e-4)
        outs[bayer] = o64.interpolate_red_blue(m, g, bayer, 1e-4)
    ref = outs[2]
    for bayer, (sx, sy) in shifts.items():
        a = outs[bayer][6:-6, 6:-6]
        b = ref[6 + sy:h - 6 + sy, 6 + sx:w - 6 + sx]
        np.testing.assert_allclose(a, b, atol=1e-12, err_msg=f"bayer {bayer}")


# ------------------------------------------------------- colour stages
def test_highlights_identity_and_invariants(o64):
    rng = np.random.default_rng(8)
    img = np.zeros((16, 16, 4))
    img[..., :3] = rng.uniform(0, 2.0, (16, 16, 3))
    out = o64.blend_highlights(img, 1.0)
    low = np.all(img[..., :3] <= 1.0, -1)
    np.testing.assert_array_equal(out[low, :3], img[low, :3])
    # the channel sum is preserved, the chroma magnitude is the clipped pixel's
    hi = ~low
    np.testing.assert_allclose(out[hi, :3].sum(-1), img[hi, :3].sum(-1), rtol=1e-12)
    T = np.array([[np.sqrt(3), -np.sqrt(3), 0], [-1, -1, 2]])
    chroma_out = np.linalg.norm((T @ out[hi, :3].T), axis=0)
    chroma_clip = np.linalg.norm((T @ np.minimum(img[hi, :3], 1.0).T), axis=0)
    np.testing.assert_allclose(chroma_out, chroma_clip, rtol=1e-6, atol=1e-12)


def test_transform_matmul(o64):
    rng = np.random.default_rng(9)
    img = rng.normal(size=(8, 12, 4))
    M = rng.normal(size=(3, 3)).astype(np.float32)
    out = o64.transform(img, M)
    np.testing.assert_allclose(out[..., :3], img[..., :3] @ M.astype(np.float64).T, atol=1e-14)


def test_cam_ycbcr_matches_reference_table_and_inverse(o64):
    m = o64.cam_ycbcr(np.eye(3))
    # srgb_ycbcr as tabulated in cpp/demosaic_utils.cpp:830-834
    table = np.array([[0.2126, 0.7152, 0.0722], [-0.1146, -0.3854, 0.5], [0.5, -0.4542, -0.0458]])
    np.testing.assert_allclose(m, table, atol=1e-4)
    rc, inv = o64.inverse3x3(m)
    assert rc == 0
    np.testing.assert_allclose(inv, np.linalg.inv(m.astype(np.float64)), rtol=1e-6, atol=1e-6)
    # ycbcr_srgb of cpp/demosaic_utils.cpp:836-840
    np.testing.assert_allclose(inv, [[1, 0, 1.5748], [1, -0.1873, -0.4681], [1, 1.8556, 0]], atol=2e-4)
    rc, _ = o64.inverse3x3(np.ones((3, 3)))
    assert rc == 1


@pytest.mark.parametrize("double", [False, True])
def test_despeckle_brute_force(double):
    o = O.Oracle(double)
    rng = np.random.default_rng(11)
    h, w = 20, 24
    img = np.zeros((h, w, 4))
    img[..., :3] = rng.uniform(0, 1, (h, w, 3))
    img[5, 7, 0] = 9.0   # isolated outlier
    img[9, 9, 0] = -0.5  # negative luma (the max tracker starts at 0)
    va, vb = [2.1e-4, 0, 0], [2.6e-4, 0, 0]
    out = o.despeckle(img, va, vb)
    yy, xx = np.mgrid[0:h, 0:w]
    nb = np.stack([gather_clamped(img, xx + i, yy + j) for j in (-1, 0, 1) for i in (-1, 0, 1)])
    tol = 1e-6 if not double else 1e-14
    np.testing.assert_allclose(out[..., 1], np.median(nb[..., 1], 0), atol=tol)
    np.testing.assert_allclose(out[..., 2], np.median(nb[..., 2], 0), atol=tol)
    y = nb[..., 0]
    hi = np.sort(np.concatenate([y, np.zeros((2, h, w))]), 0)
    lo = np.sort(np.concatenate([y, np.full((2, h, w), 65535.0)]), 0)
    firstMax, secMax, firstMin, secMin = hi[-1], hi[-2], lo[0], lo[1]
    s = img[..., 0]
    sigma = np.sqrt(va[0] + vb[0] * s)

    def ss(e0, e1, x):
        t = np.clip((x - e0) / (e1 - e0), 0, 1)
        return t * t * (3 - 2 * t)

    minv = secMin + (firstMin - secMin) * ss(sigma, 4 * sigma, secMin - firstMin)
    maxv = secMax + (firstMax - secMax) * ss(sigma, 4 * sigma, firstMax - secMax)
    ref = np.minimum(np.maximum(s, minv), maxv)
    ok = np.isfinite(sigma)
    np.testing.assert_allclose(out[..., 0][ok], ref[ok], atol=1e-6 if not double else 1e-12)
    assert out[5, 7, 0] < 9.0 or True  # documented quirk: isolated outliers far above the gap are kept
    assert out[5, 7, 0] == pytest.approx(9.0, abs=1e-6)


# ---------------------------------------------------------------- pyramid
def test_downsample_exact_box(o64):
    rng = np.random.default_rng(12)
    h, w = 32, 48
    img = rng.uniform(0, 1, (h, w, 4))
    out = o64.downsample(img, w // 2, h // 2)
    yy, xx = np.mgrid[0:h // 2, 0:w // 2]
    acc = sum(gather_clamped(img, 2 * xx + i, 2 * yy + j) for j in range(-1, 3) for i in range(-1, 3)) / 16
    np.testing.assert_allclose(out, acc, atol=1e-14)
    imp = np.zeros((16, 16, 4)); imp[8, 8] = 1.0
    d = o64.downsample(imp, 8, 8)
    assert d[4, 4, 0] == pytest.approx(1 / 16) and d[3, 3, 0] == pytest.approx(1 / 16)
    assert d[..., 0].sum() == pytest.approx(4 / 16)
    const = np.full((15, 23, 4), 0.42)
    np.testing.assert_allclose(o64.downsample(const, 11, 7), 0.42, atol=1e-14)


def test_downsample_general_ratio_close_to_bilinear(o64):
    rng = np.random.default_rng(13)
    img = ndimage.gaussian_filter(rng.uniform(0, 1, (75, 93, 4)), (3, 3, 0))
    ow, oh = 46, 37
    out = o64.downsample(img, ow, oh)
    yy, xx = np.mgrid[0:oh, 0:ow].astype(np.float64)
    ref = np.zeros_like(out)
    for dy in (-0.5, 0.5):
        for dx in (-0.5, 0.5):
            u = (xx + 0.5 + dx) / ow * 93 - 0.5
            v = (yy + 0.5 + dy) / oh * 75 - 0.5
            for c in range(4):
                ref[..., c] += ndimage.map_coordinates(img[..., c], [v, u], order=1, mode="nearest") / 4
    # differs from exact bilinear only through the 8-bit weight rounding
    assert np.abs(out - ref).max() < 2e-3


def test_noise_statistics_uniform_filter(o64):
    rng = np.random.default_rng(14)
    img = rng.uniform(0, 1, (30, 34, 4))
    st = o64.noise_statistics(img)
    m = ndimage.uniform_filter(img[..., 0], 5, mode="nearest")
    np.testing.assert_allclose(st[..., 0], m, atol=1e-14)
    for c in range(3):
        e2 = ndimage.uniform_filter(img[..., c] ** 2, 5, mode="nearest")
        e1 = ndimage.uniform_filter(img[..., c], 5, mode="nearest")
        np.testing.assert_allclose(st[..., 1 + c], e2 - e1 ** 2, atol=1e-13)


def test_bilateral_constant_identity_and_brute_force(o64):
    const = np.full((16, 16, 4), 0.3); const[..., 3] = 0
    np.testing.assert_allclose(o64.denoise_bilateral(const, [1e-4] * 3, [1e-3] * 3), const, atol=1e-15)
    rng = np.random.default_rng(15)
    img = np.zeros((24, 28, 4)); img[..., :3] = 0.3 + rng.normal(0, 0.02, (24, 28, 3))
    # sigma -> 0: only the centre passes the threshold
    np.testing.assert_allclose(o64.denoise_bilateral(img, [1e-20] * 3, [0.0] * 3)[..., :3], img[..., :3], atol=1e-15)
    va, vb = [1e-5, 2e-5, 2e-5], [1e-3, 5e-4, 5e-4]
    out = o64.denoise_bilateral(img, va, vb, 1.3, 1.0)
    h, w = img.shape[:2]
    yy, xx = np.mgrid[0:h, 0:w]
    sig = np.sqrt(np.array(va) + np.array(vb) * img[..., :1])
    num = np.zeros((h, w, 3)); den = np.zeros((h, w, 3))
    edge = np.array([1.0, 1.3, 1.3]) * np.sqrt(2)
    for j in range(-2, 3):
        for i in range(-2, 3):
            s = gather_clamped(img, xx + i, yy + j)[..., :3]
            d = np.linalg.norm((s - img[..., :3]) / sig, axis=-1)
            wgt = (d[..., None] < edge).astype(np.float64)
            num += wgt * s; den += wgt
    np.testing.assert_allclose(out[..., :3], num / den, atol=1e-13)


def test_guided_fast_limits_and_textbook(o64):
    rng = np.random.default_rng(16)
    h, w = 20, 22
    img = np.zeros((h, w, 4)); img[..., :3] = rng.uniform(0.1, 0.9, (h, w, 3))
    box = np.stack([ndimage.uniform_filter(img[..., c], 5, mode="nearest") for c in range(3)], -1)
    big = o64.denoise_guided_fast(img, [1e6] * 3, [0.0] * 3)
    np.testing.assert_allclose(big[..., :3], box, atol=1e-6)
    small = o64.denoise_guided_fast(img, [1e-14] * 3, [0.0] * 3)
    np.testing.assert_allclose(small[..., :3], img[..., :3], atol=1e-6)
    # textbook guided filter (He et al., eq. 14-15, a and b not re-averaged)
    va, vb = [1e-4, 2e-4, 3e-4], [1e-3, 1e-3, 1e-3]
    out = o64.denoise_guided_fast(img, va, vb)
    I = img[..., :3]
    mu = box
    cov = np.stack([np.stack([ndimage.uniform_filter(I[..., a] * I[..., b], 5, mode="nearest") - mu[..., a] * mu[..., b]
                              for b in range(3)], -1) for a in range(3)], -2)
    eps = np.array(va) + np.array(vb) * I[..., :1]
    Ae = cov + np.einsum("...i,ij->...ij", eps, np.eye(3))
    a = np.linalg.solve(Ae, cov)  # a[...,k,c]: coefficient of channel k for output c (Sigma symmetric)
    b = mu - np.einsum("...kc,...k->...c", a, mu)
    q = np.einsum("...kc,...k->...c", a, I) + b
    np.testing.assert_allclose(out[..., :3], q, atol=1e-9)


def test_guided_precise_textbook(o64):
    rng = np.random.default_rng(17)
    h, w = 18, 21
    img = np.zeros((h, w, 4)); img[..., :3] = rng.uniform(0.1, 0.9, (h, w, 3))
    eps = [2e-3, 1e-3, 5e-4]
    out = o64.guided_filter(img, 5, eps)

    def box(x):
        return ndimage.uniform_filter1d(ndimage.uniform_filter1d(x, 5, axis=1, mode="nearest"), 5, axis=0,
                                        mode="nearest")
    I = img[..., :3]
    mu = np.stack([box(I[..., c]) for c in range(3)], -1)
    S = np.stack([np.stack([box(I[..., a] * I[..., b]) - mu[..., a] * mu[..., b] for b in range(3)], -1)
                  for a in range(3)], -2)
    # the reference scales the red epsilon by the local mean (cl/guided_filter.cl:58)
    E = np.zeros_like(S)
    E[..., 0, 0] = eps[0] * mu[..., 0]; E[..., 1, 1] = eps[1]; E[..., 2, 2] = eps[2]
    a = np.linalg.solve(S + E, S)
    b = mu - np.einsum("...kc,...k->...c", a, mu)
    abar = np.stack([np.stack([box(a[..., k, c]) for c in range(3)], -1) for k in range(3)], -2)
    bbar = np.stack([box(b[..., c]) for c in range(3)], -1)
    q = np.einsum("...kc,...k->...c", abar, I) + bbar
    np.testing.assert_allclose(out[..., :3], q, atol=1e-9)


def test_reassemble_identity_and_upsample(o64):
    rng = np.random.default_rng(18)
    d0 = rng.uniform(0, 1, (16, 20, 4))
    n1 = rng.uniform(0, 1, (8, 10, 4))
    out = o64.reassemble(d0, n1, n1, 1.0, [1e-4, 1e-3])
    np.testing.assert_allclose(out, d0, atol=1e-15)
    # d1 = 0 isolates the bilinear 2x upsample of n1: weights 1/4, 3/4
    z = np.zeros_like(n1)
    out = o64.reassemble(np.zeros_like(d0), n1, z, 1.0, [1e-4, 1e-3])
    h, w = 16, 20
    yy, xx = np.mgrid[0:h, 0:w]
    x0 = (xx - 1) // 2; y0 = (yy - 1) // 2
    ax = np.where(xx % 2 == 0, 0.75, 0.25); ay = np.where(yy % 2 == 0, 0.75, 0.25)
    up = ((1 - ax) * (1 - ay))[..., None] * gather_clamped(n1, x0, y0) + \
         (ax * (1 - ay))[..., None] * gather_clamped(n1, x0 + 1, y0) + \
         ((1 - ax) * ay)[..., None] * gather_clamped(n1, x0, y0 + 1) + \
         (ax * ay)[..., None] * gather_clamped(n1, x0 + 1, y0 + 1)
    np.testing.assert_allclose(out[..., 1:], -up[..., 1:], atol=1e-14)


# -------------------------------------------------------------- tone map
def test_tone_curve_endpoints_monotone_and_tanh(o32):
    ramp = np.zeros((1, 256, 4), np.float32)
    ramp[0, :, :3] = np.linspace(0, 1 / 0.95, 256)[:, None]
    o8, of = o32.convert_to_srgb(ramp, np.eye(3), contrast=1.0)
    assert np.all(of[0, 0, :3] == 0.0)
    assert np.all(np.abs(of[0, -1, :3] - 1.0) < 1e-6)
    assert np.all(np.diff(of[0, :, 0]) >= 0)
    assert o8[0, -1, 0] == 255 and o8[0, 0, 0] == 0
    s = 3.5
    sig = lambda t: 0.5 * (np.tanh(s * t - 0.3 * s) + 1)
    x = ramp[0, :, 0].astype(np.float64)
    ref = np.clip((sig(np.sqrt(0.95 * x)) - sig(0)) / (sig(1) - sig(0)), 0, 1)
    np.testing.assert_allclose(of[0, :, 0], ref, atol=2e-6)
    np.testing.assert_array_equal(o8[0, :, 0], np.rint(np.clip(of[0, :, 0], 0, 1) * 255).astype(np.uint8))


# ----------------------------------------------------------- end to end
@pytest.mark.parametrize("bayer", BAYERS)
def test_fp32_against_fp64_oracle(o32, o64, bayer):
    raw = O.generate_bayer(256, 192, bayer, 1000 + bayer, 2000 + bayer, 0.002, 1e-5, 64.0, 4095.0)
    p = O.default_params(bayer, 64.0, 4095.0, levels=5)
    r32, r64 = o32.raw_to_rgb(raw, p), o64.raw_to_rgb(raw, p)
    d = r32["rgbf"][..., :3].astype(np.float64) - r64["rgbf"][..., :3]
    psnr = 10 * np.log10(1.0 / np.mean(d ** 2))
    assert psnr >= 60.0, psnr
    np.testing.assert_allclose(r32["rawNlf"], r64["rawNlf"], rtol=1e-2, atol=1e-6)


def test_oracle_deterministic_across_thread_counts(o32):
    gomp = C.CDLL("libgomp.so.1")
    raw = O.generate_bayer(256, 192, 2, 1000, 2000, 0.004, 2e-5, 64.0, 4095.0)
    p = O.default_params(2, 64.0, 4095.0, levels=5)
    gomp.omp_set_num_threads(1)
    a = o32.raw_to_rgb(raw, p)
    gomp.omp_set_num_threads(7)
    b = o32.raw_to_rgb(raw, p)
    gomp.omp_set_num_threads(8)
    np.testing.assert_array_equal(a["rgbf"], b["rgbf"])
    np.testing.assert_array_equal(a["pyramidNlf"], b["pyramidNlf"])


def test_scene_noise_is_poisson_gaussian():
    a, b, black, white = 0.004, 2e-5, 64.0, 4095.0
    raw, clean = O.generate_bayer(512, 384, 2, 1000, 2000, a, b, black, white, clean=True)
    x = mosaic(clean[..., :3].astype(np.float64), 2)
    dn = raw.astype(np.float64)
    mu = black + x * (white - black)
    sd = np.sqrt(a * x + b) * (white - black)
    ok = (x > 0.005) & (mu + 4 * sd < white)
    z = (dn[ok] - mu[ok]) / sd[ok]
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.03
    # high-count pixels (normal branch) and low-count pixels (inversion) both fit
    lam = x / a
    for sel in (lam < 64, lam >= 64):
        zz = ((dn - mu) / sd)[ok & sel]
        assert abs(zz.var() - 1.0) < 0.05


def _fixed_bilinear32(img, i0, j0, a, b):
    """The sampler's weighted sum for given texel and weights, in float32 and
    in the sampler's order (w00 t00 + w10 t10 + w01 t01 + w11 t11)."""
    f = np.float32
    a, b = f(a), f(b)
    w00, w10, w01, w11 = (f(1) - a) * (f(1) - b), a * (f(1) - b), (f(1) - a) * b, a * b
    t00, t10 = gather_clamped(img, i0, j0), gather_clamped(img, i0 + 1, j0)
    t01, t11 = gather_clamped(img, i0, j0 + 1), gather_clamped(img, i0 + 1, j0 + 1)
    return ((w00 * t00 + w10 * t10) + w01 * t01) + w11 * t11


def test_sampler_exact_2x_ratios(o32):
    """At exact 2x ratios the linear sampler's float32 coordinate arithmetic
    puts every tap on a fixed texel with a fixed weight, for every output
    size up to 6144 (downsample) and 16384 (upsample in reassemble, the
    sharpening gradient's four samples included); the HIP
    kernels use those taps directly within these limits.  Checked
    exhaustively, and through the FP32 oracle at sizes up to the limits."""
    f = np.float32

    def weight(u):
        fu = np.floor(u)
        return fu, np.floor((u - fu) * f(256) + f(0.5)) * f(0.00390625)

    for wo in range(1, 6145):
        x = np.arange(wo, dtype=np.float32)
        nx = f(1) / f(wo)
        px, sx = (x + f(0.5)) * nx, f(0.5) * nx
        for tap, off in ((px - sx, -1), (px + sx, 1)):
            fu, a = weight(tap * f(2 * wo) - f(0.5))
            assert np.all(a == f(0.5)) and np.all(fu == 2 * x + off), wo
    for w in range(2, 16385, 2):
        x = np.arange(w)
        fu, a = weight((x.astype(np.float32) + f(0.5)) * (f(1) / f(w)) * f(w // 2) - f(0.5))
        assert np.all(a == np.where(x % 2 == 0, f(0.75), f(0.25))), w
        assert np.all(fu == np.where(x % 2 == 0, x // 2 - 1, (x - 1) // 2)), w
    # the sharpening gradient's samples at px +- 1/w (and py +- 1/h): texels
    # floor(x/2) and floor(x/2) - 1 with weight 1 - a
    for w in range(2, 16385, 2):
        x = np.arange(w)
        nx = f(1) / f(w)
        px = (x.astype(np.float32) + f(0.5)) * nx
        a_side = np.where(x % 2 == 0, f(0.25), f(0.75))
        for t, i in ((px + nx, x // 2), (px - nx, x // 2 - 1)):
            fu, a = weight(t * f(w // 2) - f(0.5))
            assert np.all(a == a_side) and np.all(fu == i), w
    # the limit is tight: 6145 is the first downsample size that breaks it
    x = np.arange(6145, dtype=np.float32)
    nx = f(1) / f(6145)
    bad = [not np.all(weight(t * f(12290) - f(0.5))[1] == f(0.5)) or
           not np.all(weight(t * f(12290) - f(0.5))[0] == 2 * x + o)
           for t, o in (((x + f(0.5)) * nx - f(0.5) * nx, -1), ((x + f(0.5)) * nx + f(0.5) * nx, 1))]
    assert any(bad)
    # the oracle's downsample and reassemble at the largest sizes
    rng = np.random.default_rng(21)
    img = rng.uniform(0, 1, (4, 2 * 6144, 4)).astype(np.float32)
    out = o32.downsample(img, 6144, 2)
    yy, xx = np.mgrid[0:2, 0:6144]
    acc = _fixed_bilinear32(img, 2 * xx - 1, 2 * yy - 1, 0.5, 0.5)
    for i0, j0 in ((2 * xx + 1, 2 * yy - 1), (2 * xx - 1, 2 * yy + 1), (2 * xx + 1, 2 * yy + 1)):
        acc = acc + _fixed_bilinear32(img, i0, j0, 0.5, 0.5)
    np.testing.assert_array_equal(out, f(0.25) * acc)
    d0 = rng.uniform(0, 1, (2, 16384, 4)).astype(np.float32)
    n1 = rng.uniform(0, 1, (1, 8192, 4)).astype(np.float32)
    d1 = rng.uniform(0, 1, (1, 8192, 4)).astype(np.float32)
    out = o32.reassemble(d0, n1, d1, 1.0, [1e-4] * 6)
    yy, xx = np.mgrid[0:2, 0:16384]
    i0 = np.where(xx % 2 == 1, (xx - 1) // 2, xx // 2 - 1)
    j0 = np.where(yy % 2 == 1, (yy - 1) // 2, yy // 2 - 1)
    a = np.where(xx % 2 == 1, f(0.25), f(0.75))[..., None]
    b = np.where(yy % 2 == 1, f(0.25), f(0.75))[..., None]
    up_n = _fixed_bilinear32(n1, i0, j0, a, b)
    up_d = _fixed_bilinear32(d1, i0, j0, a, b)
    want = d0 - (up_n - up_d)
    # d.x = mix(up(d1).x, d.x, sharpening) with sharpening = 1
    want[..., 0] = up_d[..., 0] + (want[..., 0] - up_d[..., 0]) * f(1)
    np.testing.assert_array_equal(out, want)


def test_relaxed_forms_keep_edge_semantics(o32, o64):
    """The FP32 path's contracted forms must decide like the square-root and
    division forms at their edges: zero noise in the green interpolation (a
    black patch: lumaStdDev = 0) and a negative noise variance in the
    bilateral (sigma NaN: every weight zero, the output NaN)."""
    rng = np.random.default_rng(31)
    h, w = 40, 56
    scaled = rng.uniform(0.05, 0.4, (h, w)).astype(np.float32)
    scaled[10:30, 12:40] = 0.0
    g32 = o32.interpolate_green(scaled, 2, 1e-3)
    g64 = o64.interpolate_green(scaled.astype(np.float64), 2, 1e-3)
    np.testing.assert_allclose(g32, g64, atol=2e-6)
    img = rng.uniform(-0.2, 0.6, (24, 32, 4))
    va, vb = [1e-4, 1e-8, 1e-8], [1e-3, 1e-3, 1.2e-3]
    b32 = o32.denoise_bilateral(img.astype(np.float32), va, vb)
    b64 = o64.denoise_bilateral(img, va, vb)
    assert np.isnan(b64).any()
    np.testing.assert_array_equal(np.isnan(b32), np.isnan(b64))
    ok = ~np.isnan(b64)
    np.testing.assert_allclose(b32[ok], b64[ok], atol=1e-5)
