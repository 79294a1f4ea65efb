"""Pins of the local tone mapping restatement in the CPU oracle:
localToneMappingMask (cl/demosaic.cl:1024-1107) and the LTM branch of
convertTosRGB (cl/demosaic.cl:1456-1466).

The reference has no test or fixture for this stage either; the checks are
closed forms that follow from the sampler's definition and from algebra
(a constant image, a linear guide, gray pixels), chosen so that a wrong tap
offset or scale, swapped axes, a dropped term or swapped mix operands fail."""
import numpy as np
import pytest

from oracle import oracle as O

LTM = (0.01, 1.25, 1.0, 1.1)  # LTMParameters defaults, cpp/demosaic.hpp:54-59


@pytest.fixture(scope="module")
def o64():
    return O.Oracle(double=True)


@pytest.fixture(scope="module")
def o32():
    return O.Oracle(double=False)


def ycc_to_rgb(o):
    _, y2c = o.inverse3x3(o.cam_ycbcr(np.eye(3)))
    return y2c.astype(np.float64)


def expected_mask(illum, luma, rgb, ltm):
    # the parameters as the float values the operators receive
    eps, sh, hi, det = (float(np.float32(v)) for v in ltm)
    with np.errstate(invalid="ignore"):
        L = np.sqrt(2 * rgb[..., 0] + 2 * rgb[..., 1] + 2 * rgb[..., 2])
    L = np.where((rgb >= 0).all(-1), L, np.nan)
    hc = np.where(1 < L, 1.0, L)
    tc = sh + (hi - sh) * hc
    with np.errstate(invalid="ignore"):
        return illum ** (1 / tc) * (luma / illum) ** det / luma


@pytest.mark.parametrize("double", [False, True])
def test_constant_image_closed_form(o32, o64, double):
    """Constant guide g and constant pixel: every tap reads g, the guided
    weights reduce to the box weights, illuminance = g, and the mask is
    g^(1/tc) (L/g)^detail / L."""
    o = o64 if double else o32
    T = ycc_to_rgb(o64)
    h, w = 48, 64
    for g, pix, nan in [(0.3, (0.25, 0.02, -0.01), False), (0.05, (0.6, -0.03, 0.04), False),
                        (0.3, (0.02, 0.0, 0.3), True),   # a negative camera component: NaN
                        (1.0, (0.02, 0.0, 0.3), False)]:  # ... unless the illuminance is 1: pow(1, NaN) = 1
        img = np.zeros((h, w, 4))
        img[..., :3] = pix
        guide = np.full((h // 8, w // 8, 4), g)
        m = o.ltm_mask(img, guide, LTM, T)
        rgb = np.asarray(T, np.float64) @ np.asarray(pix)
        exp = expected_mask(np.float64(g), pix[0], rgb[None, :], LTM)[0]
        if nan:
            assert np.isnan(exp) and np.isnan(m).all()
        else:
            np.testing.assert_allclose(m, exp, rtol=1e-12 if double else 3e-6)


@pytest.mark.parametrize("double", [False, True])
def test_linear_guide_pins_tap_geometry(o32, o64, double):
    """A guide linear in its texel indices is reproduced exactly by the
    linear sampler, so each tap reads a + b u + c v at its own position
    u = (x + o + 1/2) * gw / w - 1/2 (o in -1.5, 0.5, 2.0 texels of the MASK),
    with box weights (2/5, 2/5, 1/5) per axis.  The guided average then has
    the closed form  mean + (L - mean) S / (var + eps),  S = sum w (s - mean) s,
    var = sum (w s)^2 - mean^2  (the weights sum to one)."""
    o = o64 if double else o32
    h, w = 96, 128
    gh, gw = h // 8, w // 8
    a, b, c = 0.2, 0.011, 0.0043  # different slopes per axis: a swapped axis fails
    jj, ii = np.mgrid[0:gh, 0:gw].astype(np.float64)
    guide = np.zeros((gh, gw, 4))
    guide[..., 0] = a + b * ii + c * jj
    guide[..., 1:3] = 0.5  # ignored channels
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 4))
    img[..., 0] = 0.3 + 0.1 * np.sin(xx / 9) * np.cos(yy / 7)
    T = np.eye(3)
    m = o.ltm_mask(img, guide, LTM, T)

    off = np.array([-1.5, 0.5, 2.0])
    cw = np.array([0.4, 0.4, 0.2])
    U = (xx[..., None] + off + 0.5) * gw / w - 0.5          # (h, w, 3) tap columns
    V = (yy[..., None] + off + 0.5) * gh / h - 0.5          # (h, w, 3) tap rows
    s = a + b * U[..., None, :] + c * V[..., :, None]       # (h, w, 3[y], 3[x])
    wt = np.outer(cw, cw)
    mean = (wt * s).sum((-1, -2))
    var = ((wt * s) ** 2).sum((-1, -2)) - mean ** 2
    S = (wt * (s - mean[..., None, None]) * s).sum((-1, -2))
    L = img[..., 0]
    illum = mean + (L - mean) * S / (var + float(np.float32(LTM[0])))
    exp = expected_mask(illum, L, np.stack([L, 0 * L, 0 * L], -1), LTM)
    # interior: no tap reaches a clamped edge texel
    inside = (U.min(-1) >= 0) & (U.max(-1) <= gw - 1) & (V.min(-1) >= 0) & (V.max(-1) <= gh - 1)
    assert inside.sum() > 0.6 * h * w
    np.testing.assert_allclose(m[inside], exp[inside], rtol=1e-11 if double else 2e-5)


@pytest.mark.parametrize("double", [False, True])
def test_tone_ltm_branch_gray_gain_and_limits(o32, o64, double):
    """convertTosRGB's LTM branch runs after the tone curve and before the
    black lift.  On gray pixels below white (luma l = v < 1) both branches of
    the mix equal boost * v, so the mask acts as a plain gain; at white
    (l = 1) the pixel is kept; boost 1 and NaN leave the pixel alone."""
    o = o64 if double else o32
    h, w = 4, 64
    v = np.linspace(0.0, 1.2, w)
    img = np.zeros((h, w, 4))
    img[..., :3] = v[None, :, None]
    boosts = np.array([0.5, 1.0, 1.6, np.nan])
    mask = np.repeat(boosts[:, None], w, 1)
    rtol = 1e-12 if double else 2e-6
    for blacks in (0.0, 0.1):
        _, base = o.convert_to_srgb(img, np.eye(3), contrast=1.0, blacks=0.0)
        _, got = o.convert_to_srgb(img, np.eye(3), contrast=1.0, blacks=blacks, ltm_mask=mask)
        t = base[..., :3].astype(np.float64)
        assert (t[0, :, 0] == 1).any() and (t[0, :, 0] < 0.99).sum() > 40
        gain = np.where(np.isnan(boosts), 1.0, boosts)[:, None, None]
        gain = np.where((t < 0.99) | (gain < 1), gain, np.where(t == 1, 1.0, np.nan))
        keep = np.isfinite(gain)
        exp = t * np.where(keep, gain, 0)
        if blacks > 0:
            b32 = float(np.float32(blacks))
            exp = (exp - b32) / (1 - b32)
        exp = np.where(keep, exp, got[..., :3])
        np.testing.assert_allclose(got[..., :3], exp, rtol=rtol, atol=1e-6 if not double else 1e-12)


@pytest.mark.parametrize("double", [False, True])
def test_tone_ltm_branch_colour(o32, o64, double):
    """boost > 1 on a coloured pixel: mix(rgb * boost, 1 - (1 - rgb)(1 - boost l)/(1 - l), sqrt(l)),
    l the BT.709 luma of the tone-mapped pixel (the gain dominates in the
    shadows, the highlight-preserving branch in the highlights)."""
    o = o64 if double else o32
    img = np.zeros((1, 3, 4))
    img[0, :, :3] = [[0.02, 0.05, 0.01], [0.3, 0.2, 0.1], [0.9, 0.6, 0.8]]
    mask = np.full((1, 3), 1.7)
    _, base = o.convert_to_srgb(img, np.eye(3), contrast=1.0)
    _, got = o.convert_to_srgb(img, np.eye(3), contrast=1.0, ltm_mask=mask)
    rgb = base[0, :, :3].astype(np.float64)
    lum = rgb @ np.array([0.2126, 0.7152, 0.0722])
    hi = 1 - (1 - rgb) * ((1 - 1.7 * lum) / (1 - lum))[:, None]
    exp = rgb * 1.7 + (hi - rgb * 1.7) * np.sqrt(lum)[:, None]
    np.testing.assert_allclose(got[0, :, :3], exp, rtol=1e-12 if double else 3e-6)


def test_fp32_mask_matches_fp64_where_conditioned(o32, o64):
    """The two restatements on the same inputs, including guide values where
    the reference's denominator var + eps = eps - (1 - sum w^2) mean^2 (on flat
    areas) crosses zero (mean ~ 0.107): there the mask is ill-conditioned
    and only well-conditioned pixels are compared."""
    rng = np.random.default_rng(3)
    h, w = 192, 256
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 4))
    img[..., 0] = 0.02 + 0.5 * (xx / w) + 0.05 * np.sin(yy / 11) + rng.normal(0, 0.005, (h, w))
    img[..., 1] = 0.02 * np.cos(xx / 13)
    img[..., 2] = -0.015 * np.sin(yy / 17)
    gy, gx = np.mgrid[0:h // 8, 0:w // 8].astype(np.float64)
    guide = np.zeros((h // 8, w // 8, 4))
    guide[..., 0] = 0.02 + 0.5 * (gx * 8 / w) + 0.05 * np.sin(gy * 8 / 11)
    T = ycc_to_rgb(o64)
    m64, ve = o64.ltm_mask(img, guide, LTM, T, var_eps=True)
    m32 = o32.ltm_mask(img.astype(np.float32), guide.astype(np.float32), LTM, T)
    assert (np.abs(ve) < 1e-3).any()  # the band is present in this input
    good = np.abs(ve) >= 1e-3
    assert good.mean() > 0.9
    np.testing.assert_array_equal(np.isnan(m32[good]), np.isnan(m64[good]))
    f = good & np.isfinite(m64)
    rel = np.abs(m32[f] / m64[f] - 1)
    assert rel.max() < 1e-4, rel.max()


@pytest.mark.parametrize("bayer", [0, 2])
def test_pipeline_with_ltm_fp32_against_fp64(o32, o64, bayer):
    raw = O.generate_bayer(256, 192, bayer, 1000 + bayer, 2000 + bayer, 0.002, 1e-5, 64.0, 4095.0)
    p = O.default_params(bayer, 64.0, 4095.0, levels=5)
    p.localToneMapping = 1
    r32, r64 = o32.raw_to_rgb(raw, p), o64.raw_to_rgb(raw, p)
    d = r32["rgbf"][..., :3].astype(np.float64) - r64["rgbf"][..., :3]
    assert np.isfinite(d).all()
    psnr = 10 * np.log10(1.0 / np.mean(d ** 2))
    assert psnr >= 60.0, psnr
    p.localToneMapping = 0
    off = o64.raw_to_rgb(raw, p)
    assert np.abs(off["rgbf"] - r64["rgbf"]).max() > 0.05  # the stage is active
