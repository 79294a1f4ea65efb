"""Closed-form pins of the oracle's direction, high-frequency, diagonal, gradient-boost
and sharpening logic, and planted mistakes that each pin catches.

The smooth inputs of tests/test_oracle.py leave parts of these stages free:
on a constant or linear mosaic the horizontal and vertical green estimates
agree and the Laplacian term vanishes, and constant colour makes both
diagonal colour-difference means equal.  The inputs here are built so that
the reference's definitions (cl/demosaic.cl:96-129 sobel and bilateralSobel,
:160-227 interpolateGreen, :229-318 interpolateRedBlue, :735-765
denoiseImage, :867-896 reassembleImage) reduce to values derived by hand:

* a step edge (constant along one axis): the along-edge green estimate is the
  true green and the direction logic must select it;
* a gray quadratic f(u), u = x + y: both green estimates equal f + c - 2 c g
  (c the quadratic coefficient, g the high-frequency gain), the gradient is
  (16/3)(b + 2cu) in closed form and whiteness is f / (f + 2c);
* a colour-difference step along one diagonal: the other diagonal's mean is
  the true value and the alpha mix must select it;
* a ramp under reassembleImage: dx = slope / 2 exactly (the taps sit half a
  coarse texel apart), so the sharpening gate is a textbook smoothstep.

test_planted_mistakes_fail_the_pins compiles copies of oracle/clim_oracle.c
(FP64) with one plausible transcription mistake each and checks that the pin
written for it fails.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
from scipy import ndimage

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "clim_oracle.c")

# OpenCL Bayer table (cl/demosaic.cl:29-34): (x, y) of R, G, B, G2
OFFS = {0: [(1, 0), (0, 0), (0, 1), (1, 1)], 1: [(0, 1), (0, 0), (1, 0), (1, 1)],
        2: [(0, 0), (0, 1), (1, 1), (1, 0)], 3: [(1, 1), (0, 1), (0, 0), (1, 0)]}
COLOR_OF_SITE = [0, 1, 2, 1]


def site_map(bayer, w, h):
    """CFA site index (0 R, 1 G, 2 B, 3 G2) of every pixel."""
    s = np.zeros((h, w), np.int64)
    for c, (ox, oy) in enumerate(OFFS[bayer]):
        s[oy::2, ox::2] = c
    return s


def mosaic(rgb, bayer):
    h, w, _ = rgb.shape
    cm = np.take(COLOR_OF_SITE, site_map(bayer, w, h))
    return np.take_along_axis(rgb, cm[..., None], axis=2)[..., 0]


def smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


# ------------------------------------------------------------- green: direction
def step_scene(w, h, x0, vertical):
    """Dyadic R, G, B values stepping at column x0 (vertical edge) or row x0."""
    lo, hi = np.array([0.25, 0.375, 0.5]), np.array([0.75, 0.625, 0.125])
    yy, xx = np.mgrid[0:h, 0:w]
    side = (xx >= x0) if vertical else (yy >= x0)
    return np.where(side[..., None], hi, lo)


def check_green_step_edges(o):
    """Green at every non-green site of a step edge equals the true green: the
    direction logic must pick the along-edge (vertical for a vertical edge)
    estimate, whose Laplacian term is exactly zero."""
    for bayer in (0, 1, 2, 3):
        for vertical in (True, False):
            for x0 in (13, 14):
                rgb = step_scene(36, 30, x0, vertical)
                raw = mosaic(rgb, bayer)
                g = o.interpolate_green(raw, bayer, 1e-4)
                # clamped reads past the border return texels of another CFA
                # colour: compare where the stencils stay inside
                m = slice(4, -4)
                np.testing.assert_allclose(g[m, m], rgb[m, m, 1], rtol=0, atol=1e-12,
                                           err_msg=f"bayer {bayer} vertical {vertical} edge {x0}")


def test_green_step_edges_take_the_along_edge_estimate(o64):
    check_green_step_edges(o64)


# ------------------------------------------------ green: high-frequency term
def check_green_quadratic_closed_form(o):
    """Gray f(u) = f0 + b u + c u^2, u = x + y.  Both estimates are
    f + c - 2 c hf_gain (low-pass f + c, Laplacian -2c); bilateralSobel's
    weights are all 1 (tap gradients differ by < 4 sigma / 0.05), so the
    filtered gradient is the centre Sobel: |gx| = |gy| = (8b + 16cu)/sqrt(4.5),
    glen = (16/3)(b + 2cu); whiteness = f / (f + 2c)."""
    w = h = 28
    f0, b, c, lv = 0.2, 0.005, 2.5e-4, 1e-4
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    u = xx + yy
    f = f0 + b * u + c * u * u
    for bayer in (0, 1, 2, 3):
        g = o.interpolate_green(f, bayer, lv)
        site = site_map(bayer, w, h)
        nongreen = (site == 0) | (site == 2)
        glen = 16.0 / 3.0 * (b + 2 * c * u)
        sigma = np.sqrt(lv * (f + c))
        lgt = smoothstep(sigma, 4 * sigma, glen)
        whiteness = f / (f + 2 * c)
        hf_gain = lgt * np.minimum(0.5 * whiteness + smoothstep(0.0, 0.3, glen / 1.41421356), 1.0)
        want = np.where(nongreen, f + c - 2 * c * hf_gain, f)
        m = 4  # stencil reach: Sobel 1 + bilateral window 2, Laplacian 2
        assert np.all((hf_gain[m:-m, m:-m] > 0.3) & (hf_gain[m:-m, m:-m] < 0.99)), "gain strictly inside (0, 1)"
        np.testing.assert_allclose(g[m:-m, m:-m], want[m:-m, m:-m], rtol=0, atol=1e-12, err_msg=f"bayer {bayer}")


def test_green_high_frequency_term_closed_form(o64):
    check_green_quadratic_closed_form(o64)


# -------------------------------------------------- red/blue: diagonal alpha
def check_red_blue_diagonal_edges(o):
    """G = 1/2, R = 1/4 everywhere, B steps along x + y = k (constant along the
    anti-diagonal through each pixel) or along x - y = k (constant along the
    main diagonal).  At R sites the missing blue is the mean over the diagonal
    B neighbours that lie on the pixel's own constant line, which the alpha
    mix must select near the edge: the result is the true blue exactly."""
    w, h = 30, 26
    yy, xx = np.mgrid[0:h, 0:w]
    for bayer in (0, 1, 2, 3):
        for anti in (True, False):
            for k in (0, 1):
                t = (xx + yy) if anti else (xx - yy + h)
                B = np.where(t >= (w + h) // 2 + k, 0.625, 0.125)
                rgb = np.stack([np.full((h, w), 0.25), np.full((h, w), 0.5), B], -1)
                raw = mosaic(rgb, bayer)
                out = o.interpolate_red_blue(raw, rgb[..., 1], bayer, 1e-6)
                site = site_map(bayer, w, h)
                r_sites, b_sites = site == 0, site == 2
                m = np.zeros((h, w), bool)
                m[1:-1, 1:-1] = True  # diagonal neighbours inside the image
                np.testing.assert_allclose(out[..., 2][r_sites & m], B[r_sites & m], rtol=0, atol=1e-12,
                                           err_msg=f"bayer {bayer} anti {anti}")
                np.testing.assert_allclose(out[..., 0][b_sites & m], 0.25, rtol=0, atol=1e-12)


def test_red_blue_diagonal_alpha_selects_the_edge_direction(o64):
    check_red_blue_diagonal_edges(o64)


# --------------------------------------- bilateral with gradientBoost > 1
KX = np.array([[1, 0, -1], [2, 0, -2], [1, 0, -1]], np.float64)


def check_bilateral_gradient_boost(o):
    """denoiseImage with gradientBoost 2 against a numpy brute force: the luma
    edge is gB - (gB - 1) smoothstep(0.5, 2, |sobel(Y)| / sigma_Y) (Sobel by
    scipy correlation with clamped borders, divided by sqrt(4.5))."""
    rng = np.random.default_rng(31)
    h, w = 26, 30
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    img = np.zeros((h, w, 4))
    img[..., 0] = 0.3 + 0.2 * (xx > 14) + 0.004 * yy
    img[..., 1] = 0.05 * np.sin(xx / 4.0)
    img[..., 2] = -0.03 + 0.002 * xx
    img[..., :3] += rng.normal(0, 0.01, (h, w, 3))
    va, vb = np.array([1e-5, 2e-5, 2e-5]), np.array([1e-3, 5e-4, 5e-4])
    gb, cb = 2.0, 1.2
    out = o.denoise_bilateral(img, va, vb, cb, gb)
    Y = img[..., 0]
    gx = ndimage.correlate(Y, KX, mode="nearest") / np.sqrt(4.5)
    gy = ndimage.correlate(Y, KX.T, mode="nearest") / np.sqrt(4.5)
    sig = np.sqrt(va + vb * img[..., :1])
    boost = gb - (gb - 1) * smoothstep(0.5, 2.0, np.hypot(gx, gy) / sig[..., 0])
    frac = ((boost > 1.01) & (boost < 1.99)).mean()
    assert 0.05 < frac < 0.95, f"the gradient term spans its ramp on {frac:.2f} of the pixels"
    edge = np.stack([boost, np.full_like(boost, cb), np.full_like(boost, cb)], -1) * np.sqrt(2)
    num = np.zeros((h, w, 3))
    den = np.zeros((h, w, 3))
    for j in range(-2, 3):
        for i in range(-2, 3):
            s = img[np.clip(yy.astype(int) + j, 0, h - 1), np.clip(xx.astype(int) + i, 0, w - 1)][..., :3]
            d = np.linalg.norm((s - img[..., :3]) / sig, axis=-1)
            wgt = (d[..., None] < edge).astype(np.float64)
            num += wgt * s
            den += wgt
    np.testing.assert_allclose(out[..., :3], num / den, rtol=0, atol=1e-12)


def test_bilateral_gradient_boost_brute_force(o64):
    check_bilateral_gradient_boost(o64)


# --------------------------------------------- reassemble with sharpening > 1
def check_reassemble_sharpening(o):
    """n1 = d1 (so the Laplacian add is d0 itself), d1 a ramp of slope beta per
    coarse texel along x or y, d0 = up(d1) + delta.  Interior: up(d1) is the
    ramp at (x + 1/2)/2 - 1/2, the gate's central difference is beta / 2, and
    the result is up(d1) + delta * s' with
    s' = 1 + (s - 1) smoothstep(tau/4, tau, beta/2) (1 - smoothstep(0.95, 1, d))
         (0.6 + 0.4 smoothstep(0, 0.1, d)),  tau = sqrt(A + B up(d1)) / 4."""
    w1, h1 = 16, 12
    w, h = 2 * w1, 2 * h1
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    # sharpening and the noise model pass through float32 (values chosen exact)
    cases = [  # (axis, base, beta, delta, A, B, s)
        ("x", 0.3, 0.01, 0.02, 2.0 ** -20, 0.0, 1.75),        # gate open
        ("y", 0.3, 0.01, 0.02, 2.0 ** -20, 0.0, 1.75),
        ("x", 0.3, 0.01, 0.02, 1.0, 0.0, 1.75),               # flat for the noise level: gate shut
        ("x", 0.3, 0.01, 0.02, 2.0 ** -10, 2.0 ** -8, 2.5),   # gate part open
        ("y", 0.93, 0.004, 0.03, 2.0 ** -20, 0.0, 1.75),      # highlight protection (d crosses 0.95..1)
        ("x", 0.02, 0.004, -0.03, 2.0 ** -20, 0.0, 1.75),     # shadow protection (d crosses 0..0.1)
    ]
    for axis, base, beta, delta, A, B, s in cases:
        jj1, ii1 = np.mgrid[0:h1, 0:w1].astype(np.float64)
        coarse = ii1 if axis == "x" else jj1
        fine = (xx if axis == "x" else yy) / 2.0 - 0.25
        d1 = np.zeros((h1, w1, 4))
        d1[..., 0] = base + beta * coarse
        d1[..., 1] = 0.1
        up = base + beta * fine
        d0 = np.zeros((h, w, 4))
        d0[..., 0] = up + delta
        d0[..., 1:3] = np.random.default_rng(5).uniform(-0.1, 0.1, (h, w, 2))
        out = o.reassemble(d0, d1, d1, s, [A, B])
        tau = 0.25 * np.sqrt(A + B * up)
        d = up + delta
        detail = (smoothstep(0.25 * tau, tau, beta / 2.0) * (1.0 - smoothstep(0.95, 1.0, d))
                  * (0.6 + 0.4 * smoothstep(0.0, 0.1, d)))
        want = up + delta * (1.0 + (s - 1.0) * detail)
        m = slice(2, -2)
        np.testing.assert_allclose(out[m, m, 0], want[m, m], rtol=0, atol=1e-12, err_msg=str((axis, A, base)))
        np.testing.assert_array_equal(out[..., 1:3], d0[..., 1:3])


def test_reassemble_sharpening_gate_closed_form(o64):
    check_reassemble_sharpening(o64)


# ------------------------------------------------------ planted mistakes
# (pin, reference line, original text, mistaken text)
MUTANTS = [
    (check_green_step_edges, "atan2pi arguments swapped", "ATAN2PI(grady, gradx)", "ATAN2PI(gradx, grady)"),
    (check_green_step_edges, "direction mix operands swapped", "mixr(g_est_y, g_est_x, direction)",
     "mixr(g_est_x, g_est_y, direction)"),
    (check_green_step_edges, "Sobel kernel transposed", "{{1, 1}, {0, 2}, {-1, 1}},\n    {{2, 0}, {0, 0}, {-2, 0}},\n"
     "    {{1, -1}, {0, -2}, {-1, -1}},", "{{1, 1}, {2, 0}, {1, -1}},\n    {{0, 2}, {0, 0}, {0, -2}},\n"
     "    {{-1, 1}, {-2, 0}, {-1, -1}},"),
    (check_green_quadratic_closed_form, "Laplacian sign flipped", "(K(2.0) * c_xy - (c_left + c_right)) / K(4.0)",
     "((c_left + c_right) - K(2.0) * c_xy) / K(4.0)"),
    (check_green_quadratic_closed_form, "whiteness inverted",
     "mmin(c_xy, mmin(g_ave, c2_ave)) / mmax(c_xy, mmax(g_ave, c2_ave))",
     "mmax(c_xy, mmax(g_ave, c2_ave)) / mmin(c_xy, mmin(g_ave, c2_ave))"),
    (check_green_quadratic_closed_form, "gradient not scaled by 1/sqrt(2) in the gain",
     "glen / K(1.41421356)", "glen"),
    (check_green_quadratic_closed_form, "Sobel not normalised", "*gx = vx * INV_SQRT4_5;", "*gx = vx;"),
    (check_red_blue_diagonal_edges, "diagonal means swapped in the alpha mix",
     "mixr((c2_tr + c2_bl) / K(2.0), (c2_tl + c2_br) / K(2.0), alpha)",
     "mixr((c2_tl + c2_br) / K(2.0), (c2_tr + c2_bl) / K(2.0), alpha)"),
    (check_red_blue_diagonal_edges, "atan2pi arguments swapped (red/blue)", "ATAN2PI(dvy, dvx)", "ATAN2PI(dvx, dvy)"),
    (check_bilateral_gradient_boost, "gradient not divided by sigma", "len2(vx * is0, vy * is0)", "len2(vx, vy)"),
    (check_bilateral_gradient_boost, "boost ramp reversed", "lumaBoost = gb - (gb - K(1.0)) *",
     "lumaBoost = K(1.0) + (gb - K(1.0)) *"),
    (check_reassemble_sharpening, "sharpening mix operands swapped", "mixr(b1[0], d[0], sharpening)",
     "mixr(d[0], b1[0], sharpening)"),
    (check_reassemble_sharpening, "highlight protection on the coarse value",
     "(K(1.0) - smoothstepr(K(0.95), K(1.0), d[0]))", "(K(1.0) - smoothstepr(K(0.95), K(1.0), b1[0]))"),
    (check_reassemble_sharpening, "threshold without the 1/4", "real threshold = K(0.25) * RSQRT(",
     "real threshold = RSQRT("),
    (check_reassemble_sharpening, "central difference over a whole coarse texel",
     "real dx = (t0[0] - t1[0]) / K(2.0);", "real dx = (t0[0] - t1[0]);"),
]


def _mutant_oracle(tmp_path, idx, old, new):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"mutation {idx}: the original text must occur exactly once"
    path = tmp_path / f"mutant{idx}.c"
    path.write_text(src.replace(old, new))
    so = tmp_path / f"mutant{idx}.so"
    subprocess.run(["gcc", "-O1", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                    "-DORACLE_DOUBLE=1", "-I", os.path.join(ROOT, "include"), str(path), "-o", str(so), "-lm"],
                   check=True, cwd=str(tmp_path))
    o = O.Oracle(double=True)
    o.L = C.CDLL(str(so))
    return o


@pytest.mark.parametrize("idx", range(len(MUTANTS)), ids=[m[1].replace(" ", "_") for m in MUTANTS])
def test_planted_mistakes_fail_the_pins(tmp_path, idx):
    check, _, old, new = MUTANTS[idx]
    o = _mutant_oracle(tmp_path, idx, old, new)
    with pytest.raises(AssertionError):
        check(o)


@pytest.fixture(scope="module")
def o64():
    return O.Oracle(double=True)
