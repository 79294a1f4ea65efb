/*
 * clim_math.h - scalar primitives shared by the HIP kernels and the CPU oracle.
 *
 * The RAW->RGB path is a chain of threshold decisions (hard sigma filter,
 * noise-model inlier masks, despeckle clamps).  A one-ulp difference in an
 * upstream value flips such a decision somewhere in a 12 MP frame, so the GPU
 * path and the FP32 oracle must perform the same IEEE single-precision
 * operations in the same order.  Basic arithmetic, sqrtf and division are
 * correctly rounded on both targets; transcendental functions are not, so the
 * few the path needs are defined here once, in plain +,-,*,/ arithmetic that
 * both compilers evaluate identically (both sides build with
 * -ffp-contract=off).  Each primitive is checked against the C library in the
 * test suite (error bounds stated there), so sharing them does not make the
 * oracle agree with the kernels by construction on anything but these
 * scalar functions.
 *
 * Also here: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
 * numbers: as easy as 1, 2, 3", SC'11) for the seeded synthetic scenes, and
 * the 128-bit fixed-point accumulator used for order-independent, exactly
 * reproducible noise-model regression sums.
 *
 * Usable from C99, C++ and HIP device code.
 */
#ifndef CLIM_MATH_H
#define CLIM_MATH_H

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define CLIM_HD __host__ __device__ static inline
#else
#define CLIM_HD static inline
#endif

/* ---------------------------------------------------------------- bit casts */

CLIM_HD uint32_t clim_f2u(float f) { uint32_t u; __builtin_memcpy(&u, &f, 4); return u; }
CLIM_HD float clim_u2f(uint32_t u) { float f; __builtin_memcpy(&f, &u, 4); return f; }

/* OpenCL C clamp/smoothstep/mix/step semantics (OpenCL C 2.0 spec, 6.13.2/6.13.4):
 * clamp(x, lo, hi) = fmin(fmax(x, lo), hi), so a NaN x yields lo. */
CLIM_HD float clim_clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
CLIM_HD float clim_smoothstepf(float e0, float e1, float x) {
    float t = clim_clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
CLIM_HD float clim_mixf(float a, float b, float t) { return a + (b - a) * t; }
/* smoothstep(e0, e0 + span, x) with the division by the constant span done
 * as a multiplication by rspan = 1 / span and one fused correction
 * (clim_div25's form).  Equal to clim_smoothstepf for every float x where
 * that holds for the pair (e0, span): the tone map's (0.9, 2.0) and
 * (0.75, 2.0) are checked over every x between the edges (outside them both
 * forms clamp to the same end) and the special values.  An infinite x makes
 * the correction NaN; the product, of the right sign, is kept then. */
CLIM_HD float clim_smoothstep_rc(float e0, float span, float rspan, float x) {
    const float d = x - e0;
    const float q = d * rspan;
    const float c = fmaf(fmaf(-q, span, d), rspan, q);
    const float t = clim_clampf(c == c ? c : q, 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
CLIM_HD float clim_stepf(float edge, float x) { return x < edge ? 0.0f : 1.0f; }

/* ------------------------------------------------------------ atan2(y,x)/pi
 * Octant reduction to t = min/max in [0,1]; past tan(pi/8) the identity
 * atan(t) = pi/4 + atan((t-1)/(t+1)) brings the argument into
 * [-tan(pi/8), tan(pi/8)], where atan(u) = u*P(u^2) with P a degree-5
 * least-squares fit at Chebyshev nodes (max abs error 2.7e-10 before float
 * rounding).  The reduced argument is formed with one division, as
 * min/max or (min - max)/(min + max), the octant chosen by
 * min > tan(pi/8) max; the polynomial in fused steps.  Result in units of
 * pi (OpenCL atan2pi). */
CLIM_HD float clim_atan2pif(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const int far = mn > 4.142135680e-01f * mx; /* tan(pi/8) */
    const float num = far ? mn - mx : mn, den = far ? mn + mx : mx;
    const float t = den > 0.0f ? num / den : 0.0f;
    const float s = t * t;
    float p = -6.077074260e-02f;
    p = fmaf(p, s, 1.059265658e-01f);
    p = fmaf(p, s, -1.424319446e-01f);
    p = fmaf(p, s, 1.999843121e-01f);
    p = fmaf(p, s, -3.333331347e-01f);
    p = fmaf(p, s, 1.0f);
    float r = fmaf(t * p, 3.183098733e-01f, far ? 0.25f : 0.0f); /* 1/pi */
    if (ay > ax) r = 0.5f - r;
    if (clim_f2u(x) >> 31) r = 1.0f - r;
    if (clim_f2u(y) >> 31) r = -r;
    return (x != x || y != y) ? x + y : r; /* NaN in, NaN out */
}

/* ----------------------------------------------------------------- exp(x)
 * Cody-Waite reduction x = k*ln2 + r, |r| <= ln2/2, degree-7 Taylor
 * polynomial for e^r (truncation error < 6e-9 relative), scale by 2^k built
 * from the exponent bits. */
CLIM_HD float clim_expf(float x) {
    if (x != x) return x;
    if (x > 88.0f) return INFINITY;
    if (x < -103.0f) return 0.0f;
    float k = rintf(x * 1.442695022e+00f);
    float r = (x - k * 6.93359375e-01f) - k * -2.12194440e-04f;
    float p = 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    int ki = (int)k;
    /* 2^k in two steps so that k down to -149 stays representable */
    int k1 = ki / 2, k2 = ki - k1;
    float s1 = clim_u2f((uint32_t)(k1 + 127) << 23);
    float s2 = clim_u2f((uint32_t)(k2 + 127) << 23);
    return (p * s1) * s2;
}

/* ----------------------------------------------------------------- log(x)
 * x = m*2^e with m in [sqrt(1/2), sqrt(2)); log(m) = 2*atanh(s), s = (m-1)/(m+1),
 * series to s^11 (|s| <= 0.1716, truncation < 3e-10). */
CLIM_HD float clim_logf(float x) {
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }
    uint32_t b = clim_f2u(x);
    e += (int)((b >> 23) & 0xffu) - 127;
    float m = clim_u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.414213538e+00f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + 1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + 1.0f / 3.0f;
    p = p * s2 + 1.0f;
    float lm = 2.0f * s * p;
    return (float)e * 6.931471825e-01f + lm;
}

/* ---------------------------------------------------------------- pow(x,y)
 * |x|^y = exp(y * log|x|), with the special values of C99 Annex F (which
 * OpenCL's pow follows): pow(x, +-0) = pow(1, y) = 1 even for NaN;
 * pow(-1, +-inf) = 1; a finite negative x gives NaN unless y is an integer,
 * whose parity gives the sign (also for -0 and -inf).  The rounding of
 * y * log|x| dominates the error: relative error <= 4e-6 for
 * |y * log|x|| <= 20, <= 1e-6 for |y * log|x|| <= 4. */
CLIM_HD float clim_powf(float x, float y) {
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x != x || y != y) return x + y;
    const float ay = fabsf(y);
    const int is_int = ay >= 16777216.0f || y == rintf(y);  /* +-inf counts as even */
    const int odd = is_int && ay < 16777216.0f && ((int)ay & 1);
    const int neg = (clim_f2u(x) >> 31) != 0;
    const float ax = fabsf(x);
    if (neg && !is_int && ax != 0.0f && ax != INFINITY) return clim_u2f(0x7fc00000u);
    if (ax == 1.0f) return (neg && odd) ? -1.0f : 1.0f;
    float r = clim_expf(y * clim_logf(ax));
    return (neg && odd) ? -r : r;
}

/* --------------------------------------------------------------- tanh(x)
 * (1 - e) / (1 + e) with e = exp(-2|x|), the exponential inlined without
 * branches: below -20, e < 2^-28 and (1 - e) / (1 + e) rounds to 1 for any
 * such e, so the argument is clamped there (a NaN compares false and stays
 * NaN), which keeps 2^k a normal number (k >= -29).  The exponential is
 * clim_expf's method with the Cody-Waite steps and the Horner steps fused
 * (fmaf: one rounding each, the same on every IEEE machine). */
CLIM_HD float clim_tanhf(float x) {
    float y = -2.0f * fabsf(x);
    y = y < -20.0f ? -20.0f : y;
    const float k = rintf(y * 1.442695022e+00f);
    const float r = fmaf(-k, -2.12194440e-04f, fmaf(-k, 6.93359375e-01f, y));
    float p = 1.0f / 5040.0f;
    p = fmaf(p, r, 1.0f / 720.0f);
    p = fmaf(p, r, 1.0f / 120.0f);
    p = fmaf(p, r, 1.0f / 24.0f);
    p = fmaf(p, r, 1.0f / 6.0f);
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    const float ks = fmaxf(k, -29.0f); /* a NaN k (NaN x) gives a NaN p anyway */
    const float e = p * clim_u2f((uint32_t)((int)ks + 127) << 23);
    const float t = (1.0f - e) / (1.0f + e);
    return (clim_f2u(x) >> 31) ? -t : t;
}

/* -------------------------------------------------------- cos(2*pi*u)
 * Reduction in turns: v = |u - rint(u)| in [0, 1/2], reflect to [0, 1/4],
 * then a Taylor cosine or sine on |arg| <= pi/4 (error < 4e-9). */
CLIM_HD float clim_cos2pif(float u) {
    float v = fabsf(u - rintf(u));
    float sign = 1.0f;
    if (v > 0.25f) { v = 0.5f - v; sign = -1.0f; }
    float res;
    if (v <= 0.125f) {
        float a = v * 6.283185482e+00f;
        float a2 = a * a;
        float p = -1.0f / 3628800.0f;
        p = p * a2 + 1.0f / 40320.0f;
        p = p * a2 + -1.0f / 720.0f;
        p = p * a2 + 1.0f / 24.0f;
        p = p * a2 + -0.5f;
        p = p * a2 + 1.0f;
        res = p;
    } else {
        float a = (0.25f - v) * 6.283185482e+00f;
        float a2 = a * a;
        float p = 1.0f / 362880.0f;
        p = p * a2 + -1.0f / 5040.0f;
        p = p * a2 + 1.0f / 120.0f;
        p = p * a2 + -1.0f / 6.0f;
        p = p * a2 + 1.0f;
        res = a * p;
    }
    return sign * res;
}

/* ------------------------------------------------------- Philox4x32-10 */
typedef struct { uint32_t v[4]; } clim_u32x4;

CLIM_HD void clim_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

CLIM_HD clim_u32x4 clim_philox4x32_10(clim_u32x4 ctr, uint32_t k0, uint32_t k1) {
    for (int round = 0; round < 10; ++round) {
        uint32_t hi0, lo0, hi1, lo1;
        clim_mulhilo32(0xD2511F53u, ctr.v[0], &hi0, &lo0);
        clim_mulhilo32(0xCD9E8D57u, ctr.v[2], &hi1, &lo1);
        clim_u32x4 n;
        n.v[0] = hi1 ^ ctr.v[1] ^ k0;
        n.v[1] = lo1;
        n.v[2] = hi0 ^ ctr.v[3] ^ k1;
        n.v[3] = lo0;
        ctr = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return ctr;
}

/* uniform in [0,1): top 24 bits, exact in float */
CLIM_HD float clim_u01(uint32_t r) { return (float)(r >> 8) * 5.9604644775390625e-08f; }
/* uniform in (0,1]: for logarithms */
CLIM_HD float clim_u01_open0(uint32_t r) { return (float)((r >> 8) + 1u) * 5.9604644775390625e-08f; }

/* ------------------------------------------- relaxed-math stencil forms
 * Two per-tap expressions of the FP32 path in contracted form (the
 * reference's -cl-fast-relaxed-math allows contraction and reassociation);
 * the FP32 oracle and the kernels both evaluate these functions.
 *
 * bilateralSobel (cl/demosaic.cl:114-129) tap weight
 * 1 - smoothstep(e0, e0 + span, 0.05 |dl|): with c1 = 0.05 / span and
 * c0 = -e0 / span, t = clamp(fma(|dl|, c1, c0), 0, 1) and
 * w = fma(-t*t, fma(-2, t, 3), 1).  For e0 = 0 (zero noise), c0 = 0: t is
 * then 1 for |dl| > 0 and 0 for |dl| = 0, as (0.05 |dl| - 0) / 0 clamps. */
CLIM_HD float clim_green_tap_c0(float e0, float inv) { return e0 != 0.0f ? -(e0 * inv) : 0.0f; }

CLIM_HD float clim_green_tap_weight(float adl, float c1, float c0) {
    const float t = clim_clampf(fmaf(adl, c1, c0), 0.0f, 1.0f);
    return fmaf(-(t * t), fmaf(-2.0f, t, 3.0f), 1.0f);
}

/* the reciprocal variance k = 1 / sigma^2 of a noise model value s = a + b Y;
 * NaN for s < 0, where the square-root form's sigma is NaN (so every weight
 * of the pixel is zero, as there) */
CLIM_HD float clim_inv_var(float s) { return s < 0.0f ? clim_u2f(0x7fc00000u) : 1.0f / s; }

/* 1 / sqrt(s) from integer operations, multiplications and fused
 * multiply-adds only, so every IEEE machine computes the same bits: the
 * bit-level estimate 0x5f375a86 - (bits(s) >> 1) (Lomont's constant for the
 * classic estimate) refined by three Newton steps y <- y (3/2 - (s/2) y^2).
 * Within 2.5 ulp of 1/sqrt(s) for s in [2^-124, 2^124] (exhaustive over
 * [1, 4); multiplying s by 4 halves every intermediate exactly, so that
 * binade pair decides the others while no intermediate leaves the normal
 * range), within 4 ulp for every normal s; subnormal s get a coarser
 * estimate.  NaN for s < 0 and NaN, +inf for s = 0, 0 for s = +inf. */
/* the bit-level estimate and the three Newton steps alone: clim_rsqrt for
 * every positive finite s (normal or subnormal) */
CLIM_HD float clim_rsqrt_newton(float s) {
    float y = clim_u2f(0x5f375a86u - (clim_f2u(s) >> 1));
    const float h = 0.5f * s;
    for (int i = 0; i < 3; i++) y = y * fmaf(-h, y * y, 1.5f);
    return y;
}
CLIM_HD float clim_rsqrt(float s) {
    const float y = clim_rsqrt_newton(s);
    if (!(s > 0.0f)) return s == 0.0f ? clim_u2f(0x7f800000u) : clim_u2f(0x7fc00000u);
    return s == clim_u2f(0x7f800000u) ? 0.0f : y;
}

/* denoiseImage (cl/demosaic.cl:735-765) range distance |(s - c) / sigma|^2
 * in the kernels' form: u_c = (s_c - c_c) / sigma_c as one fused step
 * fma(s_c, r_c, -cr_c), with r_c = 1 / sigma_c = clim_inv_sigma(a_c + b_c Y)
 * (clim_rsqrt) and cr_c = c_c r_c per centre pixel, then the squared length
 * in two fused steps.  A zero variance gives r = inf and a negative one NaN:
 * every tap's distance is then NaN or inf (as the literal form's), so no tap
 * has weight. */
CLIM_HD float clim_inv_sigma(float s) { return clim_rsqrt(s); }
CLIM_HD float clim_range_len2(float s0, float s1, float s2, float r0, float r1, float r2, float cr0, float cr1,
                              float cr2) {
    const float u0 = fmaf(s0, r0, -cr0), u1 = fmaf(s1, r1, -cr1), u2 = fmaf(s2, r2, -cr2);
    return fmaf(u2, u2, fmaf(u1, u1, u0 * u0));
}

/* The squared range distance
 * |(s - c) / sigma|^2 as sum d_c^2 k_c, k_c = clim_inv_var(a_c + b_c Y):
 * no square root, two fused steps. */
CLIM_HD float clim_bilateral_dist(float d0, float d1, float d2, float k0, float k1, float k2) {
    const float l = (d0 * k0) * d0;
    return fmaf(d2 * k2, d2, fmaf(d1 * k1, d1, l));
}
/* The same over four channels (denoiseRawRGBA, cl/demosaic.cl:932-948:
 * length((s - c) / sigma) < 1 with sigma^2 = rawVariance * c). */
CLIM_HD float clim_raw_dist4(float d0, float d1, float d2, float d3, float k0, float k1, float k2, float k3) {
    return fmaf(d3 * k3, d3, clim_bilateral_dist(d0, d1, d2, k0, k1, k2));
}

/* --------------------------------------------------- v / 65535 for uint16 v
 * The correctly rounded quotient from one reciprocal multiplication and one
 * fused correction step: q = v * fl(1/65535), e = fma(-q, 65535, v) (exact),
 * q + e * fl(1/65535) rounded once.  Equal to the IEEE division v / 65535.0f
 * for every v in [0, 65535] (checked exhaustively). */
CLIM_HD float clim_div65535(uint32_t v) {
    const float x = (float)v, r = 1.0f / 65535.0f;
    const float q = x * r;
    return fmaf(fmaf(-q, 65535.0f, x), r, q);
}

/* ---------------------------------------------------------------- x / 25
 * One reciprocal multiplication and one fused correction (zeros passed
 * through with their sign): equal to the IEEE division x / 25.0f for every
 * finite float x (checked over all 2^32 bit patterns).  Not a general
 * method: the same form differs from x / 1.41421356f for millions of x. */
CLIM_HD float clim_div25(float x) {
    const float r = 1.0f / 25.0f;
    const float q = x * r;
    const float c = fmaf(fmaf(-q, 25.0f, x), r, q);
    return x == 0.0f ? q : c;
}

/* The same form without the zero test: equal to x / 25.0f for every finite
 * x except -0, for which it gives +0.  Under round-to-nearest a sum that
 * starts from +0 is never -0, nor is a square or the difference of two such
 * values, so the 5x5 window statistics (whose sums start from +0) use it. */
CLIM_HD float clim_div25_nz(float x) {
    const float r = 1.0f / 25.0f;
    const float q = x * r;
    return fmaf(fmaf(-q, 25.0f, x), r, q);
}

/* ------------------------------------------- x / n for a weight count n
 * The bilateral's weighted mean divides a sum by its count of unit weights,
 * n in 0..25, with rn = 1.0f / n (rn = +inf for n = 0).  Where the product
 * q = x * rn is a normal float, one fused correction gives the correctly
 * rounded quotient; where q is 0, inf or NaN, q itself is x / n (n = 0
 * included: 0 * inf = NaN = 0 / 0); where q is subnormal the IEEE division
 * is taken.  Equal to x / (float)n bit for bit for every float x and every
 * n in 0..25 (checked over all 2^32 bit patterns per n; the test suite
 * rechecks one binade and the subnormal-quotient range, which decide it).
 * Without the subnormal branch the form differs for n = 6, 10, 12, 14, ...
 * when x < 2^-125.  clim_div_count_normal is the fused form alone (equal to
 * clim_div_count wherever x * rn is a normal float). */
CLIM_HD float clim_div_count_normal(float x, float n, float rn) {
    const float q = x * rn;
    return fmaf(fmaf(-q, n, x), rn, q);
}
CLIM_HD float clim_div_count(float x, float n, float rn) {
    const float q = x * rn;
    const float c = clim_div_count_normal(x, n, rn);
    const float a = fabsf(q);
    if (a < 1.17549435e-38f && a > 0.0f) return x / n;
    return a < INFINITY && a > 0.0f ? c : q;
}

/* ----------------------------------------- the unit step (l < e0) scaled
 * The bilateral's weight (l < e0 ? 1 : 0) as one fused step whose output
 * clamp to [0, 1] the GPU applies for free: clamp(fma(l, -H, e0 H), 0, 1)
 * with H = 2^(50 - e), e the exponent of e0 (a positive normal float of
 * exponent >= -77), so e0 H is exact.  For l < e0 the exact value
 * H (e0 - l) >= H ulp(e0) / 2 = 2^26 rounds to >= 1; for l >= e0 it is
 * <= 0; a NaN l clamps to 0 (fmaxf(NaN, 0) = 0 here, the GPU's clamp with
 * DX10 clamping likewise).  The kernels evaluate this form; the oracle keeps
 * the comparison, and the test suite checks the two agree. */
CLIM_HD float clim_step_below_scaled(float l, float e0) {
    const float H = clim_u2f((304u - (clim_f2u(e0) >> 23)) << 23);
    return clim_clampf(fmaf(l, -H, e0 * H), 0.0f, 1.0f);
}

/* ------------------------------------------ 128-bit fixed-point accumulator
 * Non-negative doubles below 2^37 are truncated to multiples of 2^-90 and
 * summed as unsigned 128-bit integers.  Integer addition is associative, so
 * the sums (and everything derived from them) are identical for any
 * reduction order: GPU block trees, CPU threads or a sequential loop. */
typedef struct { uint64_t hi, lo; } clim_fx128;

CLIM_HD clim_fx128 clim_fx_from_double(double t) {
    clim_fx128 r;
    /* NaN, non-positive and >= 2^37 (never a regression term) count as 0 */
    if (!(t > 0.0) || !(t < 137438953472.0)) { r.hi = 0; r.lo = 0; return r; }
    double y = t * 1237940039285380274899124224.0;          /* 2^90 */
    double h = floor(y * 5.421010862427522170037264e-20);  /* 2^-64 */
    r.hi = (uint64_t)h;
    r.lo = (uint64_t)(y - h * 18446744073709551616.0);      /* exact */
    return r;
}

CLIM_HD clim_fx128 clim_fx_u64(uint64_t a) {
    clim_fx128 r;
    r.hi = 0;
    r.lo = a;
    return r;
}

CLIM_HD clim_fx128 clim_fx_add(clim_fx128 a, clim_fx128 b) {
    clim_fx128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
    return r;
}

/* (hi 2^64 + lo) * scale, scale a power of two */
CLIM_HD double clim_fx_to_double_scaled(clim_fx128 a, double scale) {
    return ((double)a.hi * 18446744073709551616.0 + (double)a.lo) * scale;
}

CLIM_HD double clim_fx_to_double(clim_fx128 a) {
    /* hi*2^64 + lo, scaled by 2^-90; the two parts are converted separately,
     * so the result is the correctly rounded sum of two exact-ish doubles */
    double hi = (double)a.hi * 1.490116119384765625e-08;    /* hi*2^64*2^-90 = hi*2^-26 */
    double lo = (double)a.lo * 8.077935669463160887416100e-28; /* 2^-90 */
    return hi + lo;
}

/* ------------------------------------- noise-model regression, FP32 path
 * The least-squares fits of the noise model (var = A + B * mean) take their
 * samples on a fixed grid: an in-range mean m in [0.001, 0.5] as
 * mq = floor(m * 2^32) <= 2^31 and a variance v in [0, 0.001] as
 * vq = floor(v * 2^41) < 2^32 (grid errors below 2.4e-10 and 4.6e-13, far
 * below the FP32 error of the statistics themselves).  Every regression term
 * (mq, vq, mq^2, mq vq, vq^2) is then an integer below 2^63 and every sum an
 * exact 128-bit integer, whatever the summation order. */
CLIM_HD uint32_t clim_q32_mean(float m) { return (uint32_t)(m * 4294967296.0f); }
CLIM_HD uint32_t clim_q32_var(float v) { return (uint32_t)(v * 2199023255552.0f); }

/* the exact sums of one channel */
typedef struct { clim_fx128 sx, sy, sxx, sxy, svv; } clim_qsums;

/* B = max((N s_xy - s_x s_y) / (N s_xx - s_x^2), 1e-8), A = max((s_y - B s_x) / N, 1e-8)
 * with s_x = sum(m), s_y = sum(v), s_xx = sum(m^2), s_xy = sum(m v) of the grid values */
CLIM_HD void clim_fit_solve_q(const clim_qsums* q, unsigned long long n, double* A, double* B) {
    const double N = (double)n;
    const double s_x = clim_fx_to_double_scaled(q->sx, 0x1p-32);
    const double s_y = clim_fx_to_double_scaled(q->sy, 0x1p-41);
    const double s_xx = clim_fx_to_double_scaled(q->sxx, 0x1p-64);
    const double s_xy = clim_fx_to_double_scaled(q->sxy, 0x1p-73);
    const double b = fmax((N * s_xy - s_x * s_y) / (N * s_xx - s_x * s_x), 1e-8);
    *B = b;
    *A = fmax((s_y - b * s_x) / N, 1e-8);
}

/* sum over the samples of (A + B m - v)^2 of the grid values, expanded into
 * the exact moment sums:  N A^2 + B^2 s_xx + s_vv + 2AB s_x - 2A s_y - 2B s_xy
 * (evaluated in this order in FP64, clamped at 0) */
CLIM_HD double clim_fit_rss_q(const clim_qsums* q, unsigned long long n, double A, double B) {
    const double s_x = clim_fx_to_double_scaled(q->sx, 0x1p-32);
    const double s_y = clim_fx_to_double_scaled(q->sy, 0x1p-41);
    const double s_xx = clim_fx_to_double_scaled(q->sxx, 0x1p-64);
    const double s_xy = clim_fx_to_double_scaled(q->sxy, 0x1p-73);
    const double s_vv = clim_fx_to_double_scaled(q->svv, 0x1p-82);
    const double a2 = 2.0 * A, b2 = 2.0 * B;
    double r = (double)n * (A * A);
    r = r + (B * B) * s_xx;
    r = r + s_vv;
    r = r + (a2 * B) * s_x;
    r = r - a2 * s_y;
    r = r - b2 * s_xy;
    return r > 0.0 ? r : 0.0;
}

#endif /* CLIM_MATH_H */
