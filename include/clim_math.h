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
CLIM_HD float clim_stepf(float edge, float x) { return x < edge ? 0.0f : 1.0f; }

/* ------------------------------------------------------------ atan2(y,x)/pi
 * Octant reduction to t = min/max in [0,1]; for t > tan(pi/8) the identity
 * atan(t) = pi/4 + atan((t-1)/(t+1)) brings the argument into
 * [-tan(pi/8), tan(pi/8)], where atan(u) = u*P(u^2) with P a degree-5
 * least-squares fit at Chebyshev nodes (max abs error 2.7e-10 before float
 * rounding).  Result in units of pi (OpenCL atan2pi). */
CLIM_HD float clim_atan2pif(float y, float x) {
    float ax = fabsf(x), ay = fabsf(y);
    float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    float t = mx > 0.0f ? mn / mx : 0.0f;
    if (x != x || y != y) return x + y;          /* NaN in, NaN out */
    float base = 0.0f;
    if (t > 4.142135680e-01f) {                  /* tan(pi/8) */
        t = (t - 1.0f) / (t + 1.0f);
        base = 0.25f;
    }
    float s = t * t;
    float p = -6.077074260e-02f;
    p = p * s + 1.059265658e-01f;
    p = p * s + -1.424319446e-01f;
    p = p * s + 1.999843121e-01f;
    p = p * s + -3.333331347e-01f;
    p = p * s + 1.0f;
    float r = base + (t * p) * 3.183098733e-01f; /* 1/pi */
    if (ay > ax) r = 0.5f - r;
    if (clim_f2u(x) >> 31) r = 1.0f - r;
    if (clim_f2u(y) >> 31) r = -r;
    return r;
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

/* --------------------------------------------------------------- tanh(x) */
CLIM_HD float clim_tanhf(float x) {
    float ax = fabsf(x);
    float e = clim_expf(-2.0f * ax);
    float t = (1.0f - e) / (1.0f + e);
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

/* ------------------------------------------ 128-bit fixed-point accumulator
 * Non-negative doubles below 2^37 are truncated to multiples of 2^-90 and
 * summed as unsigned 128-bit integers.  Integer addition is associative, so
 * the sums (and everything derived from them) are identical for any
 * reduction order: GPU block trees, CPU threads or a sequential loop. */
typedef struct { uint64_t hi, lo; } clim_fx128;

#define CLIM_FX_FRAC_BITS 90

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

/* m * 2^s as a 128-bit integer, truncated (s may be negative) */
CLIM_HD clim_fx128 clim_fx_shift(uint64_t m, int s) {
    clim_fx128 r;
    if (s >= 0) {
        if (s >= 128) { r.hi = 0; r.lo = 0; }               /* out of range: not reached for |t| < 2^37 */
        else if (s >= 64) { r.hi = m << (s - 64); r.lo = 0; }
        else if (s == 0) { r.hi = 0; r.lo = m; }
        else { r.hi = m >> (64 - s); r.lo = m << s; }
    } else {
        r.hi = 0;
        r.lo = -s >= 64 ? 0 : m >> (-s);
    }
    return r;
}

/* floor(f * 2^90) for a float f >= 0, by integer arithmetic: equal to
 * clim_fx_from_double((double)f) */
CLIM_HD clim_fx128 clim_fx_from_float(float f) {
    uint32_t b = clim_f2u(f) & 0x7fffffffu;
    int e = (int)(b >> 23);
    uint64_t m = b & 0x7fffffu;
    if ((clim_f2u(f) >> 31) || e >= 127 + 37) { clim_fx128 z = {0, 0}; return z; }
    if (e) m |= 0x800000u; else e = 1;
    return clim_fx_shift(m, e - 150 + CLIM_FX_FRAC_BITS);
}

/* floor(f * g * 2^90) for floats f, g >= 0: the 48-bit product of the
 * significands is exact, so this equals clim_fx_from_double((double)f * g) */
CLIM_HD clim_fx128 clim_fx_from_product(float f, float g) {
    uint32_t bf = clim_f2u(f) & 0x7fffffffu, bg = clim_f2u(g) & 0x7fffffffu;
    int ef = (int)(bf >> 23), eg = (int)(bg >> 23);
    if (((clim_f2u(f) ^ clim_f2u(g)) >> 31 && bf && bg) || ef == 255 || eg == 255 || ef + eg >= 254 + 37) {
        clim_fx128 z = {0, 0};
        return z;
    }
    uint64_t mf = bf & 0x7fffffu, mg = bg & 0x7fffffu;
    if (ef) mf |= 0x800000u; else ef = 1;
    if (eg) mg |= 0x800000u; else eg = 1;
    clim_fx128 r = clim_fx_shift(mf * mg, ef + eg - 300 + CLIM_FX_FRAC_BITS);
    if (r.hi >> 63) { r.hi = 0; r.lo = 0; }  /* product >= 2^37 */
    return r;
}

/* floor(t * 2^90) for a double t >= 0 from its bits: equal to clim_fx_from_double(t) */
CLIM_HD clim_fx128 clim_fx_from_double_bits(double t) {
    uint64_t b;
    __builtin_memcpy(&b, &t, 8);
    int e = (int)((b & 0x7fffffffffffffffull) >> 52);
    if ((b >> 63) || e >= 1023 + 37) { clim_fx128 z = {0, 0}; return z; }
    uint64_t m = b & 0xfffffffffffffull;
    if (e) m |= 0x10000000000000ull; else e = 1;
    return clim_fx_shift(m, e - 1075 + CLIM_FX_FRAC_BITS);
}

/* ---- regression terms as exact integers
 * An in-range mean m lies in [0.001, 0.5]; every float there is a multiple
 * of 2^-33 (0.001 > 2^-10), so m * 2^33 is an exact uint32 and its square an
 * exact uint64.  A variance v in [0, 0.001] is quantised to floor(v * 2^64)
 * (at most 5.4e-20 below v).  All regression sums are then sums of exact
 * integers: s_x = sum(mq) 2^-33, s_xx = sum(mq^2) 2^-66, s_y = sum(vq) 2^-64,
 * s_xy = sum(mq vq) 2^-97. */
CLIM_HD uint32_t clim_q_mean(float m) { return (uint32_t)(m * 8589934592.0f); }

CLIM_HD uint64_t clim_q_var(float v) {
    float x = v * 18446744073709551616.0f;               /* exact: v * 2^64 < 2^55 */
    uint32_t hi = (uint32_t)(x * 2.3283064365386962890625e-10f); /* floor(x / 2^32) */
    uint32_t lo = (uint32_t)(x - (float)hi * 4294967296.0f);      /* exact remainder */
    return ((uint64_t)hi << 32) | lo;
}

CLIM_HD clim_fx128 clim_fx_u64(uint64_t a) {
    clim_fx128 r;
    r.hi = 0;
    r.lo = a;
    return r;
}

/* a * b for a < 2^32, b < 2^64 as 128 bits */
CLIM_HD clim_fx128 clim_fx_mul32x64(uint32_t a, uint64_t b) {
    uint64_t lo = (uint64_t)a * (uint32_t)b;
    uint64_t hp = (uint64_t)a * (uint32_t)(b >> 32);
    clim_fx128 r;
    r.lo = lo + (hp << 32);
    r.hi = (hp >> 32) + (r.lo < lo ? 1u : 0u);
    return r;
}

/* (hi 2^64 + lo) * scale, scale a power of two */
CLIM_HD double clim_fx_to_double_scaled(clim_fx128 a, double scale) {
    return ((double)a.hi * 18446744073709551616.0 + (double)a.lo) * scale;
}

CLIM_HD clim_fx128 clim_fx_add(clim_fx128 a, clim_fx128 b) {
    clim_fx128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
    return r;
}

CLIM_HD double clim_fx_to_double(clim_fx128 a) {
    /* hi*2^64 + lo, scaled by 2^-90; the two parts are converted separately,
     * so the result is the correctly rounded sum of two exact-ish doubles */
    double hi = (double)a.hi * 1.490116119384765625e-08;    /* hi*2^64*2^-90 = hi*2^-26 */
    double lo = (double)a.lo * 8.077935669463160887416100e-28; /* 2^-90 */
    return hi + lo;
}

#endif /* CLIM_MATH_H */
