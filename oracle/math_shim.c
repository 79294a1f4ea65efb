/*
 * math_shim.c - exports the scalar primitives of include/clim_math.h over
 * arrays, so the tests can check them against numpy / the C library.
 * TEST INFRASTRUCTURE ONLY.
 */
#include <stdint.h>

#include "clim_math.h"

void om_atan2pi(const float* y, const float* x, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_atan2pif(y[i], x[i]);
}
void om_exp(const float* x, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_expf(x[i]);
}
void om_log(const float* x, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_logf(x[i]);
}
void om_pow(const float* x, const float* y, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_powf(x[i], y[i]);
}
void om_tanh(const float* x, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_tanhf(x[i]);
}
void om_cos2pi(const float* x, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_cos2pif(x[i]);
}
void om_smoothstep(float e0, float e1, const float* x, float* out, long n) {
    for (long i = 0; i < n; i++) out[i] = clim_smoothstepf(e0, e1, x[i]);
}
void om_philox(const uint32_t* ctr, uint32_t k0, uint32_t k1, uint32_t* out) {
    clim_u32x4 c = {{ctr[0], ctr[1], ctr[2], ctr[3]}};
    clim_u32x4 r = clim_philox4x32_10(c, k0, k1);
    for (int i = 0; i < 4; i++) out[i] = r.v[i];
}
/* out[2*i] = hi, out[2*i+1] = lo of the fixed-point value of t[i] */
void om_fx_from_double(const double* t, uint64_t* out, long n) {
    for (long i = 0; i < n; i++) {
        clim_fx128 f = clim_fx_from_double(t[i]);
        out[2 * i] = f.hi;
        out[2 * i + 1] = f.lo;
    }
}
/* sum of the fixed-point values of t[0..n) in order, and its double value */
double om_fx_sum(const double* t, long n, uint64_t* hi, uint64_t* lo) {
    clim_fx128 s = {0, 0};
    for (long i = 0; i < n; i++) s = clim_fx_add(s, clim_fx_from_double(t[i]));
    *hi = s.hi;
    *lo = s.lo;
    return clim_fx_to_double(s);
}

/* the grid terms of one (mean, variance) pair: out[2*i ..] = mq, vq */
void om_q32_terms(const float* m, const float* v, uint32_t* out, long n) {
    for (long i = 0; i < n; i++) {
        out[2 * i] = clim_q32_mean(m[i]);
        out[2 * i + 1] = clim_q32_var(v[i]);
    }
}

/* clim_fit_solve_q and clim_fit_rss_q of sums given as (hi, lo) pairs in the
 * order sx, sy, sxx, sxy, svv; out = A, B, rss */
void om_fit_q(const uint64_t* sums, unsigned long long n, double* out) {
    clim_qsums q;
    clim_fx128* f[5] = {&q.sx, &q.sy, &q.sxx, &q.sxy, &q.svv};
    for (int k = 0; k < 5; k++) { f[k]->hi = sums[2 * k]; f[k]->lo = sums[2 * k + 1]; }
    clim_fit_solve_q(&q, n, &out[0], &out[1]);
    out[2] = clim_fit_rss_q(&q, n, out[0], out[1]);
}
/* out[v] = 1 when clim_div65535(v) and (float)v / 65535.0f have the same bits */
void om_div65535_check(uint8_t* out) {
    for (uint32_t v = 0; v <= 65535; v++) {
        const float a = clim_div65535(v), b = (float)v / 65535.0f;
        out[v] = clim_f2u(a) == clim_f2u(b);
    }
}
/* mismatches between clim_div25(x) and x / 25.0f over the bit patterns
 * [lo, hi) (finite x only) */
long om_div25_mismatches(uint32_t lo, uint32_t hi) {
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float x = clim_u2f((uint32_t)i);
        if (!isfinite(x)) continue;
        if (clim_f2u(clim_div25(x)) != clim_f2u(x / 25.0f)) bad++;
    }
    return bad;
}

/* mismatches between clim_div25_nz(x) and x / 25.0f over the bit patterns
 * [lo, hi) (finite x other than -0) */
long om_div25nz_mismatches(uint32_t lo, uint32_t hi) {
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float x = clim_u2f((uint32_t)i);
        if (!isfinite(x) || (uint32_t)i == 0x80000000u) continue;
        if (clim_f2u(clim_div25_nz(x)) != clim_f2u(x / 25.0f)) bad++;
    }
    return bad;
}
/* mismatches between clim_smoothstep_rc(e0, e1 - e0, 1 / (e1 - e0), x) and
 * clim_smoothstepf(e0, e1, x) over the bit patterns [lo, hi) (NaN results
 * equal when both are NaN; a zero's sign is not compared) */
long om_smoothstep_rc_mismatches(float e0, float e1, uint32_t lo, uint32_t hi) {
    const float span = e1 - e0, rspan = 1.0f / span;
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float x = clim_u2f((uint32_t)i);
        const float a = clim_smoothstep_rc(e0, span, rspan, x), b = clim_smoothstepf(e0, e1, x);
        if (a != a || b != b) {
            if (!(a != a && b != b)) bad++;
        } else if (a != b) {
            bad++;
        }
    }
    return bad;
}
/* mismatches between clim_div_count(x, n, 1 / n) and x / n over the bit
 * patterns [lo, hi) (NaN results equal when both are NaN) */
long om_div_count_mismatches(int n, uint32_t lo, uint32_t hi) {
    const float fn = (float)n, rn = 1.0f / fn;
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float x = clim_u2f((uint32_t)i);
        const float a = clim_div_count(x, fn, rn), b = x / fn;
        if (a != a || b != b) {
            if (!(a != a && b != b)) bad++;
        } else if (clim_f2u(a) != clim_f2u(b)) {
            bad++;
        }
    }
    return bad;
}
/* the same count for clim_div_count without its subnormal branch (a planted
 * mistake the test must catch) */
long om_div_count_nobranch_mismatches(int n, uint32_t lo, uint32_t hi) {
    const float fn = (float)n, rn = 1.0f / fn;
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float x = clim_u2f((uint32_t)i);
        const float q = x * rn;
        const float c = fmaf(fmaf(-q, fn, x), rn, q);
        const float a = fabsf(q) < INFINITY && q != 0.0f ? c : q, b = x / fn;
        if (a != a || b != b) {
            if (!(a != a && b != b)) bad++;
        } else if (clim_f2u(a) != clim_f2u(b)) {
            bad++;
        }
    }
    return bad;
}
/* mismatches between clim_step_below_scaled(l, e0) and (l < e0 ? 1 : 0)
 * over the bit patterns [lo, hi) of l (both signs' NaNs and infinities
 * included when the range holds them) */
long om_step_scaled_mismatches(float e0, uint32_t lo, uint32_t hi) {
    long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float l = clim_u2f((uint32_t)i);
        const float w = clim_step_below_scaled(l, e0);
        if (clim_f2u(w) != clim_f2u(l < e0 ? 1.0f : 0.0f)) bad++;
    }
    return bad;
}
/* clim_div25_nz at one value (its -0 case) */
float om_div25_nz(float x) { return clim_div25_nz(x); }

/* the largest |clim_rsqrt(s) - 1/sqrt(s)| / (1/sqrt(s)) over the bit patterns
 * [lo, hi) (positive finite floats), the exact value in double precision */
double om_rsqrt_max_rel_err(uint32_t lo, uint32_t hi) {
    double worst = 0.0;
#pragma omp parallel for reduction(max : worst) schedule(static)
    for (int64_t i = lo; i < (int64_t)hi; i++) {
        const float s = clim_u2f((uint32_t)i);
        const double ex = 1.0 / sqrt((double)s);
        const double e = fabs((double)clim_rsqrt(s) - ex) / ex;
        if (e > worst) worst = e;
    }
    return worst;
}

float om_rsqrt(float s) { return clim_rsqrt(s); }
