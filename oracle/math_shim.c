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

/* the integer conversions must equal clim_fx_from_double on the exact value:
 * out[4*i .. 4*i+3] = (hi, lo) of fx_from_float(f[i]) and fx_from_double(f[i]) */
void om_fx_float_check(const float* f, uint64_t* out, long n) {
    for (long i = 0; i < n; i++) {
        clim_fx128 a = clim_fx_from_float(f[i]), b = clim_fx_from_double((double)f[i]);
        out[4 * i] = a.hi; out[4 * i + 1] = a.lo; out[4 * i + 2] = b.hi; out[4 * i + 3] = b.lo;
    }
}
void om_fx_product_check(const float* f, const float* g, uint64_t* out, long n) {
    for (long i = 0; i < n; i++) {
        clim_fx128 a = clim_fx_from_product(f[i], g[i]), b = clim_fx_from_double((double)f[i] * (double)g[i]);
        out[4 * i] = a.hi; out[4 * i + 1] = a.lo; out[4 * i + 2] = b.hi; out[4 * i + 3] = b.lo;
    }
}
void om_fx_double_bits_check(const double* t, uint64_t* out, long n) {
    for (long i = 0; i < n; i++) {
        clim_fx128 a = clim_fx_from_double_bits(t[i]), b = clim_fx_from_double(t[i]);
        out[4 * i] = a.hi; out[4 * i + 1] = a.lo; out[4 * i + 2] = b.hi; out[4 * i + 3] = b.lo;
    }
}
