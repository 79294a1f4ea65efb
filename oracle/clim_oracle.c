/*
 * clim_oracle.c - CPU restatement of the reference RAW->RGB path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  Nothing in the product links against it.
 *
 * Compiled twice (see Makefile):
 *   ORACLE_DOUBLE=0 -> o32_*  FP32 storage and arithmetic, the "reference
 *                             implementation" the GPU path is compared with
 *                             (the reference's USE_FP16_FLOATS=0 build);
 *   ORACLE_DOUBLE=1 -> o64_*  FP64 storage and arithmetic, C library
 *                             transcendentals: the FP64 oracle errors are
 *                             reported against.
 *
 * Each function cites the reference kernel it restates (cl/ = CLImage/app/
 * src/main/cpp/OpenCL/, cpp/ = CLImage/app/src/main/cpp/).  Conventions
 * pinned here that the reference leaves to the OpenCL implementation:
 *  - sampler-less reads clamp coordinates to the image (clamp-to-edge);
 *  - CL_FILTER_LINEAR: unnormalised u = x*W, i0 = floor(u-0.5), weight
 *    frac(u-0.5) rounded to 8 fractional bits (texture-unit precision),
 *    clamp-to-edge;
 *  - -cl-single-precision-constant: every literal is a float literal in the
 *    FP32 build;
 *  - -cl-fast-relaxed-math transcendentals: atan2pi, tanh and exp come from
 *    include/clim_math.h in the FP32 build (the HIP kernels use the same
 *    primitives, each checked against the C library by the tests) and from
 *    libm in the FP64 build;
 *  - native_powr(t, 0.5) == sqrt(t), NaN for t < 0;
 *  - the despeckle stage computes in FP32 (the reference uses half);
 *  - the FP64 regression sums are accumulated exactly (128-bit fixed point,
 *    2^-90 resolution), so their value does not depend on summation order.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "clim_math.h"

#if ORACLE_DOUBLE
typedef double real;
#define K(x) x
#define P(name) o64_##name
#define RSQRT(x) sqrt(x)
#define RABS(x) fabs(x)
#define RFMIN(a, b) fmin(a, b)
#define RFMAX(a, b) fmax(a, b)
#define RFLOOR(x) floor(x)
#define ATAN2PI(y, x) (atan2(y, x) / 3.14159265358979323846)
#define TANH(x) tanh(x)
#define EXP(x) exp(x)
#define POW(x, y) pow(x, y)
#define INV_SQRT4_5 0.47140452079103168
#else
typedef float real;
#define K(x) x##f
#define P(name) o32_##name
#define RSQRT(x) sqrtf(x)
#define RABS(x) fabsf(x)
#define RFMIN(a, b) fminf(a, b)
#define RFMAX(a, b) fmaxf(a, b)
#define RFLOOR(x) floorf(x)
#define ATAN2PI(y, x) clim_atan2pif(y, x)
#define TANH(x) clim_tanhf(x)
#define EXP(x) clim_expf(x)
#define POW(x, y) clim_powf(x, y)
#define INV_SQRT4_5 0.471404521f
#endif

/* OpenCL built-ins (OpenCL C 2.0, 6.13.2 / 6.13.4) */
static inline real clampr(real x, real lo, real hi) { return RFMIN(RFMAX(x, lo), hi); }
static inline real smoothstepr(real e0, real e1, real x) {
    real t = clampr((x - e0) / (e1 - e0), K(0.0), K(1.0));
    return t * t * (K(3.0) - K(2.0) * t);
}
static inline real mixr(real a, real b, real t) { return a + (b - a) * t; }
/* smoothstep with the division by (e1 - e0) given as a reciprocal `inv` */
static inline real smoothstep_inv(real e0, real inv, real x) {
    real t = clampr((x - e0) * inv, K(0.0), K(1.0));
    return t * t * (K(3.0) - K(2.0) * t);
}
/* the kernel file's own min/max macros, cl/demosaic.cl:57-63 */
static inline real mmin(real a, real b) { return a < b ? a : b; }
static inline real mmax(real a, real b) { return a > b ? a : b; }
static inline real len2(real x, real y) { return RSQRT(x * x + y * y); }
static inline real len3(real x, real y, real z) { return RSQRT(x * x + y * y + z * z); }

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* sampler-less read_imagef, clamp-to-edge */
static inline real rd1(const real* img, int w, int h, int x, int y) {
    return img[(size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)];
}
static inline const real* rd4(const real* img, int w, int h, int x, int y) {
    return img + ((size_t)clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)) * 4;
}

/* OpenCL Bayer table, cl/demosaic.cl:29-34, entries (x, y) for
 * raw_red, raw_green, raw_blue, raw_green2 */
static const int kBayerOffsets[4][4][2] = {
    {{1, 0}, {0, 0}, {0, 1}, {1, 1}}, /* grbg */
    {{0, 1}, {0, 0}, {1, 0}, {1, 1}}, /* gbrg */
    {{0, 0}, {0, 1}, {1, 1}, {1, 0}}, /* rggb */
    {{1, 1}, {0, 1}, {0, 0}, {1, 0}}  /* bggr */
};
enum { raw_red = 0, raw_green = 1, raw_blue = 2, raw_green2 = 3 };

/* ------------------------------------------------------------ scaleRawData
 * cl/demosaic.cl:69-78: max(scaleMul[c] * (raw/65535 - black), 0) per site. */
void P(scale_raw)(const uint16_t* raw, int w, int h, int bayer, const float scale_mul[4],
                  float black_level, real* out) {
#pragma omp parallel for schedule(static)
    for (int qy = 0; qy < h / 2; qy++) {
        for (int qx = 0; qx < w / 2; qx++) {
            for (int c = 0; c < 4; c++) {
                int x = 2 * qx + kBayerOffsets[bayer][c][0];
                int y = 2 * qy + kBayerOffsets[bayer][c][1];
                real v = (real)raw[(size_t)y * w + x] / K(65535.0);
                real s = (real)scale_mul[c] * (v - (real)black_level);
                out[(size_t)y * w + x] = mmax(s, K(0.0));
            }
        }
    }
}

/* 5x5 window sums, evaluated as the vertical sum of five horizontal 5-tap
 * sums: S = ((((H(-2) + H(-1)) + H(0)) + H(1)) + H(2)) with
 * H(j) = ((((x(j,-2) + x(j,-1)) + x(j,0)) + x(j,1)) + x(j,2)).  The
 * reference writes one sequential 25-term loop; -cl-fast-relaxed-math leaves
 * the association to the compiler, and this one lets a GPU share the
 * horizontal sums between the five rows that use them. */

/* ------------------------------------------------------ rawNoiseStatistics
 * cl/demosaic.cl:1288-1324: 5x5-quad mean and variance of (R, G, B, G2). */
void P(raw_noise_statistics)(const real* raw, int w, int h, int bayer, real* mean, real* var) {
    const int qw = w / 2, qh = h / 2;
#pragma omp parallel for schedule(static)
    for (int qy = 0; qy < qh; qy++) {
        for (int qx = 0; qx < qw; qx++) {
            real sum[4] = {0, 0, 0, 0}, sumSq[4] = {0, 0, 0, 0};
            for (int j = -2; j <= 2; j++) {
                real hs[4] = {0, 0, 0, 0}, hq[4] = {0, 0, 0, 0};
                for (int i = -2; i <= 2; i++) {
                    for (int c = 0; c < 4; c++) {
                        real s = rd1(raw, w, h, 2 * (qx + i) + kBayerOffsets[bayer][c][0],
                                     2 * (qy + j) + kBayerOffsets[bayer][c][1]);
                        hs[c] += s;
                        hq[c] += s * s;
                    }
                }
                for (int c = 0; c < 4; c++) {
                    sum[c] += hs[c];
                    sumSq[c] += hq[c];
                }
            }
            real* m = mean + ((size_t)qy * qw + qx) * 4;
            real* v = var + ((size_t)qy * qw + qx) * 4;
            for (int c = 0; c < 4; c++) {
                m[c] = sum[c] / K(25.0);
                v[c] = (sumSq[c] - (sum[c] * sum[c]) / K(25.0)) / K(25.0);
            }
        }
    }
}

/* ------------------------------------------------------ noise model fit
 * cpp/pyramidal_denoise.cpp:161-256 (levels, scalar mean, 3 variances) and
 * :259-362 (raw, 4 means, 4 variances).  Sample i has means
 * m[i*m_step + c] (c < m_nch; m_nch == 1 means one shared mean) and
 * variances v[i*v_step + c] (c < nch).  Samples with every mean in
 * [0.001, 0.5] and every variance in [0, 0.001] enter an FP64 least-squares
 * fit var = A + B*mean per channel (B and A floored at 1e-8); the RMS error
 * of that fit selects inliers (squared residual <= RMS in every channel)
 * for a second fit.  Returns 0 on success, 1 if no sample is in range
 * (out untouched), 2 if the refit had no inliers (first fit returned).
 * out = A[0..nch), B[0..nch).
 * Sums are exact.  The FP64 build follows the reference's three passes
 * (sums, squared residuals, sums over inliers) on the 2^-90 fixed-point values
 * of the FP64 terms.  The FP32 build (the definition the HIP kernels follow)
 * takes the sums over the grid values of clim_q32_mean / clim_q32_var, adds
 * the sum of squared variances, and gets the sum of squared residuals from
 * those moments (clim_fit_rss_q) instead of a second pass over the samples. */
#define FIT_MAXCH 4

/* loads sample i; returns 1 when every mean is in [0.001, 0.5] and every
 * variance in [0, 0.001] (inRange, cpp/pyramidal_denoise.cpp:151-159) */
static inline int fit_sample(int nch, int m_nch, const real* m, int m_step, const real* v, int v_step,
                             long i, double* mm, double* vv) {
    int ok = 1;
    for (int c = 0; c < nch; c++) {
        mm[c] = (double)m[i * m_step + (m_nch == 1 ? 0 : c)];
        vv[c] = (double)v[i * v_step + c];
        if (mm[c] < 0.001 || mm[c] > 0.5) ok = 0;
        if (vv[c] < 0 || vv[c] > 0.001) ok = 0;
    }
    return ok;
}

/* 1 when (A + B*m - v)^2 <= err2 in every channel (cpp/pyramidal_denoise.cpp:324-343) */
static inline int fit_inlier(int nch, const double* mm, const double* vv, const double* A, const double* B,
                             const double* err2) {
    for (int c = 0; c < nch; c++) {
        double diff = (A[c] + B[c] * mm[c]) - vv[c];
        if (!(diff * diff <= err2[c])) return 0;
    }
    return 1;
}

#if ORACLE_DOUBLE
/* sums (s_x, s_y, s_xx, s_xy) per channel over in-range samples; with
 * err2 != NULL only over inliers */
static long fit_sums(int nch, int m_nch, const real* m, int m_step, const real* v, int v_step, long count,
                     const double* A, const double* B, const double* err2, clim_fx128* acc) {
    long n = 0;
    for (int k = 0; k < 4 * FIT_MAXCH; k++) acc[k].hi = acc[k].lo = 0;
#pragma omp parallel
    {
        clim_fx128 l[4 * FIT_MAXCH];
        long ln = 0;
        for (int k = 0; k < 4 * FIT_MAXCH; k++) l[k].hi = l[k].lo = 0;
#pragma omp for schedule(static)
        for (long i = 0; i < count; i++) {
            double mm[FIT_MAXCH], vv[FIT_MAXCH];
            if (!fit_sample(nch, m_nch, m, m_step, v, v_step, i, mm, vv)) continue;
            if (err2 && !fit_inlier(nch, mm, vv, A, B, err2)) continue;
            for (int c = 0; c < nch; c++) {
                l[4 * c + 0] = clim_fx_add(l[4 * c + 0], clim_fx_from_double(mm[c]));
                l[4 * c + 1] = clim_fx_add(l[4 * c + 1], clim_fx_from_double(vv[c]));
                l[4 * c + 2] = clim_fx_add(l[4 * c + 2], clim_fx_from_double(mm[c] * mm[c]));
                l[4 * c + 3] = clim_fx_add(l[4 * c + 3], clim_fx_from_double(mm[c] * vv[c]));
            }
            ln++;
        }
#pragma omp critical
        {
            for (int k = 0; k < 4 * FIT_MAXCH; k++) acc[k] = clim_fx_add(acc[k], l[k]);
            n += ln;
        }
    }
    return n;
}

/* sum of squared residuals over in-range samples (cpp/pyramidal_denoise.cpp:300-311) */
static void fit_residuals(int nch, int m_nch, const real* m, int m_step, const real* v, int v_step,
                          long count, const double* A, const double* B, clim_fx128* acc) {
    for (int k = 0; k < FIT_MAXCH; k++) acc[k].hi = acc[k].lo = 0;
#pragma omp parallel
    {
        clim_fx128 l[FIT_MAXCH];
        for (int k = 0; k < FIT_MAXCH; k++) l[k].hi = l[k].lo = 0;
#pragma omp for schedule(static)
        for (long i = 0; i < count; i++) {
            double mm[FIT_MAXCH], vv[FIT_MAXCH];
            if (!fit_sample(nch, m_nch, m, m_step, v, v_step, i, mm, vv)) continue;
            for (int c = 0; c < nch; c++) {
                double diff = (A[c] + B[c] * mm[c]) - vv[c];
                l[c] = clim_fx_add(l[c], clim_fx_from_double(diff * diff));
            }
        }
#pragma omp critical
        {
            for (int k = 0; k < FIT_MAXCH; k++) acc[k] = clim_fx_add(acc[k], l[k]);
        }
    }
}

/* nlfB = max((N s_xy - s_x s_y) / (N s_xx - s_x^2), 1e-8), nlfA = max((s_y - B s_x) / N, 1e-8)
 * (cpp/pyramidal_denoise.cpp:296-297) */
static void fit_solve(int nch, long n, const clim_fx128* acc, double* A, double* B) {
    double N = (double)n;
    for (int c = 0; c < nch; c++) {
        double s_x = clim_fx_to_double(acc[4 * c + 0]);
        double s_y = clim_fx_to_double(acc[4 * c + 1]);
        double s_xx = clim_fx_to_double(acc[4 * c + 2]);
        double s_xy = clim_fx_to_double(acc[4 * c + 3]);
        B[c] = fmax((N * s_xy - s_x * s_y) / (N * s_xx - s_x * s_x), 1e-8);
        A[c] = fmax((s_y - B[c] * s_x) / N, 1e-8);
    }
}

int P(nlf_fit)(int nch, int m_nch, const real* m, int m_step, const real* v, int v_step, long count,
               float* out) {
    double A[FIT_MAXCH], B[FIT_MAXCH], err2[FIT_MAXCH];
    clim_fx128 acc[4 * FIT_MAXCH];
    long n0 = fit_sums(nch, m_nch, m, m_step, v, v_step, count, NULL, NULL, NULL, acc);
    if (n0 == 0) return 1;
    fit_solve(nch, n0, acc, A, B);
    fit_residuals(nch, m_nch, m, m_step, v, v_step, count, A, B, acc);
    for (int c = 0; c < nch; c++) err2[c] = sqrt(clim_fx_to_double(acc[c]) / (double)n0);
    long n1 = fit_sums(nch, m_nch, m, m_step, v, v_step, count, A, B, err2, acc);
    int rc = 2;
    if (n1 > 0) {
        fit_solve(nch, n1, acc, A, B);
        rc = 0;
    }
    for (int c = 0; c < nch; c++) { out[c] = (float)A[c]; out[nch + c] = (float)B[c]; }
    return rc;
}
#else
/* exact sums of the grid terms per channel over in-range samples; with
 * err2 != NULL only over inliers (the inlier test uses the float values) */
static long fit_qsums(int nch, int m_nch, const real* m, int m_step, const real* v, int v_step, long count,
                      const double* A, const double* B, const double* err2, clim_qsums* acc) {
    long n = 0;
    const clim_fx128 z = {0, 0};
    for (int c = 0; c < FIT_MAXCH; c++) acc[c].sx = acc[c].sy = acc[c].sxx = acc[c].sxy = acc[c].svv = z;
#pragma omp parallel
    {
        clim_qsums l[FIT_MAXCH];
        long ln = 0;
        for (int c = 0; c < FIT_MAXCH; c++) l[c].sx = l[c].sy = l[c].sxx = l[c].sxy = l[c].svv = z;
#pragma omp for schedule(static)
        for (long i = 0; i < count; i++) {
            double mm[FIT_MAXCH], vv[FIT_MAXCH];
            if (!fit_sample(nch, m_nch, m, m_step, v, v_step, i, mm, vv)) continue;
            if (err2 && !fit_inlier(nch, mm, vv, A, B, err2)) continue;
            for (int c = 0; c < nch; c++) {
                const uint64_t mq = clim_q32_mean((float)mm[c]), vq = clim_q32_var((float)vv[c]);
                l[c].sx = clim_fx_add(l[c].sx, clim_fx_u64(mq));
                l[c].sy = clim_fx_add(l[c].sy, clim_fx_u64(vq));
                l[c].sxx = clim_fx_add(l[c].sxx, clim_fx_u64(mq * mq));
                l[c].sxy = clim_fx_add(l[c].sxy, clim_fx_u64(mq * vq));
                l[c].svv = clim_fx_add(l[c].svv, clim_fx_u64(vq * vq));
            }
            ln++;
        }
#pragma omp critical
        {
            for (int c = 0; c < FIT_MAXCH; c++) {
                acc[c].sx = clim_fx_add(acc[c].sx, l[c].sx);
                acc[c].sy = clim_fx_add(acc[c].sy, l[c].sy);
                acc[c].sxx = clim_fx_add(acc[c].sxx, l[c].sxx);
                acc[c].sxy = clim_fx_add(acc[c].sxy, l[c].sxy);
                acc[c].svv = clim_fx_add(acc[c].svv, l[c].svv);
            }
            n += ln;
        }
    }
    return n;
}

int P(nlf_fit)(int nch, int m_nch, const real* m, int m_step, const real* v, int v_step, long count,
               float* out) {
    double A[FIT_MAXCH], B[FIT_MAXCH], err2[FIT_MAXCH];
    clim_qsums acc[FIT_MAXCH];
    long n0 = fit_qsums(nch, m_nch, m, m_step, v, v_step, count, NULL, NULL, NULL, acc);
    if (n0 == 0) return 1;
    for (int c = 0; c < nch; c++) {
        clim_fit_solve_q(&acc[c], (unsigned long long)n0, &A[c], &B[c]);
        err2[c] = sqrt(clim_fit_rss_q(&acc[c], (unsigned long long)n0, A[c], B[c]) / (double)n0);
    }
    long n1 = fit_qsums(nch, m_nch, m, m_step, v, v_step, count, A, B, err2, acc);
    int rc = 2;
    if (n1 > 0) {
        for (int c = 0; c < nch; c++) clim_fit_solve_q(&acc[c], (unsigned long long)n1, &A[c], &B[c]);
        rc = 0;
    }
    for (int c = 0; c < nch; c++) { out[c] = (float)A[c]; out[nch + c] = (float)B[c]; }
    return rc;
}
#endif

/* --------------------------------------------------------- interpolateGreen
 * cl/demosaic.cl:96-129 (sobel, bilateralSobel) and :160-227. */
static const real kSobel[3][3][2] = {
    {{1, 1}, {0, 2}, {-1, 1}},
    {{2, 0}, {0, 0}, {-2, 0}},
    {{1, -1}, {0, -2}, {-1, -1}},
};

static inline void sobel(const real* img, int w, int h, int x, int y, real* gx, real* gy) {
    real vx = 0, vy = 0;
    for (int j = -1; j <= 1; j++) {
        for (int i = -1; i <= 1; i++) {
            real s = rd1(img, w, h, x + i, y + j);
            vx += kSobel[j + 1][i + 1][0] * s;
            vy += kSobel[j + 1][i + 1][1] * s;
        }
    }
    /* value / sqrt(4.5), as a multiplication by the reciprocal */
    *gx = vx * INV_SQRT4_5;
    *gy = vy * INV_SQRT4_5;
}

static inline void bilateral_sobel(const real* img, int w, int h, int x, int y, real sigma,
                                   real* ox, real* oy) {
    real sx = 0, sy = 0, wx = 0, wy = 0;
    real cx, cy;
    sobel(img, w, h, x, y, &cx, &cy);
    cx = RABS(cx);
    cy = RABS(cy);
    /* smoothstep(4 sigma, 16 sigma, d) with the division by the (per-pixel)
     * edge distance hoisted out of the 5x5 loop as a reciprocal */
    const real e0 = K(4.0) * sigma, inv = K(1.0) / (K(16.0) * sigma - K(4.0) * sigma);
#if !ORACLE_DOUBLE
    /* the tap weight by clim_green_tap_weight, the sums in fused form */
    const real c1 = K(0.05) * inv, c0 = clim_green_tap_c0(e0, inv);
#endif
    for (int j = -2; j <= 2; j++) {
        for (int i = -2; i <= 2; i++) {
            real gx, gy;
            sobel(img, w, h, x + i, y + j, &gx, &gy);
            gx = RABS(gx);
            gy = RABS(gy);
#if ORACLE_DOUBLE
            real d = K(0.05) * RABS(len2(gx, gy) - len2(cx, cy));
            real wt = K(1.0) - smoothstep_inv(e0, inv, d);
            sx += wt * gx;
            sy += wt * gy;
#else
            real wt = clim_green_tap_weight(RABS(len2(gx, gy) - len2(cx, cy)), c1, c0);
            sx = fmaf(wt, gx, sx);
            sy = fmaf(wt, gy, sy);
#endif
            wx += wt;
            wy += wt;
        }
    }
#if ORACLE_DOUBLE
    *ox = sx / wx;
    *oy = sy / wy;
#else
    /* the two quotients by one reciprocal (the weight sums are equal) */
    const real iw = K(1.0) / wx;
    *ox = sx * iw;
    *oy = sy * iw;
#endif
}

void P(interpolate_green)(const real* raw, int w, int h, int bayer, float luma_variance, real* green) {
    real lumaVariance = mmax((real)luma_variance, K(1e-4));
    const int gxo = kBayerOffsets[bayer][raw_green][0];
    const int gyo = kBayerOffsets[bayer][raw_green][1];
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            int x0 = (y & 1) == (gyo & 1) ? gxo + 1 : gxo;
            if ((x0 & 1) != (x & 1)) {
                green[(size_t)y * w + x] = rd1(raw, w, h, x, y);
                continue;
            }
            real g_left = rd1(raw, w, h, x - 1, y);
            real g_right = rd1(raw, w, h, x + 1, y);
            real g_up = rd1(raw, w, h, x, y - 1);
            real g_down = rd1(raw, w, h, x, y + 1);
            real c_xy = rd1(raw, w, h, x, y);
            real c_left = rd1(raw, w, h, x - 2, y);
            real c_right = rd1(raw, w, h, x + 2, y);
            real c_up = rd1(raw, w, h, x, y - 2);
            real c_down = rd1(raw, w, h, x, y + 2);
            real c2_tl = rd1(raw, w, h, x - 1, y - 1);
            real c2_tr = rd1(raw, w, h, x + 1, y - 1);
            real c2_bl = rd1(raw, w, h, x - 1, y + 1);
            real c2_br = rd1(raw, w, h, x + 1, y + 1);
            real c2_ave = (c2_tl + c2_tr + c2_bl + c2_br) / K(4.0);

            real g_ave = (g_left + g_right + g_up + g_down) / K(4.0);
            real lumaStdDev = RSQRT(lumaVariance * g_ave);
            real gradx, grady;
            bilateral_sobel(raw, w, h, x, y, lumaStdDev, &gradx, &grady);

            real g_lf_x = (g_left + g_right) / K(2.0), g_lf_y = (g_up + g_down) / K(2.0);
            real g_hf_x = (K(2.0) * c_xy - (c_left + c_right)) / K(4.0);
            real g_hf_y = (K(2.0) * c_xy - (c_up + c_down)) / K(4.0);

            real whiteness = clampr(mmin(c_xy, mmin(g_ave, c2_ave)) / mmax(c_xy, mmax(g_ave, c2_ave)),
                                    K(0.0), K(1.0));
            real glen = len2(gradx, grady);
            real lgt = smoothstepr(lumaStdDev, K(4.0) * lumaStdDev, glen);
#if ORACLE_DOUBLE
            real glen_s2 = glen / K(1.41421356);
#else
            real glen_s2 = glen * K(0.707106781); /* / sqrt(2) as a multiplication */
#endif
            real hf_gain = lgt * mmin(K(0.5) * whiteness + smoothstepr(K(0.0), K(0.3), glen_s2), K(1.0));
            real g_est_x = g_lf_x + hf_gain * g_hf_x;
            real g_est_y = g_lf_y + hf_gain * g_hf_y;

            real direction = K(2.0) * ATAN2PI(grady, gradx);
            direction = direction < K(0.5)
                            ? mixr(direction, K(0.0),
                                   K(1.0) - smoothstepr(K(0.3) * lgt, K(0.45) * lgt, direction))
                            : mixr(direction, K(1.0),
                                   smoothstepr((K(1.0) - K(0.45)) * lgt, (K(1.0) - K(0.3)) * lgt, direction));
            direction = mixr(K(0.5), direction, lgt);
            green[(size_t)y * w + x] = mixr(g_est_y, g_est_x, direction);
        }
    }
}

/* ------------------------------------------------------- interpolateRedBlue
 * cl/demosaic.cl:229-318.  rgb is RGBA (alpha 0).  rotate_180 writes to
 * (W - x, H - y) exactly as the reference does (texels outside the image are
 * dropped, row/column 0 are not written). */
void P(interpolate_red_blue)(const real* raw, const real* green, int w, int h, int bayer,
                             float chroma_variance, int rotate_180, real* rgb) {
    real chromaVariance = mmax((real)chroma_variance, K(1e-6));
    const int rx = kBayerOffsets[bayer][raw_red][0] & 1, ry = kBayerOffsets[bayer][raw_red][1] & 1;
    const int gx = kBayerOffsets[bayer][raw_green][0] & 1, gy = kBayerOffsets[bayer][raw_green][1] & 1;
    const int bx = kBayerOffsets[bayer][raw_blue][0] & 1, by = kBayerOffsets[bayer][raw_blue][1] & 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            int color = (rx == (x & 1) && ry == (y & 1))   ? raw_red
                        : (gx == (x & 1) && gy == (y & 1)) ? raw_green
                        : (bx == (x & 1) && by == (y & 1)) ? raw_blue
                                                           : raw_green2;
            real g = rd1(green, w, h, x, y);
            real red, blue;
            if (color == raw_red || color == raw_blue) {
                real c1 = rd1(raw, w, h, x, y);
                real g_tl = rd1(green, w, h, x - 1, y - 1);
                real g_tr = rd1(green, w, h, x + 1, y - 1);
                real g_bl = rd1(green, w, h, x - 1, y + 1);
                real g_br = rd1(green, w, h, x + 1, y + 1);
                real c2_tl = g_tl - rd1(raw, w, h, x - 1, y - 1);
                real c2_tr = g_tr - rd1(raw, w, h, x + 1, y - 1);
                real c2_bl = g_bl - rd1(raw, w, h, x - 1, y + 1);
                real c2_br = g_br - rd1(raw, w, h, x + 1, y + 1);
                real chromaStdDev = RSQRT(chromaVariance * g);
                real dvx = RABS(c2_tl - c2_br), dvy = RABS(c2_tr - c2_bl);
                real flatness = K(1.0) - smoothstepr(K(0.25) * chromaStdDev, chromaStdDev, len2(dvx, dvy));
                real alpha = mixr(K(2.0) * ATAN2PI(dvy, dvx), K(0.5), flatness);
                real c2 = g - mixr((c2_tr + c2_bl) / K(2.0), (c2_tl + c2_br) / K(2.0), alpha);
                if (color == raw_red) { red = c1; blue = c2; } else { blue = c1; red = c2; }
            } else {
                real g_left = rd1(green, w, h, x - 1, y);
                real g_right = rd1(green, w, h, x + 1, y);
                real g_up = rd1(green, w, h, x, y - 1);
                real g_down = rd1(green, w, h, x, y + 1);
                real c1_left = g_left - rd1(raw, w, h, x - 1, y);
                real c1_right = g_right - rd1(raw, w, h, x + 1, y);
                real c2_up = g_up - rd1(raw, w, h, x, y - 1);
                real c2_down = g_down - rd1(raw, w, h, x, y + 1);
                real c1 = g - (c1_left + c1_right) / K(2.0);
                real c2 = g - (c2_up + c2_down) / K(2.0);
                int redIsC1 = color == ((bayer == 3 || bayer == 0) ? raw_green : raw_green2);
                if (redIsC1) { red = c1; blue = c2; } else { blue = c1; red = c2; }
            }
            int ox = x, oy = y;
            if (rotate_180) {
                ox = w - x;
                oy = h - y;
                if (ox >= w || oy >= h) continue;
            }
            real* o = rgb + ((size_t)oy * w + ox) * 4;
            o[0] = red; o[1] = g; o[2] = blue; o[3] = 0;
        }
    }
}

/* ---------------------------------------------------- blendHighlightsImage
 * cl/demosaic.cl:338-380 (IPHONE_LENS_SHADING is false). In-place safe. */
void P(blend_highlights)(const real* in, int w, int h, float clip_f, real* out) {
    const real s3 = K(1.7320508);
    const real trans[3][3] = {{1, 1, 1}, {s3, -s3, 0}, {-1, -1, 2}};
    const real itrans[3][3] = {{1, s3 / K(2.0), K(-0.5)}, {1, -s3 / K(2.0), K(-0.5)}, {1, 0, 1}};
    const real clip = (real)clip_f;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)w * h; i++) {
        const real* p = in + i * 4;
        real px = p[0], py = p[1], pz = p[2];
        if (px > clip || py > clip || pz > clip) {
            real cam[2][3] = {{px, py, pz}, {mmin(px, clip), mmin(py, clip), mmin(pz, clip)}};
            real lab[2][3], sum[2];
            for (int k = 0; k < 2; k++) {
                for (int r = 0; r < 3; r++)
                    lab[k][r] = trans[r][0] * cam[k][0] + trans[r][1] * cam[k][1] + trans[r][2] * cam[k][2];
                sum[k] = lab[k][1] * lab[k][1] + lab[k][2] * lab[k][2];
            }
            real chratio = sum[0] > 0 ? RSQRT(sum[1] / sum[0]) : K(1.0);
            lab[0][1] *= chratio;
            lab[0][2] *= chratio;
            real q[3];
            for (int r = 0; r < 3; r++)
                q[r] = (itrans[r][0] * lab[0][0] + itrans[r][1] * lab[0][1] + itrans[r][2] * lab[0][2]) / K(3.0);
            px = q[0]; py = q[1]; pz = q[2];
        }
        real* o = out + i * 4;
        o[0] = px; o[1] = py; o[2] = pz; o[3] = 0;
    }
}

/* ------------------------------------------------------------ transformImage
 * cl/demosaic.cl:728-733, row-major 3x3.  In-place safe. */
void P(transform)(const real* in, int w, int h, const float m[9], real* out) {
    real M[9];
    for (int k = 0; k < 9; k++) M[k] = (real)m[k];
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)w * h; i++) {
        real x = in[i * 4 + 0], y = in[i * 4 + 1], z = in[i * 4 + 2];
        out[i * 4 + 0] = M[0] * x + M[1] * y + M[2] * z;
        out[i * 4 + 1] = M[3] * x + M[4] * y + M[5] * z;
        out[i * 4 + 2] = M[6] * x + M[7] * y + M[8] * z;
        out[i * 4 + 3] = 0;
    }
}

/* -------------------------------------------- despeckleLumaMedianChromaImage
 * cl/demosaic.cl:384-413 (forgetful-selection median network) and
 * :443-479 (luma second-extreme tracker), evaluated in FP32/FP64. */
#define SWAPMM(a, b) do { real _t = a; a = mmin(a, b); b = mmax(_t, b); } while (0)
#define MINMAX6(a0, a1, a2, a3, a4, a5) SWAPMM(a0, a1); SWAPMM(a2, a3); SWAPMM(a4, a5); SWAPMM(a0, a2); SWAPMM(a1, a3); SWAPMM(a0, a4); SWAPMM(a3, a5);
#define MINMAX5(a0, a1, a2, a3, a4) SWAPMM(a0, a1); SWAPMM(a2, a3); SWAPMM(a0, a2); SWAPMM(a1, a3); SWAPMM(a0, a4); SWAPMM(a3, a4);
#define MINMAX4(a0, a1, a2, a3) SWAPMM(a0, a1); SWAPMM(a2, a3); SWAPMM(a0, a2); SWAPMM(a1, a3);
#define MINMAX3(a0, a1, a2) SWAPMM(a0, a1); SWAPMM(a0, a2); SWAPMM(a1, a2);

void P(despeckle)(const real* in, int w, int h, const float var_a[3], const float var_b[3], real* out) {
    /* read order of fast_median3x3 */
    static const int ord[9][2] = {{0, -1}, {1, -1}, {0, 0}, {1, 0}, {0, 1}, {1, 1}, {-1, 1}, {-1, 0}, {-1, -1}};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            real sample = 0, firstMax = 0, secMax = 0;
            real firstMin = K(65535.0), secMin = K(65535.0);
            real cb[9], cr[9];
            for (int k = 0; k < 9; k++) {
                const real* p = rd4(in, w, h, x + ord[k][0], y + ord[k][1]);
                real v = p[0];
                secMax = v <= firstMax && v > secMax ? v : secMax;
                secMax = v > firstMax ? firstMax : secMax;
                firstMax = v > firstMax ? v : firstMax;
                secMin = v >= firstMin && v < secMin ? v : secMin;
                secMin = v < firstMin ? firstMin : secMin;
                firstMin = v < firstMin ? v : firstMin;
                if (ord[k][0] == 0 && ord[k][1] == 0) sample = v;
                cb[k] = p[1];
                cr[k] = p[2];
            }
            real med[2];
            for (int c = 0; c < 2; c++) {
                real* q = c == 0 ? cb : cr;
                real a0 = q[0], a1 = q[1], a2 = q[2], a3 = q[3], a4 = q[4], a5 = q[5];
                MINMAX6(a0, a1, a2, a3, a4, a5);
                a0 = q[6];
                MINMAX5(a0, a1, a2, a3, a4);
                a0 = q[7];
                MINMAX4(a0, a1, a2, a3);
                a0 = q[8];
                MINMAX3(a0, a1, a2);
                med[c] = a1;
            }
            real sigma = RSQRT((real)var_a[0] + (real)var_b[0] * sample);
            real minVal = mixr(secMin, firstMin, smoothstepr(sigma, K(4.0) * sigma, secMin - firstMin));
            real maxVal = mixr(secMax, firstMax, smoothstepr(sigma, K(4.0) * sigma, firstMax - secMax));
            sample = clampr(sample, minVal, maxVal);
            real* o = out + ((size_t)y * w + x) * 4;
            o[0] = sample; o[1] = med[0]; o[2] = med[1]; o[3] = 0;
        }
    }
}

/* ------------------------------------------------ linear sampler (RGBA)
 * read_imagef(img, CLK_NORMALIZED_COORDS_TRUE | CLK_ADDRESS_CLAMP_TO_EDGE |
 * CLK_FILTER_LINEAR, (u, v)) with the pinned 8-bit weight precision. */
static inline void bilinear4(const real* img, int w, int h, real un, real vn, real out[4]) {
    real u = un * (real)w - K(0.5);
    real v = vn * (real)h - K(0.5);
    real fu = RFLOOR(u), fv = RFLOOR(v);
    real a = u - fu, b = v - fv;
    a = RFLOOR(a * K(256.0) + K(0.5)) * K(0.00390625);
    b = RFLOOR(b * K(256.0) + K(0.5)) * K(0.00390625);
    int i0 = (int)fu, j0 = (int)fv;
    const real* t00 = rd4(img, w, h, i0, j0);
    const real* t10 = rd4(img, w, h, i0 + 1, j0);
    const real* t01 = rd4(img, w, h, i0, j0 + 1);
    const real* t11 = rd4(img, w, h, i0 + 1, j0 + 1);
    real w00 = (K(1.0) - a) * (K(1.0) - b), w10 = a * (K(1.0) - b);
    real w01 = (K(1.0) - a) * b, w11 = a * b;
    for (int c = 0; c < 4; c++) out[c] = w00 * t00[c] + w10 * t10[c] + w01 * t01[c] + w11 * t11[c];
}

/* ------------------------------------------------------------ downsampleImage
 * cl/demosaic.cl:853-865 via resampleImage, cpp/demosaic_cl.cpp:157-178. */
void P(downsample)(const real* in, int iw, int ih, real* out, int ow, int oh) {
    const real nx = K(1.0) / (real)ow, ny = K(1.0) / (real)oh;
    const real sx = K(0.5) * nx, sy = K(0.5) * ny;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < oh; y++) {
        for (int x = 0; x < ow; x++) {
            real px = ((real)x + K(0.5)) * nx, py = ((real)y + K(0.5)) * ny;
            real acc[4], t[4];
            bilinear4(in, iw, ih, px - sx, py - sy, acc);
            bilinear4(in, iw, ih, px + sx, py - sy, t);
            for (int c = 0; c < 4; c++) acc[c] += t[c];
            bilinear4(in, iw, ih, px - sx, py + sy, t);
            for (int c = 0; c < 4; c++) acc[c] += t[c];
            bilinear4(in, iw, ih, px + sx, py + sy, t);
            for (int c = 0; c < 4; c++) acc[c] += t[c];
            real* o = out + ((size_t)y * ow + x) * 4;
            for (int c = 0; c < 4; c++) o[c] = K(0.25) * acc[c];
        }
    }
}

/* ------------------------------------------------------------ noiseStatistics
 * cl/demosaic.cl:1267-1286: (mean Y, var Y, var Cb, var Cr) over 5x5. */
void P(noise_statistics)(const real* in, int w, int h, real* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            real sum[3] = {0, 0, 0}, sumSq[3] = {0, 0, 0};
            for (int j = -2; j <= 2; j++) {
                real hs[3] = {0, 0, 0}, hq[3] = {0, 0, 0};
                for (int i = -2; i <= 2; i++) {
                    const real* p = rd4(in, w, h, x + i, y + j);
                    for (int c = 0; c < 3; c++) {
                        hs[c] += p[c];
                        hq[c] += p[c] * p[c];
                    }
                }
                for (int c = 0; c < 3; c++) {
                    sum[c] += hs[c];
                    sumSq[c] += hq[c];
                }
            }
            real* o = out + ((size_t)y * w + x) * 4;
            o[0] = sum[0] / K(25.0);
            for (int c = 0; c < 3; c++) o[1 + c] = (sumSq[c] - (sum[c] * sum[c]) / K(25.0)) / K(25.0);
        }
    }
}

/* ----------------------------------------------------- learned strength map
 * The learned-strength network's per-level multipliers (no reference
 * counterpart; climage_amd/csrc/kernels.h strength_at): a map of 5 planes
 * of gw x gh floats (the grid of pyramid level 3), plane l level l's
 * multipliers, sampled bilinearly (clamped) at the level texel's centre, grid
 * coordinate (x + 1/2) 2^(l - 3) - 1/2, in FP32 in both builds.  The
 * bilateral range threshold (squared) is multiplied by m^2, the guided
 * epsilon by m.  Level l reads the map only when open[l] != 0. */
typedef struct {
    const float* map;
    int gw, gh;
    const int* open;
} o_strength;

#ifndef O_STRENGTH_AT
#define O_STRENGTH_AT
static float o_strength_at(const o_strength* s, int level, int x, int y) {
    if (!s || !s->map || !s->open[level]) return 0.0f;
    const float sc = level == 0 ? 0.125f : level == 1 ? 0.25f : level == 2 ? 0.5f : level == 3 ? 1.0f : 2.0f;
    const float gx = ((float)x + 0.5f) * sc - 0.5f, gy = ((float)y + 0.5f) * sc - 0.5f;
    const float fx = floorf(gx), fy = floorf(gy);
    const float ax = gx - fx, ay = gy - fy;
    const int ix = (int)fx, iy = (int)fy;
    const int x0 = ix < 0 ? 0 : (ix > s->gw - 1 ? s->gw - 1 : ix);
    const int x1 = ix + 1 < 0 ? 0 : (ix + 1 > s->gw - 1 ? s->gw - 1 : ix + 1);
    const int y0 = iy < 0 ? 0 : (iy > s->gh - 1 ? s->gh - 1 : iy);
    const int y1 = iy + 1 < 0 ? 0 : (iy + 1 > s->gh - 1 ? s->gh - 1 : iy + 1);
    const float* m = s->map + (size_t)level * s->gw * s->gh;
    const float v00 = m[(size_t)y0 * s->gw + x0], v01 = m[(size_t)y0 * s->gw + x1];
    const float v10 = m[(size_t)y1 * s->gw + x0], v11 = m[(size_t)y1 * s->gw + x1];
    const float top = v00 + (v01 - v00) * ax, bot = v10 + (v11 - v10) * ax;
    return top + (bot - top) * ay;
}
#endif

/* --------------------------------------------------------------- denoiseImage
 * cl/demosaic.cl:735-765: 5x5 hard-threshold sigma filter.  The per-tap
 * division by sigma is a multiplication by 1/sigma (computed once per pixel),
 * and length(d) < edge is evaluated as dot(d, d) < edge * edge. */
static void P(denoise_bilateral_s)(const real* in, int w, int h, const float var_a[3], const float var_b[3],
                                   float chroma_boost, float gradient_boost, const o_strength* sm, int level,
                                   real* out);
void P(denoise_bilateral)(const real* in, int w, int h, const float var_a[3], const float var_b[3],
                          float chroma_boost, float gradient_boost, real* out) {
    P(denoise_bilateral_s)(in, w, h, var_a, var_b, chroma_boost, gradient_boost, NULL, 0, out);
}
static void P(denoise_bilateral_s)(const real* in, int w, int h, const float var_a[3], const float var_b[3],
                                   float chroma_boost, float gradient_boost, const o_strength* sm, int level,
                                   real* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const real* c0 = rd4(in, w, h, x, y);
            real cY = c0[0], cCb = c0[1], cCr = c0[2];
#if ORACLE_DOUBLE
            real sigma[3];
            for (int c = 0; c < 3; c++) sigma[c] = RSQRT((real)var_a[c] + (real)var_b[c] * cY);
#endif
            real lumaBoost = K(1.0);
            if ((real)gradient_boost > K(1.0)) {
                /* sobel() of the x channel (cl/demosaic.cl:745) */
                real vx = 0, vy = 0;
                for (int j = -1; j <= 1; j++)
                    for (int i = -1; i <= 1; i++) {
                        real s = rd4(in, w, h, x + i, y + j)[0];
                        vx += kSobel[j + 1][i + 1][0] * s;
                        vy += kSobel[j + 1][i + 1][1] * s;
                    }
                vx = vx * INV_SQRT4_5;
                vy = vy * INV_SQRT4_5;
                real gb = (real)gradient_boost;
#if ORACLE_DOUBLE
                real is0 = K(1.0) / sigma[0];
#else
                real is0 = K(1.0) / RSQRT((real)var_a[0] + (real)var_b[0] * cY);
#endif
                lumaBoost = gb - (gb - K(1.0)) * smoothstepr(K(0.5), K(2.0), len2(vx * is0, vy * is0));
            }
            real edge[3] = {lumaBoost * K(1.41421356), (real)chroma_boost * K(1.41421356),
                            (real)chroma_boost * K(1.41421356)};
            for (int c = 0; c < 3; c++) edge[c] = edge[c] * edge[c];
            {
                /* learned strength: the range sigma times m */
                const float m = o_strength_at(sm, level, x, y);
                if (m > 0.0f) {
                    const real m2 = (real)m * (real)m;
                    for (int c = 0; c < 3; c++) edge[c] = edge[c] * m2;
                }
            }
#if ORACLE_DOUBLE
            const real isig[3] = {K(1.0) / sigma[0], K(1.0) / sigma[1], K(1.0) / sigma[2]};
#else
            /* r_c = 1 / sigma_c, the distance by clim_range_len2 */
            real rr[3], cr[3];
            const real cc[3] = {cY, cCb, cCr};
            for (int c = 0; c < 3; c++) {
                rr[c] = clim_inv_sigma((real)var_a[c] + (real)var_b[c] * cY);
                cr[c] = cc[c] * rr[c];
            }
#endif
            real fp[3] = {0, 0, 0}, kn[3] = {0, 0, 0};
            for (int j = -2; j <= 2; j++) {
                for (int i = -2; i <= 2; i++) {
                    const real* s = rd4(in, w, h, x + i, y + j);
#if ORACLE_DOUBLE
                    real d0 = (s[0] - cY) * isig[0];
                    real d1 = (s[1] - cCb) * isig[1];
                    real d2 = (s[2] - cCr) * isig[2];
                    real l = d0 * d0 + d1 * d1 + d2 * d2;
#else
                    real l = clim_range_len2(s[0], s[1], s[2], rr[0], rr[1], rr[2], cr[0], cr[1], cr[2]);
#endif
                    for (int c = 0; c < 3; c++) {
                        real wt = K(1.0) - (l < edge[c] ? K(0.0) : K(1.0));
                        fp[c] += wt * s[c];
                        kn[c] += wt;
                    }
                }
            }
            real* o = out + ((size_t)y * w + x) * 4;
            for (int c = 0; c < 3; c++) o[c] = fp[c] / kn[c];
            o[3] = 0;
        }
    }
}

/* ------------------------------------------------------- denoiseImageGuided
 * cl/demosaic.cl:767-851: one-pass 5x5 self-guided colour guided filter. */
static void P(denoise_guided_fast_s)(const real* in, int w, int h, const float var_a[3], const float var_b[3],
                                     const o_strength* sm, int level, real* out);
void P(denoise_guided_fast)(const real* in, int w, int h, const float var_a[3], const float var_b[3],
                            real* out) {
    P(denoise_guided_fast_s)(in, w, h, var_a, var_b, NULL, 0, out);
}
static void P(denoise_guided_fast_s)(const real* in, int w, int h, const float var_a[3], const float var_b[3],
                                     const o_strength* sm, int level, real* out) {
    const real norm = K(1.0) / K(25.0);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const real* c0 = rd4(in, w, h, x, y);
            real ix = c0[0], iy = c0[1], iz = c0[2];
            real eps[3];
            for (int c = 0; c < 3; c++) eps[c] = (real)var_a[c] + (real)var_b[c] * ix;
            {
                /* learned strength: epsilon times m */
                const float m = o_strength_at(sm, level, x, y);
                if (m > 0.0f)
                    for (int c = 0; c < 3; c++) eps[c] = eps[c] * (real)m;
            }
            /* 9 box sums: mean of I and of its 6 products (separable order) */
            real acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            for (int j = -2; j <= 2; j++) {
                real hsum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int i = -2; i <= 2; i++) {
                    const real* s = rd4(in, w, h, x + i, y + j);
                    const real v[9] = {s[0], s[1], s[2], s[0] * s[0], s[0] * s[1], s[0] * s[2],
                                       s[1] * s[1], s[1] * s[2], s[2] * s[2]};
                    for (int k = 0; k < 9; k++) hsum[k] += v[k];
                }
                for (int k = 0; k < 9; k++) acc[k] += hsum[k];
            }
            real mI[3] = {acc[0], acc[1], acc[2]};
            real rr = acc[3], rg = acc[4], rb = acc[5], gg = acc[6], gb = acc[7], bb = acc[8];
            mI[0] *= norm; mI[1] *= norm; mI[2] *= norm;
            rr *= norm; rg *= norm; rb *= norm; gg *= norm; gb *= norm; bb *= norm;
            real vrr = rr - mI[0] * mI[0], vrg = rg - mI[0] * mI[1], vrb = rb - mI[0] * mI[2];
            real vgg = gg - mI[1] * mI[1], vgb = gb - mI[1] * mI[2], vbb = bb - mI[2] * mI[2];
            real vrr_e = vrr + eps[0], vgg_e = vgg + eps[1], vbb_e = vbb + eps[2];
            real invrr = vgg_e * vbb_e - vgb * vgb;
            real invrg = vgb * vrb - vrg * vbb_e;
            real invrb = vrg * vgb - vgg_e * vrb;
            real invgg = vrr_e * vbb_e - vrb * vrb;
            real invgb = vrb * vrg - vrr_e * vgb;
            real invbb = vrr_e * vgg_e - vrg * vrg;
            real invCovDet = K(1.0) / (invrr * vrr_e + invrg * vrg + invrb * vrb);
            invrr *= invCovDet; invrg *= invCovDet; invrb *= invCovDet;
            invgg *= invCovDet; invgb *= invCovDet; invbb *= invCovDet;
            const real cr[3] = {vrr, vrg, vrb}, cg[3] = {vrg, vgg, vgb}, cb[3] = {vrb, vgb, vbb};
            real ar[3], ag[3], ab[3], b[3], q[3];
            for (int c = 0; c < 3; c++) {
                ar[c] = invrr * cr[c] + invrg * cg[c] + invrb * cb[c];
                ag[c] = invrg * cr[c] + invgg * cg[c] + invgb * cb[c];
                ab[c] = invrb * cr[c] + invgb * cg[c] + invbb * cb[c];
            }
            for (int c = 0; c < 3; c++) b[c] = mI[c] - ar[c] * mI[0] - ag[c] * mI[1] - ab[c] * mI[2];
            for (int c = 0; c < 3; c++) q[c] = ar[c] * ix + ag[c] * iy + ab[c] * iz + b[c];
            real* o = out + ((size_t)y * w + x) * 4;
            o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = 0;
        }
    }
}

/* ----------------------------------------------------------- GuidedFilter
 * cpp/guided_filter.cpp:20-225, cl/guided_filter.cl:16-154. */

/* BoxBlur::blur: two transposing boxBlurX passes (cl/guided_filter.cl:16-28);
 * nc channels per texel (4 or 2). */
static void box_blur(const real* in, int w, int h, int nc, int fs, real* out, real* tmp) {
    /* pass 1: rows of in -> tmp (h x w image, tmp[x][y]) */
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            real s[4] = {0, 0, 0, 0};
            for (int d = -fs / 2; d <= fs / 2; d++) {
                const real* p = in + ((size_t)y * w + clampi(x + d, 0, w - 1)) * nc;
                for (int c = 0; c < nc; c++) s[c] += p[c];
            }
            real* t = tmp + ((size_t)x * h + y) * nc;
            for (int c = 0; c < nc; c++) t[c] = s[c] / (real)fs;
        }
    }
    /* pass 2: rows of tmp (width h, height w) -> out transposed back */
#pragma omp parallel for schedule(static)
    for (int x = 0; x < w; x++) {
        for (int y = 0; y < h; y++) {
            real s[4] = {0, 0, 0, 0};
            for (int d = -fs / 2; d <= fs / 2; d++) {
                const real* p = tmp + ((size_t)x * h + clampi(y + d, 0, h - 1)) * nc;
                for (int c = 0; c < nc; c++) s[c] += p[c];
            }
            real* o = out + ((size_t)y * w + x) * nc;
            for (int c = 0; c < nc; c++) o[c] = s[c] / (real)fs;
        }
    }
}

void P(guided_filter)(const real* in, int w, int h, int fs, const float eps_f[3], real* out) {
    size_t n = (size_t)w * h;
    real* mean_I = (real*)malloc(n * 4 * sizeof(real));
    real* s1 = (real*)malloc(n * 4 * sizeof(real));
    real* s2 = (real*)malloc(n * 2 * sizeof(real));
    real* ipr = (real*)malloc(n * 4 * sizeof(real));
    real* ipg = (real*)malloc(n * 4 * sizeof(real));
    real* ipb = (real*)malloc(n * 4 * sizeof(real));
    real* tmp = (real*)malloc(n * 4 * sizeof(real));
    const real eps[3] = {(real)eps_f[0], (real)eps_f[1], (real)eps_f[2]};

    box_blur(in, w, h, 4, fs, mean_I, tmp);
    /* covMatrixProducts, cl/guided_filter.cl:30-44 */
    for (size_t i = 0; i < n; i++) {
        const real* p = in + i * 4;
        s1[i * 4 + 0] = p[0] * p[0]; s1[i * 4 + 1] = p[0] * p[1];
        s1[i * 4 + 2] = p[0] * p[2]; s1[i * 4 + 3] = p[1] * p[1];
        s2[i * 2 + 0] = p[1] * p[2]; s2[i * 2 + 1] = p[2] * p[2];
    }
    box_blur(s1, w, h, 4, fs, s1, tmp);
    box_blur(s2, w, h, 2, fs, s2, tmp);
    /* invSigma, cl/guided_filter.cl:46-83 */
    for (size_t i = 0; i < n; i++) {
        const real* mI = mean_I + i * 4;
        real* p1 = s1 + i * 4;
        real* p2 = s2 + i * 2;
        real vrr = p1[0] - mI[0] * mI[0] + eps[0] * mI[0];
        real vrg = p1[1] - mI[0] * mI[1];
        real vrb = p1[2] - mI[0] * mI[2];
        real vgg = p1[3] - mI[1] * mI[1] + eps[1];
        real vgb = p2[0] - mI[1] * mI[2];
        real vbb = p2[1] - mI[2] * mI[2] + eps[2];
        real invrr = vgg * vbb - vgb * vgb;
        real invrg = vgb * vrb - vrg * vbb;
        real invrb = vrg * vgb - vgg * vrb;
        real invgg = vrr * vbb - vrb * vrb;
        real invgb = vrb * vrg - vrr * vgb;
        real invbb = vrr * vgg - vrg * vrg;
        real covDet = invrr * vrr + invrg * vrg + invrb * vrb;
        p1[0] = invrr / covDet; p1[1] = invrg / covDet; p1[2] = invrb / covDet; p1[3] = invgg / covDet;
        p2[0] = invgb / covDet; p2[1] = invbb / covDet;
    }
    /* meanIpProducts, cl/guided_filter.cl:85-95 */
    for (size_t i = 0; i < n; i++) {
        const real* p = in + i * 4;
        for (int c = 0; c < 3; c++) {
            ipr[i * 4 + c] = p[0] * p[c];
            ipg[i * 4 + c] = p[1] * p[c];
            ipb[i * 4 + c] = p[2] * p[c];
        }
        ipr[i * 4 + 3] = ipg[i * 4 + 3] = ipb[i * 4 + 3] = 0;
    }
    box_blur(ipr, w, h, 4, fs, ipr, tmp);
    box_blur(ipg, w, h, 4, fs, ipg, tmp);
    box_blur(ipb, w, h, 4, fs, ipb, tmp);
    /* computeAb, cl/guided_filter.cl:97-137 */
    for (size_t i = 0; i < n; i++) {
        const real* mp = mean_I + i * 4;
        real cr[3], cg[3], cb[3];
        for (int c = 0; c < 3; c++) {
            cr[c] = ipr[i * 4 + c] - mp[0] * mp[c];
            cg[c] = ipg[i * 4 + c] - mp[1] * mp[c];
            cb[c] = ipb[i * 4 + c] - mp[2] * mp[c];
        }
        real invrr = s1[i * 4 + 0], invrg = s1[i * 4 + 1], invrb = s1[i * 4 + 2], invgg = s1[i * 4 + 3];
        real invgb = s2[i * 2 + 0], invbb = s2[i * 2 + 1];
        real ar[3], ag[3], ab[3], b[3];
        for (int c = 0; c < 3; c++) {
            ar[c] = invrr * cr[c] + invrg * cg[c] + invrb * cb[c];
            ag[c] = invrg * cr[c] + invgg * cg[c] + invgb * cb[c];
            ab[c] = invrb * cr[c] + invgb * cg[c] + invbb * cb[c];
        }
        for (int c = 0; c < 3; c++) b[c] = mp[c] - ar[c] * mp[0] - ag[c] * mp[1] - ab[c] * mp[2];
        for (int c = 0; c < 3; c++) { ipr[i * 4 + c] = ar[c]; ipg[i * 4 + c] = ag[c]; ipb[i * 4 + c] = ab[c]; }
        ipr[i * 4 + 3] = b[0]; ipg[i * 4 + 3] = b[1]; ipb[i * 4 + 3] = b[2];
    }
    box_blur(ipr, w, h, 4, fs, ipr, tmp);
    box_blur(ipg, w, h, 4, fs, ipg, tmp);
    box_blur(ipb, w, h, 4, fs, ipb, tmp);
    /* computeResult, cl/guided_filter.cl:139-154 */
    for (size_t i = 0; i < n; i++) {
        const real* p = in + i * 4;
        const real* r = ipr + i * 4;
        const real* g = ipg + i * 4;
        const real* b = ipb + i * 4;
        real q[3];
        for (int c = 0; c < 3; c++) q[c] = r[c] * p[0] + g[c] * p[1] + b[c] * p[2];
        q[0] += r[3]; q[1] += g[3]; q[2] += b[3];
        out[i * 4 + 0] = q[0]; out[i * 4 + 1] = q[1]; out[i * 4 + 2] = q[2]; out[i * 4 + 3] = 0;
    }
    free(mean_I); free(s1); free(s2); free(ipr); free(ipg); free(ipb); free(tmp);
}

/* ------------------------------------------------------------ reassembleImage
 * cl/demosaic.cl:867-896.  out may alias d0. */
void P(reassemble)(const real* d0, const real* n1, const real* d1, int w, int h, int w1, int h1,
                   float sharpening_f, const float nlf[2], real* out) {
    const real nx = K(1.0) / (real)w, ny = K(1.0) / (real)h;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            real px = ((real)x + K(0.5)) * nx, py = ((real)y + K(0.5)) * ny;
            const real* p0 = d0 + ((size_t)y * w + x) * 4;
            real a1[4], b1[4], d[4];
            bilinear4(n1, w1, h1, px, py, a1);
            bilinear4(d1, w1, h1, px, py, b1);
            for (int c = 0; c < 4; c++) d[c] = p0[c] - (a1[c] - b1[c]);
            real sharpening = (real)sharpening_f;
            if (sharpening > K(1.0)) {
                real t0[4], t1[4];
                bilinear4(d1, w1, h1, px + nx, py, t0);
                bilinear4(d1, w1, h1, px - nx, py, t1);
                real dx = (t0[0] - t1[0]) / K(2.0);
                bilinear4(d1, w1, h1, px, py + ny, t0);
                bilinear4(d1, w1, h1, px, py - ny, t1);
                real dy = (t0[0] - t1[0]) / K(2.0);
                real threshold = K(0.25) * RSQRT((real)nlf[0] + (real)nlf[1] * b1[0]);
                real detail = smoothstepr(K(0.25) * threshold, threshold, len2(dx, dy)) *
                              (K(1.0) - smoothstepr(K(0.95), K(1.0), d[0])) *
                              (K(0.6) + K(0.4) * smoothstepr(K(0.0), K(0.1), d[0]));
                sharpening = K(1.0) + (sharpening - K(1.0)) * detail;
            }
            d[0] = mixr(b1[0], d[0], sharpening);
            real* o = out + ((size_t)y * w + x) * 4;
            for (int c = 0; c < 4; c++) o[c] = d[c];
        }
    }
}

/* ------------------------------------------------------ localToneMappingMask
 * cl/demosaic.cl:1024-1107 (localToneMappingMaskImage, localToneMappingMask),
 * cpp/demosaic_cl.cpp:339-370.  in: the denoised YCbCr image (w x h RGBA);
 * guide: an RGBA image (pyramid level 3 in the pipeline,
 * cpp/raw_converter.cpp:164-167), read through its .x; mask: w x h.
 * ltm: guidedFilterEps, shadows, highlights, detail.  transform: YCbCr ->
 * camera RGB (inverse(cam_to_ycbcr) * exposure_multiplier).
 *
 * The box taps sit at (-1.5, 0.5, 2.0) + 0.5 texels of the MASK's size (the
 * reference scales them by 1 / output size), each axis with weight 2/5, 2/5,
 * 1/5 (a 5-texel box read by linear filtering: two pairs and a single); the
 * guide is sampled with the linear sampler (bilinear4).  The input pixel is
 * the texel itself (a linear read at its centre: the weights quantise to
 * 0 / 1).  The reference's "variance" is sum((w s)^2) - (sum w s)^2, as
 * written; min(L, 1) is the OpenCL definition "1 if 1 < L, else L" (NaN
 * stays NaN).  FP32 relaxed forms: the per-tap division by var + eps is a
 * multiplication by (luma - mean) / (var + eps), and
 * length(sqrt(2 rgb)) = sqrt(2r + 2g + 2b), NaN when a component is
 * negative (as the component's sqrt is). */
static const real ltm_off[3] = {K(-1.5), K(0.5), K(2.0)};
static const real ltm_w[3][3] = {{K(0.16), K(0.16), K(0.08)},   /* [y tap][x tap] */
                                 {K(0.16), K(0.16), K(0.08)},
                                 {K(0.08), K(0.08), K(0.04)}};

static inline real ltm_hclip(real r, real g, real b) {
#if ORACLE_DOUBLE
    const real a = sqrt(2 * r), c = sqrt(2 * g), d = sqrt(2 * b);
    const real L = sqrt(a * a + c * c + d * d);
#else
    const real L = (r >= K(0.0) && g >= K(0.0) && b >= K(0.0)) ? RSQRT(K(2.0) * r + K(2.0) * g + K(2.0) * b)
                                                               : clim_u2f(0x7fc00000u);
#endif
    return K(1.0) < L ? K(1.0) : L;
}

void P(ltm_mask)(const real* in, int w, int h, const real* guide, int gw, int gh, const float ltm[4],
                 const float transform[9], real* mask, real* var_eps) {
    real M[9];
    for (int k = 0; k < 9; k++) M[k] = (real)transform[k];
    const real eps = ltm[0], shadows = ltm[1], highlights = ltm[2], detail = ltm[3];
    const real nx = K(1.0) / (real)w, ny = K(1.0) / (real)h;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            const real px = (real)x * nx, py = (real)y * ny;
            real s[9], sum = 0, sumSq = 0;
            for (int j = 0; j < 3; j++) {
                for (int i = 0; i < 3; i++) {
                    real t[4];
                    bilinear4(guide, gw, gh, px + (ltm_off[i] + K(0.5)) * nx, py + (ltm_off[j] + K(0.5)) * ny, t);
                    s[j * 3 + i] = t[0];
                    const real ws = ltm_w[j][i] * t[0];
                    sum += ws;
                    sumSq += ws * ws;
                }
            }
            const real mean = sum, var = sumSq - sum * sum;
            if (var_eps) var_eps[(size_t)y * w + x] = var + eps;  /* conditioning, for the tests */
            const real* p = in + ((size_t)y * w + x) * 4;
            const real luma = p[0];
            real filtered = 0, norm = 0;
#if !ORACLE_DOUBLE
            const real q = (luma - mean) / (var + eps);
#endif
            for (int k = 0; k < 9; k++) {
#if ORACLE_DOUBLE
                const real sw = ltm_w[k / 3][k % 3] * (1 + (s[k] - mean) * (luma - mean) / (var + eps));
#else
                const real sw = ltm_w[k / 3][k % 3] * (K(1.0) + (s[k] - mean) * q);
#endif
                filtered += sw * s[k];
                norm += sw;
            }
            const real illuminance = filtered / norm;
            const real reflectance = luma / illuminance;
            const real r = M[0] * p[0] + M[1] * p[1] + M[2] * p[2];
            const real g = M[3] * p[0] + M[4] * p[1] + M[5] * p[2];
            const real b = M[6] * p[0] + M[7] * p[1] + M[8] * p[2];
            const real hc = ltm_hclip(r, g, b);
            const real tc = mixr(shadows, highlights, hc);
            mask[(size_t)y * w + x] = POW(illuminance, K(1.0) / tc) * POW(reflectance, detail) / luma;
        }
    }
}

/* --------------------------------------------------------------- convertTosRGB
 * cl/demosaic.cl:1391-1474.  out8 gets UNORM8 (round to nearest even,
 * saturating), outf (nullable) the float values.  ltm (nullable): the
 * localToneMapping mask (w x h), applied after the tone curve
 * (cl/demosaic.cl:1456-1466); FP32 relaxed forms there: pow(luma, 0.5) is
 * sqrt(luma) (luma >= 0 after the clamp) and the division by 1 - luma is
 * done once per pixel. */
static inline real sigmoidr(real x, real s) { return K(0.5) * (TANH(s * x - K(0.3) * s) + K(1.0)); }

static inline void ltm_boost(real rgb[3], real boost) {
    if (boost > K(1.0)) {
        const real luma = K(0.2126) * rgb[0] + K(0.7152) * rgb[1] + K(0.0722) * rgb[2];
#if ORACLE_DOUBLE
        const real t = pow(luma, 0.5);
        for (int c = 0; c < 3; c++) {
            const real hi = luma < 1 ? 1 - (1 - rgb[c]) * (1 - boost * luma) / (1 - luma) : rgb[c];
            rgb[c] = mixr(rgb[c] * boost, hi, t);
        }
#else
        const real t = sqrtf(luma);
        const real f = luma < K(1.0) ? (K(1.0) - boost * luma) / (K(1.0) - luma) : K(0.0);
        for (int c = 0; c < 3; c++) {
            const real hi = luma < K(1.0) ? K(1.0) - (K(1.0) - rgb[c]) * f : rgb[c];
            rgb[c] = mixr(rgb[c] * boost, hi, t);
        }
#endif
    } else if (boost < K(1.0)) {
        for (int c = 0; c < 3; c++) rgb[c] *= boost;
    }
}

void P(convert_to_srgb)(const real* in, int w, int h, const float transform[9], float contrast_f,
                        float saturation_f, float slope_f, float exposure_bias_f, float blacks_f,
                        const real* ltm, uint8_t* out8, real* outf) {
    real M[9];
    for (int k = 0; k < 9; k++) M[k] = (real)transform[k];
    const real contrast = contrast_f, saturation = saturation_f, s = slope_f;
    const real eb = exposure_bias_f, blacks = blacks_f;
#if ORACLE_DOUBLE
    const real ebmul = eb != 0 ? pow(2.0, eb) : 1.0;
#else
    const real ebmul = eb != K(0.0) ? clim_expf(eb * K(0.693147182)) : K(1.0);
#endif
    const real sig0 = sigmoidr(K(0.0), s), sig1 = sigmoidr(K(1.0), s);
    /* the two per-pixel divisions by constants as multiplications by their
     * reciprocals (a relaxed-math form) */
    const real inv_range = K(1.0) / (sig1 - sig0), inv_blacks = K(1.0) / (K(1.0) - blacks);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)w * h; i++) {
        real v[3] = {in[i * 4 + 0], in[i * 4 + 1], in[i * 4 + 2]};
        for (int c = 0; c < 3; c++) v[c] *= ebmul;
        if (saturation != K(1.0)) {
            real luma = K(0.2126) * v[0] + K(0.7152) * v[1] + K(0.0722) * v[2];
            for (int c = 0; c < 3; c++)
                v[c] = mixr(luma, v[c], mixr(saturation, K(1.0), smoothstepr(K(0.75), K(2.0), v[c])));
        }
        if (contrast != K(1.0)) {
            for (int c = 0; c < 3; c++)
                v[c] = mixr(K(0.2), v[c], mixr(contrast, K(1.0), smoothstepr(K(0.9), K(2.0), v[c])));
        }
        real rgb[3];
        for (int r = 0; r < 3; r++) rgb[r] = M[r * 3 + 0] * v[0] + M[r * 3 + 1] * v[1] + M[r * 3 + 2] * v[2];
        for (int c = 0; c < 3; c++) {
            real t = RSQRT(K(0.95) * rgb[c]);
            real tc = (sigmoidr(t, s) - sig0) * inv_range;
            rgb[c] = clampr(tc, K(0.0), K(1.0));
        }
        if (ltm) ltm_boost(rgb, ltm[i]);
        for (int c = 0; c < 3; c++)
            if (blacks > K(0.0)) rgb[c] = (rgb[c] - blacks) * inv_blacks;
        for (int c = 0; c < 3; c++) {
            real q = clampr(rgb[c], K(0.0), K(1.0)) * K(255.0);
#if ORACLE_DOUBLE
            out8[i * 4 + c] = (uint8_t)nearbyint(q);
#else
            out8[i * 4 + c] = (uint8_t)rintf(q);
#endif
            if (outf) outf[i * 4 + c] = rgb[c];
        }
        out8[i * 4 + 3] = 0;
        if (outf) outf[i * 4 + 3] = 0;
    }
}

/* ------------------------------------------------------------ host matrices
 * cam_ycbcr (cpp/demosaic_utils.cpp:809-822) with xyz_rgb of
 * cpp/demosaic.hpp:91-95, and the 3x3 cofactor inverse of
 * src/gls_linalg.hpp:544-557; both in FP64, rounded to float. */
void P(cam_ycbcr)(const float rgb_cam[9], float out[9]) {
    const double xyz_y[3] = {0.2126729, 0.7151522, 0.0721750};
    double k[3];
    for (int j = 0; j < 3; j++)
        k[j] = xyz_y[0] * (double)rgb_cam[0 * 3 + j] + xyz_y[1] * (double)rgb_cam[1 * 3 + j] +
               xyz_y[2] * (double)rgb_cam[2 * 3 + j];
    const double KR = k[0], KG = k[1], KB = k[2];
    const double m[9] = {KR, KG, KB,
                         -0.5 * KR / (1.0 - KB), -0.5 * KG / (1.0 - KB), 0.5,
                         0.5, -0.5 * KG / (1.0 - KR), -0.5 * KB / (1.0 - KR)};
    for (int i = 0; i < 9; i++) out[i] = (float)m[i];
}

int P(inverse3x3)(const float mf[9], double scale, float out[9]) {
    double m[9];
    for (int i = 0; i < 9; i++) m[i] = mf[i];
    double c00 = m[4] * m[8] - m[5] * m[7];
    double c01 = m[5] * m[6] - m[3] * m[8];
    double c02 = m[3] * m[7] - m[4] * m[6];
    double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (det == 0.0) return 1;
    double inv[9] = {c00, m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                     c01, m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                     c02, m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
    for (int i = 0; i < 9; i++) out[i] = (float)(inv[i] / det * scale);
    return 0;
}

/* ---------------------------------------------------------- fasteDebayer
 * cl/demosaic.cl:320-334: one (R, (G + G2) / 2, B, 0) texel per quad. */
void P(fast_debayer)(const real* raw, int w, int h, int bayer, real* rgba) {
    const int qw = w / 2, qh = h / 2;
#pragma omp parallel for schedule(static)
    for (int qy = 0; qy < qh; qy++) {
        for (int qx = 0; qx < qw; qx++) {
            real v[4];
            for (int c = 0; c < 4; c++)
                v[c] = raw[(size_t)(2 * qy + kBayerOffsets[bayer][c][1]) * w + 2 * qx + kBayerOffsets[bayer][c][0]];
            real* o = rgba + ((size_t)qy * qw + qx) * 4;
            o[0] = v[raw_red];
            o[1] = (v[raw_green] + v[raw_green2]) / K(2.0);
            o[2] = v[raw_blue];
            o[3] = 0;
        }
    }
}

/* ----------------------------------------------------- high-noise raw branch
 * cpp/raw_converter.cpp:111-123 (noiseLevel > 0.6): bayerToRawRGBA,
 * despeckleRawRGBAImage, denoiseRawRGBAImage, rawRGBAToBayer
 * (cl/demosaic.cl:898-929, 530-553, 932-949 + 1187-1201).  The RGBA raw
 * image holds one (R, G, B, G2) quad per texel, w/2 x h/2. */
void P(bayer_to_raw_rgba)(const real* raw, int w, int h, int bayer, real* rgba) {
    const int qw = w / 2, qh = h / 2;
#pragma omp parallel for schedule(static)
    for (int qy = 0; qy < qh; qy++)
        for (int qx = 0; qx < qw; qx++)
            for (int c = 0; c < 4; c++)
                rgba[((size_t)qy * qw + qx) * 4 + c] =
                    raw[(size_t)(2 * qy + kBayerOffsets[bayer][c][1]) * w + 2 * qx + kBayerOffsets[bayer][c][0]];
}

void P(raw_rgba_to_bayer)(const real* rgba, int qw, int qh, int bayer, real* raw) {
    const int w = 2 * qw;
#pragma omp parallel for schedule(static)
    for (int qy = 0; qy < qh; qy++)
        for (int qx = 0; qx < qw; qx++)
            for (int c = 0; c < 4; c++)
                raw[(size_t)(2 * qy + kBayerOffsets[bayer][c][1]) * w + 2 * qx + kBayerOffsets[bayer][c][0]] =
                    rgba[((size_t)qy * qw + qx) * 4 + c];
}

/* despeckle_3x3x4: per channel, the centre clamped to the second smallest and
 * second largest value of its 3x3 window (x outer, y inner), the trackers
 * starting at 0 (maxima) and 65535 (minima) as in the reference, so the bounds
 * are the second extremes of the window plus those two initial values; a
 * NaN is never selected.  Comparisons only: FP32 and FP64 agree exactly. */
void P(despeckle_raw_rgba)(const real* in, int qw, int qh, real* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < qh; y++) {
        for (int x = 0; x < qw; x++) {
            for (int c = 0; c < 4; c++) {
                real fmx = 0, smx = 0, fmn = K(65535.0), smn = K(65535.0);
                for (int dx = -1; dx <= 1; dx++) {
                    for (int dy = -1; dy <= 1; dy++) {
                        const real v = rd4(in, qw, qh, x + dx, y + dy)[c];
                        smx = (v <= fmx && v > smx) ? v : smx;
                        smx = v > fmx ? fmx : smx;
                        fmx = v > fmx ? v : fmx;
                        smn = (v >= fmn && v < smn) ? v : smn;
                        smn = v < fmn ? fmn : smn;
                        fmn = v < fmn ? v : fmn;
                    }
                }
                out[((size_t)y * qw + x) * 4 + c] = clampr(in[((size_t)y * qw + x) * 4 + c], smn, smx);
            }
        }
    }
}

/* denoiseRawRGBA: 5x5 average of the taps whose distance
 * length((s - c) / sigma) < 1, sigma = sqrt(rawVariance * c) per channel.  A
 * channel with c = 0 has sigma = 0: every distance is then NaN or infinite,
 * no tap counts and the quad is 0/0 = NaN, as the reference's expression is
 * under IEEE arithmetic.  FP32 relaxed form: sum (s - c)^2 / sigma^2 < 1
 * (clim_raw_dist4, 1/sigma^2 NaN for a negative product as the reference's
 * sqrt is), 0/1 weights accumulated by FMA (exact). */
void P(denoise_raw_rgba)(const real* in, int qw, int qh, const float var_f[4], real* out) {
    real var[4];
    for (int c = 0; c < 4; c++) var[c] = var_f[c];
#pragma omp parallel for schedule(static)
    for (int y = 0; y < qh; y++) {
        for (int x = 0; x < qw; x++) {
            const real* p = in + ((size_t)y * qw + x) * 4;
#if ORACLE_DOUBLE
            real sigma[4];
            for (int c = 0; c < 4; c++) sigma[c] = sqrt(var[c] * p[c]);
#else
            real k[4];
            for (int c = 0; c < 4; c++) k[c] = clim_inv_var(var[c] * p[c]);
#endif
            real acc[4] = {0, 0, 0, 0}, norm = 0;
            for (int dy = -2; dy <= 2; dy++) {
                for (int dx = -2; dx <= 2; dx++) {
                    const real* sv = rd4(in, qw, qh, x + dx, y + dy);
#if ORACLE_DOUBLE
                    real l = 0;
                    for (int c = 0; c < 4; c++) {
                        const real d = (sv[c] - p[c]) / sigma[c];
                        l += d * d;
                    }
                    const real wgt = sqrt(l) < 1 ? 1 : 0;
                    for (int c = 0; c < 4; c++) acc[c] += wgt * sv[c];
#else
                    const real l = clim_raw_dist4(sv[0] - p[0], sv[1] - p[1], sv[2] - p[2], sv[3] - p[3], k[0], k[1],
                                                  k[2], k[3]);
                    const real wgt = l < K(1.0) ? K(1.0) : K(0.0);
                    for (int c = 0; c < 4; c++) acc[c] = fmaf(wgt, sv[c], acc[c]);
#endif
                    norm += wgt;
                }
            }
            for (int c = 0; c < 4; c++) out[((size_t)y * qw + x) * 4 + c] = acc[c] / norm;
        }
    }
}

/* ------------------------------------------------------------- orchestration
 * PyramidalDenoise<levels>::denoise, cpp/pyramidal_denoise.cpp:103-149 and
 * RawConverter::demosaicImage, cpp/raw_converter.cpp:73-186. */
typedef struct {
    float luma, chroma, chromaBoost, gradientBoost, sharpening;
} o_denoise_params;

/* nlf_inout: levels x 6 (input used as fallback when a fit has no sample).
 * image: level 0 (w x h RGBA), out: w x h RGBA.  Returns a bit mask of the
 * levels whose fit fell back. */
static int P(pyramidal_denoise_s)(const real* image, int w, int h, int levels, int algorithm,
                                  const o_denoise_params* dp, float* nlf_inout, real* out, real* level3_out,
                                  const o_strength* sm, int half);
int P(pyramidal_denoise)(const real* image, int w, int h, int levels, int algorithm,
                         const o_denoise_params* dp, float* nlf_inout, real* out, real* level3_out) {
    return P(pyramidal_denoise_s)(image, w, h, levels, algorithm, dp, nlf_inout, out, level3_out, NULL, 0);
}

/* FP16 storage (the mixed-precision form's level images, and the reference's
 * ARM build, src/gls_image.hpp:32,189-210): each value rounded to the
 * nearest IEEE half, ties to even, with subnormals, overflow to infinity. */
static real P(half_round)(real x) {
    if (!(x == x) || x == 0) return x;
    const real ax = x < 0 ? -x : x;
    if (ax >= K(65520.0)) return x < 0 ? -(real)INFINITY : (real)INFINITY;
    int e;
    frexp((double)ax, &e);                 /* ax in [2^(e-1), 2^e) */
    if (e < -13) e = -13;                  /* below 2^-14: the subnormal step 2^-24 */
    const double ulp = ldexp(1.0, e - 11);
    const double r = nearbyint((double)ax / ulp) * ulp;
    return (real)(x < 0 ? -r : r);
}
void P(half_round_img)(real* img, size_t values) {
    for (size_t i = 0; i < values; i++) img[i] = P(half_round)(img[i]);
}
/* sm (nullable): the learned strength of the level filters (Bilateral and
 * GuidedFast; GuidedPrecise stays plain) */
/* half: the level images are stored as FP16 (the downsampled levels, the
 * denoised levels and the reassembled levels above 0; level 0's reassembly
 * goes straight to the tone map) */
static int P(pyramidal_denoise_s)(const real* image, int w, int h, int levels, int algorithm,
                                  const o_denoise_params* dp, float* nlf_inout, real* out, real* level3_out,
                                  const o_strength* sm, int half) {
    real* pyr[5] = {0};
    real* den[5] = {0};
    int lw[6], lh[6];
    int fallbacks = 0;
    lw[0] = w; lh[0] = h;
    for (int i = 1, scale = 2; i < levels; i++, scale *= 2) { lw[i] = w / scale; lh[i] = h / scale; }
    for (int i = 1; i < levels; i++) pyr[i] = (real*)malloc((size_t)lw[i] * lh[i] * 4 * sizeof(real));
    for (int i = 1; i < levels; i++) den[i] = (real*)malloc((size_t)lw[i] * lh[i] * 4 * sizeof(real));
    den[0] = out;
    const real* lvl[5];
    lvl[0] = image;
    for (int i = 1; i < levels; i++) lvl[i] = pyr[i];

    float cal[5][6];
    for (int i = 0; i < levels; i++) {
        if (i < levels - 1) {
            P(downsample)(lvl[i], lw[i], lh[i], pyr[i + 1], lw[i + 1], lh[i + 1]);
            if (half) P(half_round_img)(pyr[i + 1], (size_t)lw[i + 1] * lh[i + 1] * 4);
        }
        real* st = (real*)malloc((size_t)lw[i] * lh[i] * 4 * sizeof(real));
        P(noise_statistics)(lvl[i], lw[i], lh[i], st);
        float fit[6];
        int rc = P(nlf_fit)(3, 1, st, 4, st + 1, 4, (long)lw[i] * lh[i], fit);
        free(st);
        if (rc == 1) fallbacks |= 1 << i;
        else for (int k = 0; k < 6; k++) nlf_inout[i * 6 + k] = fit[k];
        float lm = dp[i].luma * dp[i].luma, cm = dp[i].chroma * dp[i].chroma;
        const float mul[6] = {lm, cm, cm, lm, cm, cm};
        for (int k = 0; k < 6; k++) cal[i][k] = nlf_inout[i * 6 + k] * mul[k];
    }

    for (int i = levels - 1; i >= 0; i--) {
        const float* np = cal[i];
        int use_bilateral = algorithm == 0 || (algorithm == 1 && i == levels - 1);
        if (algorithm == 2) {
            P(guided_filter)(lvl[i], lw[i], lh[i], 5, np + 3, den[i]);
        } else if (use_bilateral) {
            P(denoise_bilateral_s)(lvl[i], lw[i], lh[i], np, np + 3, dp[i].chromaBoost, dp[i].gradientBoost, sm, i,
                                   den[i]);
        } else {
            P(denoise_guided_fast_s)(lvl[i], lw[i], lh[i], np, np + 3, sm, i, den[i]);
        }
        if (half) P(half_round_img)(den[i], (size_t)lw[i] * lh[i] * 4);
        if (i < levels - 1) {
            const float nl[2] = {np[0], np[3]};
            P(reassemble)(den[i], lvl[i + 1], den[i + 1], lw[i], lh[i], lw[i + 1], lh[i + 1],
                          dp[i].sharpening, nl, den[i]);
            if (half && i > 0) P(half_round_img)(den[i], (size_t)lw[i] * lh[i] * 4);
        }
    }
    /* denoisedImagePyramid[3], the local tone mapping guide */
    if (level3_out && levels > 3) memcpy(level3_out, den[3], (size_t)lw[3] * lh[3] * 4 * sizeof(real));
    for (int i = 1; i < levels; i++) { free(pyr[i]); free(den[i]); }
    return fallbacks;
}

typedef struct {
    int bayer;
    float black_level, white_level, exposure_multiplier;
    float scale_mul[4];
    float rgb_cam[9];
    float rawNlf[4];
    float pyramidNlf[5][6];
    o_denoise_params dp[5];
    float contrast, saturation, toneCurveSlope, exposureBias, blacks;
    int levels, algorithm;
    int localToneMapping;  /* needs levels >= 4 */
    float ltm[4];          /* guidedFilterEps, shadows, highlights, detail */
    float noiseLevel;      /* > 0.6: the high-noise raw branch */
    int rotate_180;        /* RawConverter::demosaicImage's rotate_180 */
    int fp16_storage;      /* the level images stored as FP16 (the mixed-precision form's storage) */
} o_pipeline_params;

/* Full path.  nlf_out: rawNlf[4] then levels x 6.  Optional stage dumps
 * (nullable): ycc0 (despeckled level-0 YCbCr, RGBA), den0 (denoised level 0,
 * YCbCr).  Returns fallback mask (bit 8: raw fit). */
/* RawConverter::fastDemosaicImage, cpp/raw_converter.cpp:188-224: the
 * half-resolution preview (out8 / outf: w/2 x h/2). */
void P(fast_demosaic)(const uint16_t* raw, int w, int h, const o_pipeline_params* pp, uint8_t* out8, real* outf) {
    const int qw = w / 2, qh = h / 2;
    real* scaled = (real*)malloc((size_t)w * h * sizeof(real));
    real* rgb = (real*)malloc((size_t)qw * qh * 4 * sizeof(real));
    P(scale_raw)(raw, w, h, pp->bayer, pp->scale_mul, pp->black_level / K(65535.0), scaled);
    P(fast_debayer)(scaled, w, h, pp->bayer, rgb);
    P(blend_highlights)(rgb, qw, qh, 1.0f, rgb);
    P(convert_to_srgb)(rgb, qw, qh, pp->rgb_cam, pp->contrast, pp->saturation, pp->toneCurveSlope,
                       pp->exposureBias, pp->blacks, NULL, out8, outf);
    free(scaled);
    free(rgb);
}

/* The full path with the learned strength map (map, gw, gh, open as in
 * o_strength; map null: the plain path). */
int P(raw_to_rgb_strength)(const uint16_t* raw, int w, int h, const o_pipeline_params* pp, uint8_t* out8,
                           real* outf, float* nlf_out, real* ycc_dump, real* den_dump, const float* map, int gw,
                           int gh, const int* open);
int P(raw_to_rgb)(const uint16_t* raw, int w, int h, const o_pipeline_params* pp, uint8_t* out8,
                  real* outf, float* nlf_out, real* ycc_dump, real* den_dump) {
    return P(raw_to_rgb_strength)(raw, w, h, pp, out8, outf, nlf_out, ycc_dump, den_dump, NULL, 0, 0, NULL);
}
int P(raw_to_rgb_strength)(const uint16_t* raw, int w, int h, const o_pipeline_params* pp, uint8_t* out8,
                           real* outf, float* nlf_out, real* ycc_dump, real* den_dump, const float* map, int sgw,
                           int sgh, const int* open) {
    const o_strength strength = {map, sgw, sgh, open};
    size_t n = (size_t)w * h;
    int fallbacks = 0;
    real* scaled = (real*)malloc(n * sizeof(real));
    real* green = (real*)malloc(n * sizeof(real));
    real* rgbA = (real*)malloc(n * 4 * sizeof(real));
    real* rgbB = (real*)malloc(n * 4 * sizeof(real));

    P(scale_raw)(raw, w, h, pp->bayer, pp->scale_mul, pp->black_level / K(65535.0), scaled);

    float rawNlf[4];
    memcpy(rawNlf, pp->rawNlf, sizeof(rawNlf));
    {
        int qw = w / 2, qh = h / 2;
        real* mean = (real*)malloc((size_t)qw * qh * 4 * sizeof(real));
        real* var = (real*)malloc((size_t)qw * qh * 4 * sizeof(real));
        P(raw_noise_statistics)(scaled, w, h, pp->bayer, mean, var);
        float fit[8];
        int rc = P(nlf_fit)(4, 4, mean, 4, var, 4, (long)qw * qh, fit);
        if (rc == 1) fallbacks |= 1 << 8;
        else memcpy(rawNlf, fit + 4, sizeof(rawNlf));
        free(mean);
        free(var);
    }

    if (pp->noiseLevel > 0.6f) {
        /* cpp/raw_converter.cpp:111-123, with the fitted rawNlf */
        const int qw = w / 2, qh = h / 2;
        real* q0 = (real*)malloc((size_t)qw * qh * 4 * sizeof(real));
        real* q1 = (real*)malloc((size_t)qw * qh * 4 * sizeof(real));
        P(bayer_to_raw_rgba)(scaled, w, h, pp->bayer, q0);
        P(despeckle_raw_rgba)(q0, qw, qh, q1);
        P(denoise_raw_rgba)(q1, qw, qh, rawNlf, q0);
        P(raw_rgba_to_bayer)(q0, qw, qh, pp->bayer, scaled);
        free(q0);
        free(q1);
    }

    P(interpolate_green)(scaled, w, h, pp->bayer, rawNlf[1], green);
    /* cpp/raw_converter.cpp:127-128 passes rotate_180 to interpolateRedBlue,
     * whose rotated writes never reach row 0 or column 0: they start at zero */
    if (pp->rotate_180) memset(rgbA, 0, n * 4 * sizeof(real));
    P(interpolate_red_blue)(scaled, green, w, h, pp->bayer, (rawNlf[0] + rawNlf[2]) / 2.0f, pp->rotate_180, rgbA);
    P(blend_highlights)(rgbA, w, h, 1.0f, rgbA);
    float c2y[9], y2c[9];
    P(cam_ycbcr)(pp->rgb_cam, c2y);
    P(transform)(rgbA, w, h, c2y, rgbA);
    /* the rotated YCbCr image is stored before its despeckle (the unrotated
     * one goes to the despeckle without a store) */
    if (pp->fp16_storage && pp->rotate_180) P(half_round_img)(rgbA, n * 4);
    P(despeckle)(rgbA, w, h, pp->pyramidNlf[0], pp->pyramidNlf[0] + 3, rgbB);
    if (pp->fp16_storage) P(half_round_img)(rgbB, n * 4);
    if (ycc_dump) memcpy(ycc_dump, rgbB, n * 4 * sizeof(real));

    float nlf[5 * 6];
    memcpy(nlf, pp->pyramidNlf, sizeof(nlf));
    const int ltm = pp->localToneMapping && pp->levels > 3;
    const int gw = w >> 3, gh = h >> 3;  /* w/2/2/2 with truncation at every level */
    real* guide = ltm ? (real*)malloc((size_t)gw * gh * 4 * sizeof(real)) : NULL;
    fallbacks |= P(pyramidal_denoise_s)(rgbB, w, h, pp->levels, pp->algorithm, pp->dp, nlf, rgbA, guide,
                                        map ? &strength : NULL, pp->fp16_storage);
    if (den_dump) memcpy(den_dump, rgbA, n * 4 * sizeof(real));

    if (P(inverse3x3)(c2y, (double)pp->exposure_multiplier, y2c)) fallbacks |= 1 << 9;
    real* mask = NULL;
    if (ltm) {
        /* cpp/raw_converter.cpp:164-167: on the denoised YCbCr image */
        mask = (real*)malloc(n * sizeof(real));
        P(ltm_mask)(rgbA, w, h, guide, gw, gh, pp->ltm, y2c, mask, NULL);
    }
    P(transform)(rgbA, w, h, y2c, rgbA);
    P(convert_to_srgb)(rgbA, w, h, pp->rgb_cam, pp->contrast, pp->saturation, pp->toneCurveSlope,
                       pp->exposureBias, pp->blacks, mask, out8, outf);
    free(guide);
    free(mask);
    if (nlf_out) {
        memcpy(nlf_out, rawNlf, sizeof(rawNlf));
        memcpy(nlf_out + 4, nlf, (size_t)pp->levels * 6 * sizeof(float));
    }
    free(scaled); free(green); free(rgbA); free(rgbB);
    return fallbacks;
}
