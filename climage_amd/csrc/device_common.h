// device_common.h - image views, bounds-clamped loads and launch helpers
// shared by the gfx950 kernels of the RAW->RGB path.
//
// Every stencil kernel reproduces the FP32 arithmetic of the path operation
// for operation (the library is built with -ffp-contract=off), so that the
// threshold decisions downstream (hard sigma filter, noise-model inlier masks,
// despeckle clamps) come out the same as in any other IEEE FP32 evaluation of
// the same expressions.  Reuse across pixels (LDS tiles, shared Sobel
// responses) is fine; re-association of a sum is not.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clim_math.h"

namespace clim {

// pitch-linear device image view
struct Img {
    char* p;
    int w, h, pitch;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ const float* row_f(const Img& im, int y) {
    return reinterpret_cast<const float*>(im.p + (size_t)y * im.pitch);
}
__device__ __forceinline__ const float4* row_f4(const Img& im, int y) {
    return reinterpret_cast<const float4*>(im.p + (size_t)y * im.pitch);
}
__device__ __forceinline__ float* wrow_f(const Img& im, int y) {
    return reinterpret_cast<float*>(im.p + (size_t)y * im.pitch);
}
__device__ __forceinline__ float4* wrow_f4(const Img& im, int y) {
    return reinterpret_cast<float4*>(im.p + (size_t)y * im.pitch);
}

// sampler-less read with clamp-to-edge
__device__ __forceinline__ float ld1(const Img& im, int x, int y) {
    return row_f(im, clampi(y, 0, im.h - 1))[clampi(x, 0, im.w - 1)];
}
__device__ __forceinline__ float4 ld4(const Img& im, int x, int y) {
    return row_f4(im, clampi(y, 0, im.h - 1))[clampi(x, 0, im.w - 1)];
}

// OpenCL built-ins
__device__ __forceinline__ float smoothstep(float e0, float e1, float x) { return clim_smoothstepf(e0, e1, x); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return clim_mixf(a, b, t); }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return clim_clampf(x, lo, hi); }
// the min/max/abs macros of cl/demosaic.cl:57-66
__device__ __forceinline__ float mmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float mmax(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float len2(float x, float y) { return sqrtf(x * x + y * y); }
__device__ __forceinline__ float len3(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// Bayer table of cl/demosaic.cl:29-34, resolved on the host for one pattern:
// ox/oy[c] = (x, y) of raw_red, raw_green, raw_blue, raw_green2 in the quad;
// site(x, y) = which of the four a pixel is (2 bits per parity class, packed
// so that no kernel indexes an argument array dynamically).
struct Bayer {
    int ox[4], oy[4];
    int site_packed;
    __device__ __forceinline__ int site(int x, int y) const {
        return (site_packed >> ((((y & 1) << 1) | (x & 1)) << 1)) & 3;
    }
};

// Noise-model scalars a kernel needs: taken from device memory (written by the
// on-device regression) when `dev` is set, else from the host values.
struct ParamRef {
    const float* dev;
    float host[8];
    __device__ __forceinline__ float get(int i) const { return dev ? dev[i] : host[i]; }
};

// 128-bit fixed-point helpers for wave reductions
__device__ __forceinline__ clim_fx128 fx_shfl_xor(clim_fx128 a, int m) {
    clim_fx128 r;
    r.hi = __shfl_xor(a.hi, m, 64);
    r.lo = __shfl_xor(a.lo, m, 64);
    return r;
}

}  // namespace clim
