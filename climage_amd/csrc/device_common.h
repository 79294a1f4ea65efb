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

#include <type_traits>

#include "clim_math.h"

namespace clim {

// pitch-linear device image view.  RGBA images hold FP32 texels, or FP16
// texels (8 bytes) when `half` is set: the mixed-precision pyramid's storage.
// Single-channel images are FP32 (or uint16 mosaics).
struct Img {
    char* p;
    int w, h, pitch;
    int half = 0;
};

typedef _Float16 half4_t __attribute__((ext_vector_type(4)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

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
// ld1 of (xg .. xg + 3, y): one 16-byte load when the four lie in the image
// and xg is a multiple of 4 (rows are 16-byte aligned)
__device__ __forceinline__ float4 ld1x4(const Img& im, int xg, int y) {
    if ((xg & 3) == 0 && xg >= 0 && xg + 3 < im.w)
        return *reinterpret_cast<const float4*>(row_f(im, clampi(y, 0, im.h - 1)) + xg);
    return make_float4(ld1(im, xg, y), ld1(im, xg + 1, y), ld1(im, xg + 2, y), ld1(im, xg + 3, y));
}
// RGBA texel (x, y) as FP32, from FP16 storage when H (the kernels of the
// pyramid take H as a template parameter, so that neither form branches per
// access; ldt4 / st4 / ld4 below decide at run time, for the other kernels)
template <bool H>
__device__ __forceinline__ float4 ldt4_(const Img& im, int x, int y) {
    if (H) {
        const half4_t v = reinterpret_cast<const half4_t*>(im.p + (size_t)y * im.pitch)[x];
        return make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    }
    return row_f4(im, y)[x];
}
// store of RGBA texel (x, y) (FP16 storage rounds to nearest even)
template <bool H>
__device__ __forceinline__ void st4_(const Img& im, int x, int y, float4 v) {
    if (H) {
        half4_t h;
        h.x = (_Float16)v.x;
        h.y = (_Float16)v.y;
        h.z = (_Float16)v.z;
        h.w = (_Float16)v.w;
        reinterpret_cast<half4_t*>(im.p + (size_t)y * im.pitch)[x] = h;
    } else {
        wrow_f4(im, y)[x] = v;
    }
}
template <bool H>
__device__ __forceinline__ float4 ld4_(const Img& im, int x, int y) {
    return ldt4_<H>(im, clampi(x, 0, im.w - 1), clampi(y, 0, im.h - 1));
}
__device__ __forceinline__ float4 ldt4(const Img& im, int x, int y) {
    return im.half ? ldt4_<true>(im, x, y) : ldt4_<false>(im, x, y);
}
__device__ __forceinline__ void st4(const Img& im, int x, int y, float4 v) {
    if (im.half)
        st4_<true>(im, x, y, v);
    else
        st4_<false>(im, x, y, v);
}
__device__ __forceinline__ float4 ld4(const Img& im, int x, int y) {
    return ldt4(im, clampi(x, 0, im.w - 1), clampi(y, 0, im.h - 1));
}

// Runs f(std::integral_constant<bool, H>) with H = the image's storage:
// launchers pick the kernel instance this way.
template <typename F>
inline void with_storage(const Img& im, F f) {
    if (im.half)
        f(std::true_type{});
    else
        f(std::false_type{});
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

// blendHighlightsImage on one pixel (cl/demosaic.cl:338-380): above the clip,
// the opponent-space chroma is rescaled to that of the clipped pixel
__device__ __forceinline__ void blend_highlights_px(float& px, float& py, float& pz, float clip) {
    // cl/demosaic.cl:338-370
    if (px > clip || py > clip || pz > clip) {
        const float s3 = 1.7320508f;
        const float c0[3] = {px, py, pz};
        const float c1[3] = {mmin(px, clip), mmin(py, clip), mmin(pz, clip)};
        const float l00 = 1.0f * c0[0] + 1.0f * c0[1] + 1.0f * c0[2];
        const float l01 = s3 * c0[0] + -s3 * c0[1] + 0.0f * c0[2];
        const float l02 = -1.0f * c0[0] + -1.0f * c0[1] + 2.0f * c0[2];
        const float l11 = s3 * c1[0] + -s3 * c1[1] + 0.0f * c1[2];
        const float l12 = -1.0f * c1[0] + -1.0f * c1[1] + 2.0f * c1[2];
        const float sum0 = l01 * l01 + l02 * l02;
        const float sum1 = l11 * l11 + l12 * l12;
        const float chratio = sum0 > 0 ? sqrtf(sum1 / sum0) : 1.0f;
        const float a = l00, bb = l01 * chratio, cc = l02 * chratio;
        const float h = s3 / 2.0f;
        px = (1.0f * a + h * bb + -0.5f * cc) / 3.0f;
        py = (1.0f * a + -h * bb + -0.5f * cc) / 3.0f;
        pz = (1.0f * a + 0.0f * bb + 1.0f * cc) / 3.0f;
    }
}

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

// Workgroup -> tile order of a 2-D grid.  The dispatcher hands workgroups to
// the device's 8 XCDs round robin (linear id mod 8) and each XCD has its own
// L2, so in launch order horizontally adjacent tiles run on different XCDs and
// every tile fetches its halo from HBM.  Renumbered, XCD x takes the x-th
// contiguous run of tiles in row-major order: neighbouring tiles, and the halo
// rows they share, meet in the same L2.  A bijection of [0, tiles).
struct TileId {
    int x, y;
};
__device__ __forceinline__ TileId xcd_tile() {
    const int gx = gridDim.x, n = gx * gridDim.y;
    const int lin = blockIdx.x + blockIdx.y * gx;
    const int q = n >> 3, r = n & 7, xcd = lin & 7;
    const int t = xcd * q + min(xcd, r) + (lin >> 3);
    return TileId{t % gx, t / gx};
}

// The same renumbering for a 1-D grid of persistent blocks that take
// contiguous runs of tiles: XCD x runs logical blocks [x n/8, (x+1) n/8), so
// vertically adjacent runs (whose halos overlap) share an L2.  A bijection.
__device__ __forceinline__ int xcd_block() {
    const int n = gridDim.x, lin = blockIdx.x;
    const int q = n >> 3, r = n & 7, xcd = lin & 7;
    return xcd * q + min(xcd, r) + (lin >> 3);
}

// Stages N elements into LDS with NT threads: every thread issues all of its
// global loads before it waits for any (a loop with a run-time trip count
// would wait for each load before issuing the next).
template <int N, int NT, typename LoadF, typename StoreF>
__device__ __forceinline__ void stage(int tid, LoadF load, StoreF store) {
    constexpr int IT = (N + NT - 1) / NT;
    using T = decltype(load(0));
    T v[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int i = tid + k * NT;
        if (i < N) v[k] = load(i);
    }
#pragma unroll
    for (int k = 0; k < IT; k++) {
        const int i = tid + k * NT;
        if (i < N) store(i, v[k]);
    }
}

// stage() in two steps: fetch() issues the global loads into registers and
// put() stores them to LDS.  A persistent kernel fetches its next tile before
// it computes the current one, so the loads are in flight during the compute.
template <int N, int NT, typename T>
struct Staged {
    static constexpr int IT = (N + NT - 1) / NT;
    T v[IT];
    template <typename LoadF>
    __device__ __forceinline__ void fetch(int tid, LoadF load) {
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * NT;
            if (i < N) v[k] = load(i);
        }
    }
    template <typename StoreF>
    __device__ __forceinline__ void put(int tid, StoreF store) const {
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * NT;
            if (i < N) store(i, v[k]);
        }
    }
};

// One raw quad (texels (2qx + ox[c], 2qy + oy[c]), c = R, G, B, G2) of the
// uint16 mosaic, scaled as scaleRawData does (cl/demosaic.cl:69-78), with
// texel coordinates clamped to the image.  Two 32-bit loads (the quad's two
// rows); the mosaic width is even.
__device__ __forceinline__ float scale_texel(uint32_t v16, int site, float4 smul, float black) {
    const float sm = site == 0 ? smul.x : site == 1 ? smul.y : site == 2 ? smul.z : smul.w;
    const float s = sm * (clim_div65535(v16) - black);  // = v16 / 65535.0f, see clim_div65535
    return s > 0.0f ? s : 0.0f;
}

// The mosaic as FP32 texels with clamp-to-edge: the scaled FP32 mosaic
// (U16 = false), or the uint16 mosaic scaled on the fly exactly as
// scaleRawData does (U16 = true), which saves writing and re-reading the
// scaled plane.
template <bool U16>
struct MosaicSrc {
    Img im;
    Bayer b;
    float4 smul;
    float black;
    __device__ __forceinline__ float operator()(int x, int y) const {
        if (!U16) return ld1(im, x, y);
        const int xc = clampi(x, 0, im.w - 1), yc = clampi(y, 0, im.h - 1);
        const uint32_t v16 = reinterpret_cast<const uint16_t*>(im.p + (size_t)yc * im.pitch)[xc];
        return scale_texel(v16, b.site(xc, yc), smul, black);
    }
    // texels (xg .. xg + 3, y), each the value operator() gives: one 8-byte
    // (16-byte) load when the four lie in the image and xg is a multiple of 4
    __device__ __forceinline__ float4 row4(int xg, int y) const {
        const int yc = clampi(y, 0, im.h - 1);
        if ((xg & 3) == 0 && xg >= 0 && xg + 3 < im.w) {
            const char* row = im.p + (size_t)yc * im.pitch;
            if (U16) {
                const uint2 w = *reinterpret_cast<const uint2*>(row + 2 * (size_t)xg);
                const int s0 = b.site(xg, yc), s1 = b.site(xg + 1, yc);  // sites alternate along the row
                return make_float4(scale_texel(w.x & 0xffffu, s0, smul, black), scale_texel(w.x >> 16, s1, smul, black),
                                   scale_texel(w.y & 0xffffu, s0, smul, black), scale_texel(w.y >> 16, s1, smul, black));
            }
            return *reinterpret_cast<const float4*>(row + 4 * (size_t)xg);
        }
        return make_float4((*this)(xg, y), (*this)(xg + 1, y), (*this)(xg + 2, y), (*this)(xg + 3, y));
    }
};

struct RawQuadLoader {
    Img raw;
    Bayer b;
    float4 smul;
    float black;
    // returns the raw pair (x even, x+1) of row y, with both clamped
    __device__ __forceinline__ uint32_t pair(int qx, int y) const {
        const int yc = clampi(y, 0, raw.h - 1);
        const int qc = clampi(qx, 0, raw.w / 2 - 1);
        uint32_t w = reinterpret_cast<const uint32_t*>(raw.p + (size_t)yc * raw.pitch)[qc];
        if (qx < 0) w = (w & 0xffffu) * 0x10001u;          // both clamp to x = 0
        if (qx > qc) w = (w >> 16) * 0x10001u;              // both clamp to x = W-1
        return w;
    }
    // the quad's two rows (loads only)
    __device__ __forceinline__ uint2 words(int qx, int qy) const {
        return make_uint2(pair(qx, 2 * qy), pair(qx, 2 * qy + 1));
    }
    // (R, G, B, G2) of quad (qx, qy), scaled
    __device__ __forceinline__ float4 quad(int qx, int qy) const { return quad(words(qx, qy), qx, qy); }
    __device__ __forceinline__ float4 quad(uint2 wds, int qx, int qy) const {
        const uint32_t r0 = wds.x, r1 = wds.y;
        if (qx >= 0 && qy >= 0 && qx < raw.w / 2 && qy < raw.h / 2) {
            // no texel is clamped: texel c sits at a site of colour c
            float v[4];
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const uint32_t row = b.oy[c] ? r1 : r0;
                v[c] = scale_texel(b.ox[c] ? (row >> 16) : (row & 0xffffu), c, smul, black);
            }
            return make_float4(v[0], v[1], v[2], v[3]);
        }
        // site parity of the clamped rows: row 2qy clamps to 0 or H-1 only
        // at the borders, where both rows clamp to the same row
        const int y0 = clampi(2 * qy, 0, raw.h - 1), y1 = clampi(2 * qy + 1, 0, raw.h - 1);
        const int x0 = qx < 0 ? 0 : (qx >= raw.w / 2 ? raw.w - 1 : 2 * qx);
        const int x1 = qx < 0 ? 0 : (qx >= raw.w / 2 ? raw.w - 1 : 2 * qx + 1);
        float v[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const uint32_t row = b.oy[c] ? r1 : r0;
            const uint32_t v16 = b.ox[c] ? (row >> 16) : (row & 0xffffu);
            const int xs = b.ox[c] ? x1 : x0, ys = b.oy[c] ? y1 : y0;
            v[c] = scale_texel(v16, b.site(xs, ys), smul, black);
        }
        return make_float4(v[0], v[1], v[2], v[3]);
    }
};

// 128-bit fixed-point helpers for wave reductions
__device__ __forceinline__ clim_fx128 fx_shfl_xor(clim_fx128 a, int m) {
    clim_fx128 r;
    r.hi = __shfl_xor(a.hi, m, 64);
    r.lo = __shfl_xor(a.lo, m, 64);
    return r;
}

}  // namespace clim
