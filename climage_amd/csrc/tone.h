// tone.h - the per-pixel convertTosRGB (cl/demosaic.cl:1380-1474) shared by
// the kernels that end in the tone map: the level-0 reassembly of the
// pipeline, the stand-alone operator and the half-resolution preview.
#pragma once

#include "kernels.h"

namespace clim {

// cl/demosaic.cl:1391-1474, optionally preceded by
// transformImage with `pre` (the YCbCr->camera matrix of
// cpp/raw_converter.cpp:170).
struct ToneArgs {
    float pre[9];
    float cam[9];
    float contrast, saturation, slope, ebmul, blacks;
    float sig0, inv_range, inv_blacks;  // sigmoid(0), 1 / (sigmoid(1) - sigmoid(0)), 1 / (1 - blacks)
    int has_pre;
};

__device__ __forceinline__ float sigmoidf(float x, float s) { return 0.5f * (clim_tanhf(s * x - 0.3f * s) + 1.0f); }

// the optional transformImage in front of the tone map
__device__ __forceinline__ void tone_pre(float4 p, const ToneArgs& a, float v[3]) {
    v[0] = p.x; v[1] = p.y; v[2] = p.z;
    if (a.has_pre) {
        const float t0 = a.pre[0] * v[0] + a.pre[1] * v[1] + a.pre[2] * v[2];
        const float t1 = a.pre[3] * v[0] + a.pre[4] * v[1] + a.pre[5] * v[2];
        const float t2 = a.pre[6] * v[0] + a.pre[7] * v[1] + a.pre[8] * v[2];
        v[0] = t0; v[1] = t1; v[2] = t2;
    }
}

// The local tone mapping gain of cl/demosaic.cl:1456-1466, on the
// tone-mapped pixel.  pow(luma, 0.5) is sqrt(luma) (luma >= 0 after the
// clamp) and the division by 1 - luma is done once.
__device__ __forceinline__ void ltm_boost(float rgb[3], float boost) {
    if (boost > 1.0f) {
        const float luma = 0.2126f * rgb[0] + 0.7152f * rgb[1] + 0.0722f * rgb[2];
        const float t = sqrtf(luma);
        const float f = luma < 1.0f ? (1.0f - boost * luma) / (1.0f - luma) : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float hi = luma < 1.0f ? 1.0f - (1.0f - rgb[c]) * f : rgb[c];
            rgb[c] = mixf(rgb[c] * boost, hi, t);
        }
    } else if (boost < 1.0f) {
#pragma unroll
        for (int c = 0; c < 3; c++) rgb[c] *= boost;
    }
}

// one pixel from camera RGB v: RGBA8 (packed) and the FP32 RGB before
// quantisation; LTMB applies the local tone mapping gain `boost`
template <bool LTMB>
__device__ __forceinline__ uint32_t tone_rest(float v[3], const ToneArgs& a, float rgb[3], float boost) {
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] *= a.ebmul;
    // smoothstep's division by the constant edge spans as a reciprocal
    // multiplication with one fused correction: the same values
    // (clim_smoothstep_rc)
    if (a.saturation != 1.0f) {
        const float luma = 0.2126f * v[0] + 0.7152f * v[1] + 0.0722f * v[2];
        constexpr float span = 2.0f - 0.75f;
#pragma unroll
        for (int c = 0; c < 3; c++)
            v[c] = mixf(luma, v[c], mixf(a.saturation, 1.0f, clim_smoothstep_rc(0.75f, span, 1.0f / span, v[c])));
    }
    if (a.contrast != 1.0f) {
        constexpr float span = 2.0f - 0.9f;
#pragma unroll
        for (int c = 0; c < 3; c++)
            v[c] = mixf(0.2f, v[c], mixf(a.contrast, 1.0f, clim_smoothstep_rc(0.9f, span, 1.0f / span, v[c])));
    }
#pragma unroll
    for (int r = 0; r < 3; r++) rgb[r] = a.cam[r * 3 + 0] * v[0] + a.cam[r * 3 + 1] * v[1] + a.cam[r * 3 + 2] * v[2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float t = sqrtf(0.95f * rgb[c]);
        // divisions by constants as multiplications by their reciprocals
        const float tc = (sigmoidf(t, a.slope) - a.sig0) * a.inv_range;
        rgb[c] = clampf(tc, 0.0f, 1.0f);
    }
    if (LTMB) ltm_boost(rgb, boost);
    uint32_t packed = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (a.blacks > 0.0f) rgb[c] = (rgb[c] - a.blacks) * a.inv_blacks;
        const uint32_t q = (uint32_t)rintf(clampf(rgb[c], 0.0f, 1.0f) * 255.0f);
        packed |= q << (8 * c);
    }
    return packed;
}

template <bool LTMB>
__device__ __forceinline__ void tone_store_v(const Img& out8, const Img& outf, int x, int y, float v[3],
                                             const ToneArgs& a, float boost) {
    float rgb[3];
    const uint32_t packed = tone_rest<LTMB>(v, a, rgb, boost);
    reinterpret_cast<uint32_t*>(out8.p + (size_t)y * out8.pitch)[x] = packed;
    if (outf.p) wrow_f4(outf, y)[x] = make_float4(rgb[0], rgb[1], rgb[2], 0.0f);
}

__device__ __forceinline__ void tone_store(const Img& out8, const Img& outf, int x, int y, float4 p,
                                           const ToneArgs& a) {
    float v[3];
    tone_pre(p, a, v);
    tone_store_v<false>(out8, outf, x, y, v, a, 1.0f);
}

// host: the launch constants
static inline ToneArgs make_tone_args(const float* pre, const float cam[9], float contrast, float saturation, float slope,
                               float exposureBias, float blacks) {
    ToneArgs a;
    a.has_pre = pre != nullptr;
    for (int i = 0; i < 9; i++) {
        a.pre[i] = pre ? pre[i] : 0.0f;
        a.cam[i] = cam[i];
    }
    a.contrast = contrast;
    a.saturation = saturation;
    a.slope = slope;
    // powr(2, exposureBias) as exp(exposureBias * ln 2)
    a.ebmul = exposureBias != 0.0f ? clim_expf(exposureBias * 0.693147182f) : 1.0f;
    a.blacks = blacks;
    a.sig0 = 0.5f * (clim_tanhf(slope * 0.0f - 0.3f * slope) + 1.0f);
    const float sig1 = 0.5f * (clim_tanhf(slope * 1.0f - 0.3f * slope) + 1.0f);
    a.inv_range = 1.0f / (sig1 - a.sig0);
    a.inv_blacks = 1.0f / (1.0f - blacks);
    return a;
}

}  // namespace clim
