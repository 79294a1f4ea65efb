// k_preview.hip - the half-resolution preview of RawConverter::fastDemosaicImage
// (cpp/raw_converter.cpp:188-224) as one kernel: scaleRawData, fasteDebayer
// (cl/demosaic.cl:320-334), blendHighlightsImage (clip 1) and convertTosRGB
// on each 2x2 quad, which becomes one output texel.  The reference runs four
// launches through two half-resolution FP32 images; here the uint16 mosaic is
// read once and RGBA8 written once (2 + 1 bytes per mosaic pixel).
#include "tone.h"

namespace clim {

template <bool U16>
__device__ __forceinline__ float4 preview_quad(const MosaicArg& m, int qx, int qy) {
    if (U16) return RawQuadLoader{m.im, m.b, m.smul, m.black}.quad(qx, qy);
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = row_f(m.im, 2 * qy + m.b.oy[c])[2 * qx + m.b.ox[c]];
    return make_float4(v[0], v[1], v[2], v[3]);
}

// fasteDebayer on one quad: (R, (G + G2) / 2, B)
__device__ __forceinline__ void fast_debayer_px(float4 q, float v[3]) {
    v[0] = q.x;
    v[1] = (q.y + q.w) / 2.0f;
    v[2] = q.z;
}

template <bool U16>
__global__ void __launch_bounds__(256) k_fast_preview(MosaicArg m, Img out8, Img outf, ToneArgs ta) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= out8.w || y >= out8.h) return;
    float v[3];
    fast_debayer_px(preview_quad<U16>(m, x, y), v);
    blend_highlights_px(v[0], v[1], v[2], 1.0f);
    tone_store_v<false>(out8, outf, x, y, v, ta, 1.0f);
}

__global__ void __launch_bounds__(256) k_fast_debayer(Img raw, Img rgba, Bayer b) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= rgba.w || y >= rgba.h) return;
    const MosaicArg m{raw, b, make_float4(0.0f, 0.0f, 0.0f, 0.0f), 0.0f, false};
    float v[3];
    fast_debayer_px(preview_quad<false>(m, x, y), v);
    wrow_f4(rgba, y)[x] = make_float4(v[0], v[1], v[2], 0.0f);
}

void launch_fast_debayer(const Img& raw, const Img& rgba, const Bayer& b, hipStream_t st) {
    dim3 blk(64, 4), grd((rgba.w + 63) / 64, (rgba.h + 3) / 4);
    hipLaunchKernelGGL(k_fast_debayer, grd, blk, 0, st, raw, rgba, b);
}

void launch_fast_preview(const MosaicArg& m, const Img& out8, const Img& outf, const float cam[9], float contrast,
                         float saturation, float slope, float exposureBias, float blacks, hipStream_t st) {
    const ToneArgs ta = make_tone_args(nullptr, cam, contrast, saturation, slope, exposureBias, blacks);
    dim3 blk(64, 4), grd((out8.w + 63) / 64, (out8.h + 3) / 4);
    if (m.u16)
        hipLaunchKernelGGL(k_fast_preview<true>, grd, blk, 0, st, m, out8, outf, ta);
    else
        hipLaunchKernelGGL(k_fast_preview<false>, grd, blk, 0, st, m, out8, outf, ta);
}

}  // namespace clim
