// k_guided.hip - the multi-pass colour guided filter (GuidedPrecise).
//
// GuidedFilter::filter, cpp/guided_filter.cpp:195-225, cl/guided_filter.cl.
// The box filter keeps the reference's arithmetic (row sum of filterSize
// texels / filterSize, then column sum of those / filterSize, clamp-to-edge)
// but writes both passes untransposed, so every access is coalesced.
#include "kernels.h"

namespace clim {

template <int NC>
__device__ __forceinline__ const float* px(const Img& im, int x, int y) {
    return reinterpret_cast<const float*>(im.p + (size_t)y * im.pitch) + x * NC;
}
template <int NC>
__device__ __forceinline__ float* wpx(const Img& im, int x, int y) {
    return reinterpret_cast<float*>(im.p + (size_t)y * im.pitch) + x * NC;
}

// boxBlurX pass along rows (cl/guided_filter.cl:16-28)
template <int NC>
__global__ void __launch_bounds__(256) k_hblur(Img in, Img out, int fs) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    float s[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) s[c] = 0.0f;
    for (int d = -fs / 2; d <= fs / 2; d++) {
        const float* p = px<NC>(in, clampi(x + d, 0, in.w - 1), y);
#pragma unroll
        for (int c = 0; c < NC; c++) s[c] += p[c];
    }
    float* o = wpx<NC>(out, x, y);
#pragma unroll
    for (int c = 0; c < NC; c++) o[c] = s[c] / (float)fs;
}

// the second boxBlurX pass: along columns of the row-blurred image
template <int NC>
__global__ void __launch_bounds__(256) k_vblur(Img in, Img out, int fs) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    float s[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) s[c] = 0.0f;
    for (int d = -fs / 2; d <= fs / 2; d++) {
        const float* p = px<NC>(in, x, clampi(y + d, 0, in.h - 1));
#pragma unroll
        for (int c = 0; c < NC; c++) s[c] += p[c];
    }
    float* o = wpx<NC>(out, x, y);
#pragma unroll
    for (int c = 0; c < NC; c++) o[c] = s[c] / (float)fs;
}

// covMatrixProducts, cl/guided_filter.cl:30-44
__global__ void __launch_bounds__(256) k_cov_products(Img in, Img s1, Img s2) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    const float* p = px<4>(in, x, y);
    float* a = wpx<4>(s1, x, y);
    float* b = wpx<2>(s2, x, y);
    a[0] = p[0] * p[0]; a[1] = p[0] * p[1]; a[2] = p[0] * p[2]; a[3] = p[1] * p[1];
    b[0] = p[1] * p[2]; b[1] = p[2] * p[2];
}

// invSigma, cl/guided_filter.cl:46-83 (in place on s1, s2)
__global__ void __launch_bounds__(256) k_inv_sigma(Img meanI, Img s1, Img s2, ParamRef eps) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= meanI.w || y >= meanI.h) return;
    const float e0 = eps.get(0), e1 = eps.get(1), e2 = eps.get(2);
    const float* mI = px<4>(meanI, x, y);
    float* p1 = wpx<4>(s1, x, y);
    float* p2 = wpx<2>(s2, x, y);
    const float vrr = p1[0] - mI[0] * mI[0] + e0 * mI[0];
    const float vrg = p1[1] - mI[0] * mI[1];
    const float vrb = p1[2] - mI[0] * mI[2];
    const float vgg = p1[3] - mI[1] * mI[1] + e1;
    const float vgb = p2[0] - mI[1] * mI[2];
    const float vbb = p2[1] - mI[2] * mI[2] + e2;
    const float invrr = vgg * vbb - vgb * vgb;
    const float invrg = vgb * vrb - vrg * vbb;
    const float invrb = vrg * vgb - vgg * vrb;
    const float invgg = vrr * vbb - vrb * vrb;
    const float invgb = vrb * vrg - vrr * vgb;
    const float invbb = vrr * vgg - vrg * vrg;
    const float covDet = invrr * vrr + invrg * vrg + invrb * vrb;
    p1[0] = invrr / covDet; p1[1] = invrg / covDet; p1[2] = invrb / covDet; p1[3] = invgg / covDet;
    p2[0] = invgb / covDet; p2[1] = invbb / covDet;
}

// meanIpProducts, cl/guided_filter.cl:85-95
__global__ void __launch_bounds__(256) k_mean_ip(Img in, Img ipr, Img ipg, Img ipb) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    const float* p = px<4>(in, x, y);
    float* r = wpx<4>(ipr, x, y);
    float* g = wpx<4>(ipg, x, y);
    float* b = wpx<4>(ipb, x, y);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        r[c] = p[0] * p[c];
        g[c] = p[1] * p[c];
        b[c] = p[2] * p[c];
    }
    r[3] = g[3] = b[3] = 0.0f;
}

// computeAb, cl/guided_filter.cl:97-137 (in place on ipr, ipg, ipb)
__global__ void __launch_bounds__(256) k_compute_ab(Img meanI, Img ipr, Img ipg, Img ipb, Img s1, Img s2) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= meanI.w || y >= meanI.h) return;
    const float* mp = px<4>(meanI, x, y);
    float* r = wpx<4>(ipr, x, y);
    float* g = wpx<4>(ipg, x, y);
    float* b = wpx<4>(ipb, x, y);
    float cr[3], cg[3], cb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cr[c] = r[c] - mp[0] * mp[c];
        cg[c] = g[c] - mp[1] * mp[c];
        cb[c] = b[c] - mp[2] * mp[c];
    }
    const float* p1 = px<4>(s1, x, y);
    const float* p2 = px<2>(s2, x, y);
    const float invrr = p1[0], invrg = p1[1], invrb = p1[2], invgg = p1[3], invgb = p2[0], invbb = p2[1];
    float ar[3], ag[3], ab[3], bb[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        ar[c] = invrr * cr[c] + invrg * cg[c] + invrb * cb[c];
        ag[c] = invrg * cr[c] + invgg * cg[c] + invgb * cb[c];
        ab[c] = invrb * cr[c] + invgb * cg[c] + invbb * cb[c];
    }
#pragma unroll
    for (int c = 0; c < 3; c++) bb[c] = mp[c] - ar[c] * mp[0] - ag[c] * mp[1] - ab[c] * mp[2];
#pragma unroll
    for (int c = 0; c < 3; c++) { r[c] = ar[c]; g[c] = ag[c]; b[c] = ab[c]; }
    r[3] = bb[0]; g[3] = bb[1]; b[3] = bb[2];
}

// computeResult, cl/guided_filter.cl:139-154
__global__ void __launch_bounds__(256) k_guided_result(Img in, Img ipr, Img ipg, Img ipb, Img out) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x, y = tl.y * 4 + threadIdx.y;
    if (x >= in.w || y >= in.h) return;
    const float* p = px<4>(in, x, y);
    const float* r = px<4>(ipr, x, y);
    const float* g = px<4>(ipg, x, y);
    const float* b = px<4>(ipb, x, y);
    float q[3];
#pragma unroll
    for (int c = 0; c < 3; c++) q[c] = r[c] * p[0] + g[c] * p[1] + b[c] * p[2];
    q[0] += r[3]; q[1] += g[3]; q[2] += b[3];
    wrow_f4(out, y)[x] = make_float4(q[0], q[1], q[2], 0.0f);
}

template <int NC>
static void box_blur(const Img& in, const Img& out, const Img& tmp, int fs, dim3 grd, dim3 blk, hipStream_t st) {
    hipLaunchKernelGGL(k_hblur<NC>, grd, blk, 0, st, in, tmp, fs);
    hipLaunchKernelGGL(k_vblur<NC>, grd, blk, 0, st, tmp, out, fs);
}

void launch_guided_filter(const Img& in, const Img& out, int fs, const ParamRef& eps, const GuidedScratch& g,
                          hipStream_t st) {
    dim3 blk(64, 4), grd((in.w + 63) / 64, (in.h + 3) / 4);
    box_blur<4>(in, g.mean_I, g.tmp4, fs, grd, blk, st);
    hipLaunchKernelGGL(k_cov_products, grd, blk, 0, st, in, g.s1, g.s2);
    box_blur<4>(g.s1, g.s1, g.tmp4, fs, grd, blk, st);
    box_blur<2>(g.s2, g.s2, g.tmp2, fs, grd, blk, st);
    hipLaunchKernelGGL(k_inv_sigma, grd, blk, 0, st, g.mean_I, g.s1, g.s2, eps);
    hipLaunchKernelGGL(k_mean_ip, grd, blk, 0, st, in, g.ipr, g.ipg, g.ipb);
    box_blur<4>(g.ipr, g.ipr, g.tmp4, fs, grd, blk, st);
    box_blur<4>(g.ipg, g.ipg, g.tmp4, fs, grd, blk, st);
    box_blur<4>(g.ipb, g.ipb, g.tmp4, fs, grd, blk, st);
    hipLaunchKernelGGL(k_compute_ab, grd, blk, 0, st, g.mean_I, g.ipr, g.ipg, g.ipb, g.s1, g.s2);
    box_blur<4>(g.ipr, g.ipr, g.tmp4, fs, grd, blk, st);
    box_blur<4>(g.ipg, g.ipg, g.tmp4, fs, grd, blk, st);
    box_blur<4>(g.ipb, g.ipb, g.tmp4, fs, grd, blk, st);
    hipLaunchKernelGGL(k_guided_result, grd, blk, 0, st, in, g.ipr, g.ipg, g.ipb, out);
}

}  // namespace clim
