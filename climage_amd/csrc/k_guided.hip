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

// ---------------------------------------------------------------------------
// The whole filter in one launch for filter sizes 3 and 5 (R = 1, 2).
//
// A block owns a TX x TY tile of outputs and keeps every intermediate in LDS:
//   IN  the input's three colour channels over the tile grown by 2R
//       (texels at clamped coordinates);
//   H   row sums / FS of I (3) and of the six covariance products
//       (rr, rg, rb, gg, gb, bb) for the rows of IN and the columns of AB;
//   AB  the box means of those 9 channels (column sums of H / FS), turned in
//       place into invSigma and computeAb's a_r, a_g, a_b, b (12 channels)
//       over the tile grown by R, each texel at its clamped image position,
//       as the second box filter reads ab clamp-to-edge;
//   HA  row sums / FS of AB (aliases H), then column sums / FS give the
//       box-filtered ab of each output texel and computeResult.
// meanIpProducts needs no pass of its own: p_c * p_c' is the covariance
// product of (c, c') (multiplication commutes exactly), so the box means of
// ip_r, ip_g, ip_b are (rr, rg, rb), (rg, gg, gb), (rb, gb, bb) bit for bit.
// Every sum runs in the order of the two-pass filter (taps -R..R from 0,
// then / FS), so the result equals it bit for bit.
namespace gp {
constexpr int TX = 32, TY = 16, NT = 256;
template <int R> struct Dims {
    static constexpr int FS = 2 * R + 1;
    static constexpr int INW = TX + 4 * R, INH = TY + 4 * R;  // IN
    static constexpr int ABW = TX + 2 * R, ABH = TY + 2 * R;  // AB; H is INH rows x ABW columns
    static constexpr int H_SIZE = 9 * INH * ABW, HA_SIZE = 12 * ABH * TX;
    static constexpr int HH_SIZE = H_SIZE > HA_SIZE ? H_SIZE : HA_SIZE;
    static constexpr int LDS = 3 * INW * INH + HH_SIZE + 12 * ABH * ABW;  // floats
};
}  // namespace gp

template <int R, bool H>
__global__ void __launch_bounds__(gp::NT, 2) k_guided_precise(Img in, Img out, ParamRef eps) {
    using D = gp::Dims<R>;
    constexpr int TX = gp::TX, TY = gp::TY, NT = gp::NT, FS = D::FS;
    constexpr int INW = D::INW, INH = D::INH, ABW = D::ABW, ABH = D::ABH;
    __shared__ float lds[D::LDS];
    float* IN = lds;                          // [3][INH][INW]
    float* HH = IN + 3 * INW * INH;           // H [9][INH][ABW], later HA [12][ABH][TX]
    float* AB = HH + D::HH_SIZE;              // [12][ABH][ABW]
    const TileId tl = xcd_tile();
    const int x0 = tl.x * TX, y0 = tl.y * TY, tid = threadIdx.x;
    const int w = in.w, h = in.h;

    // IN: texel (x0 - 2R + c, y0 - 2R + r) at clamped coordinates
    stage<INW * INH, NT>(
        tid,
        [&](int i) {
            const int r = i / INW, c = i - r * INW;
            return ldt4_<H>(in, clampi(x0 - 2 * R + c, 0, w - 1), clampi(y0 - 2 * R + r, 0, h - 1));
        },
        [&](int i, float4 v) {
            IN[i] = v.x;
            IN[INW * INH + i] = v.y;
            IN[2 * INW * INH + i] = v.z;
        });
    __syncthreads();

    // H: row r of IN, column i of AB (image column clamp(x0 - R + i))
    for (int k = tid; k < INH * ABW; k += NT) {
        const int r = k / ABW, i = k - r * ABW;
        const int c0 = clampi(x0 - R + i, 0, w - 1) - (x0 - 2 * R);
        float s[9];
#pragma unroll
        for (int c = 0; c < 9; c++) s[c] = 0.0f;
#pragma unroll
        for (int d = -R; d <= R; d++) {
            const int o = r * INW + c0 + d;
            const float p0 = IN[o], p1 = IN[INW * INH + o], p2 = IN[2 * INW * INH + o];
            s[0] += p0; s[1] += p1; s[2] += p2;
            s[3] += p0 * p0; s[4] += p0 * p1; s[5] += p0 * p2;
            s[6] += p1 * p1; s[7] += p1 * p2; s[8] += p2 * p2;
        }
#pragma unroll
        for (int c = 0; c < 9; c++) HH[(c * INH + r) * ABW + i] = s[c] / (float)FS;
    }
    __syncthreads();

    // AB: box means at (clamp(x0 - R + i), clamp(y0 - R + j)), invSigma, computeAb
    {
        const float e0 = eps.get(0), e1 = eps.get(1), e2 = eps.get(2);
        for (int k = tid; k < ABH * ABW; k += NT) {
            const int j = k / ABW, i = k - j * ABW;
            const int r0 = clampi(y0 - R + j, 0, h - 1) - (y0 - 2 * R);
            float m[9];
#pragma unroll
            for (int c = 0; c < 9; c++) m[c] = 0.0f;
#pragma unroll
            for (int d = -R; d <= R; d++) {
#pragma unroll
                for (int c = 0; c < 9; c++) m[c] += HH[(c * INH + r0 + d) * ABW + i];
            }
#pragma unroll
            for (int c = 0; c < 9; c++) m[c] = m[c] / (float)FS;
            const float mI[3] = {m[0], m[1], m[2]};
            // invSigma (cl/guided_filter.cl:46-83)
            const float vrr = m[3] - mI[0] * mI[0] + e0 * mI[0];
            const float vrg = m[4] - mI[0] * mI[1];
            const float vrb = m[5] - mI[0] * mI[2];
            const float vgg = m[6] - mI[1] * mI[1] + e1;
            const float vgb = m[7] - mI[1] * mI[2];
            const float vbb = m[8] - mI[2] * mI[2] + e2;
            float invrr = vgg * vbb - vgb * vgb;
            float invrg = vgb * vrb - vrg * vbb;
            float invrb = vrg * vgb - vgg * vrb;
            float invgg = vrr * vbb - vrb * vrb;
            float invgb = vrb * vrg - vrr * vgb;
            float invbb = vrr * vgg - vrg * vrg;
            const float covDet = invrr * vrr + invrg * vrg + invrb * vrb;
            invrr = invrr / covDet; invrg = invrg / covDet; invrb = invrb / covDet;
            invgg = invgg / covDet; invgb = invgb / covDet; invbb = invbb / covDet;
            // computeAb (cl/guided_filter.cl:97-137): the box means of ip_r, ip_g, ip_b
            const float ipr[3] = {m[3], m[4], m[5]}, ipg[3] = {m[4], m[6], m[7]}, ipb[3] = {m[5], m[7], m[8]};
            float ar[3], ag[3], ab[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float cr = ipr[c] - mI[0] * mI[c];
                const float cg = ipg[c] - mI[1] * mI[c];
                const float cb = ipb[c] - mI[2] * mI[c];
                ar[c] = invrr * cr + invrg * cg + invrb * cb;
                ag[c] = invrg * cr + invgg * cg + invgb * cb;
                ab[c] = invrb * cr + invgb * cg + invbb * cb;
            }
            float* o = AB + j * ABW + i;
            constexpr int P = ABH * ABW;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                o[c * P] = ar[c];
                o[(4 + c) * P] = ag[c];
                o[(8 + c) * P] = ab[c];
                o[(3 + 4 * c) * P] = mI[c] - ar[c] * mI[0] - ag[c] * mI[1] - ab[c] * mI[2];
            }
        }
    }
    __syncthreads();

    // HA: row sums of AB for the tile's columns
    for (int k = tid; k < ABH * TX; k += NT) {
        const int j = k / TX, i = k - j * TX;
        float s[12];
#pragma unroll
        for (int c = 0; c < 12; c++) s[c] = 0.0f;
#pragma unroll
        for (int d = -R; d <= R; d++) {
#pragma unroll
            for (int c = 0; c < 12; c++) s[c] += AB[(c * ABH + j) * ABW + i + R + d];
        }
#pragma unroll
        for (int c = 0; c < 12; c++) HH[(c * ABH + j) * TX + i] = s[c] / (float)FS;
    }
    __syncthreads();

    // column sums of HA and computeResult (cl/guided_filter.cl:139-154)
    for (int k = tid; k < TY * TX; k += NT) {
        const int j = k / TX, i = k - j * TX;
        const int x = x0 + i, y = y0 + j;
        if (x >= w || y >= h) continue;
        float a[12];
#pragma unroll
        for (int c = 0; c < 12; c++) a[c] = 0.0f;
#pragma unroll
        for (int d = -R; d <= R; d++) {
#pragma unroll
            for (int c = 0; c < 12; c++) a[c] += HH[(c * ABH + j + R + d) * TX + i];
        }
#pragma unroll
        for (int c = 0; c < 12; c++) a[c] = a[c] / (float)FS;
        const int o = (j + 2 * R) * INW + i + 2 * R;
        const float p0 = IN[o], p1 = IN[INW * INH + o], p2 = IN[2 * INW * INH + o];
        float q[3];
#pragma unroll
        for (int c = 0; c < 3; c++) q[c] = a[c] * p0 + a[4 + c] * p1 + a[8 + c] * p2;
        q[0] += a[3]; q[1] += a[7]; q[2] += a[11];
        st4_<H>(out, x, y, make_float4(q[0], q[1], q[2], 0.0f));
    }
}

bool guided_filter_fused(int fs) { return fs == 3 || fs == 5; }

void launch_guided_filter_fused(const Img& in, const Img& out, int fs, const ParamRef& eps, hipStream_t st) {
    dim3 grd((in.w + gp::TX - 1) / gp::TX, (in.h + gp::TY - 1) / gp::TY);
    with_storage(in, [&](auto H) {  // in and out share their storage
        if (fs == 5)
            hipLaunchKernelGGL((k_guided_precise<2, H()>), grd, dim3(gp::NT), 0, st, in, out, eps);
        else
            hipLaunchKernelGGL((k_guided_precise<1, H()>), grd, dim3(gp::NT), 0, st, in, out, eps);
    });
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
