// k_rawrgba.hip - gfx950 kernels for the high-noise raw branch of
// RawConverter::demosaicImage (noiseLevel > 0.6, cpp/raw_converter.cpp:111-123):
// bayerToRawRGBA, despeckleRawRGBAImage, denoiseRawRGBAImage and
// rawRGBAToBayer (cl/demosaic.cl:898-929, 530-553, 932-949, 1187-1201).
//
// A raw RGBA image holds one (R, G, B, G2) quad per texel, w/2 x h/2, and
// reads clamp to the quad grid.  The pipeline runs the four operators as one
// kernel: each 32x8-quad block stages its quads (scaled from the uint16 mosaic)
// with a 3-quad halo in LDS, despeckles the block plus a 2-quad halo into a
// second LDS tile, denoises its own quads from that tile and writes them back
// into the FP32 mosaic that interpolateGreen / interpolateRedBlue read.
#include "kernels.h"

namespace clim {

namespace {
constexpr int QX = 32, QY = 8;  // output quads per block, one per thread
enum { kDespeckle = 0, kDenoise = 1, kFused = 2 };
}  // namespace

// despeckle_3x3x4 (cl/demosaic.cl:530-553): the reference's running second
// extremes, trackers starting at 0 and 65535, the centre clamped to
// [second min, second max].  Comparisons and selects only.
__device__ __forceinline__ void track2(float v, float& fmx, float& smx, float& fmn, float& smn) {
    smx = (v <= fmx && v > smx) ? v : smx;
    smx = v > fmx ? fmx : smx;
    fmx = v > fmx ? v : fmx;
    smn = (v >= fmn && v < smn) ? v : smn;
    smn = v < fmn ? fmn : smn;
    fmn = v < fmn ? v : fmn;
}

template <class G>
__device__ __forceinline__ float4 despeckle_quad(const G& g, int x, int y) {
    float fmx[4] = {0, 0, 0, 0}, smx[4] = {0, 0, 0, 0};
    float fmn[4] = {65535.0f, 65535.0f, 65535.0f, 65535.0f}, smn[4] = {65535.0f, 65535.0f, 65535.0f, 65535.0f};
#pragma unroll
    for (int dx = -1; dx <= 1; dx++) {
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const float4 v = g(x + dx, y + dy);
            track2(v.x, fmx[0], smx[0], fmn[0], smn[0]);
            track2(v.y, fmx[1], smx[1], fmn[1], smn[1]);
            track2(v.z, fmx[2], smx[2], fmn[2], smn[2]);
            track2(v.w, fmx[3], smx[3], fmn[3], smn[3]);
        }
    }
    const float4 c = g(x, y);
    return make_float4(clampf(c.x, smn[0], smx[0]), clampf(c.y, smn[1], smx[1]), clampf(c.z, smn[2], smx[2]),
                       clampf(c.w, smn[3], smx[3]));
}

// denoiseRawRGBA (cl/demosaic.cl:932-949): the 5x5 taps (y outer) within
// distance 1 of the centre in units of sigma = sqrt(rawVariance * centre),
// as sum (s - c)^2 / sigma^2 < 1 (clim_raw_dist4), averaged.  A zero channel
// makes every distance NaN or infinite: 0/0, as in the reference.
template <class G>
__device__ __forceinline__ float4 denoise_quad(const G& g, int x, int y, float4 var) {
    const float4 p = g(x, y);
    const float k0 = clim_inv_var(var.x * p.x), k1 = clim_inv_var(var.y * p.y);
    const float k2 = clim_inv_var(var.z * p.z), k3 = clim_inv_var(var.w * p.w);
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float norm = 0.0f;
#pragma unroll
    for (int dy = -2; dy <= 2; dy++) {
#pragma unroll
        for (int dx = -2; dx <= 2; dx++) {
            const float4 s = g(x + dx, y + dy);
            const float l = clim_raw_dist4(s.x - p.x, s.y - p.y, s.z - p.z, s.w - p.w, k0, k1, k2, k3);
            const float wgt = l < 1.0f ? 1.0f : 0.0f;
            acc.x = fmaf(wgt, s.x, acc.x);
            acc.y = fmaf(wgt, s.y, acc.y);
            acc.z = fmaf(wgt, s.z, acc.z);
            acc.w = fmaf(wgt, s.w, acc.w);
            norm += wgt;
        }
    }
    return make_float4(acc.x / norm, acc.y / norm, acc.z / norm, acc.w / norm);
}

// an LDS tile of quads whose texel (i, j) holds quad (ox + i, oy + j)
template <int W>
struct QuadTileView {
    const float4 (*t)[W];
    int ox, oy;
    __device__ __forceinline__ float4 operator()(int x, int y) const { return t[y - oy][x - ox]; }
};

// quad (qx, qy) of a mosaic, in range: scaled on the fly from uint16, or read
// from the FP32 mosaic
template <bool U16>
__device__ __forceinline__ float4 mosaic_quad(const MosaicArg& m, int qx, int qy) {
    if (U16) return RawQuadLoader{m.im, m.b, m.smul, m.black}.quad(qx, qy);
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = row_f(m.im, 2 * qy + m.b.oy[c])[2 * qx + m.b.ox[c]];
    return make_float4(v[0], v[1], v[2], v[3]);
}

// MODE kDespeckle / kDenoise: RGBA raw image `in` -> `out`.  kFused: the
// mosaic `m` -> despeckle -> denoise -> the FP32 mosaic `out`.
template <int MODE, bool U16>
__global__ void __launch_bounds__(256) k_raw_rgba_filter(Img in, MosaicArg m, Img out, ParamRef var) {
    constexpr int HA = MODE == kDespeckle ? 1 : MODE == kDenoise ? 2 : 3;
    constexpr int AW = QX + 2 * HA, AH = QY + 2 * HA;
    constexpr int BW = QX + 4, BH = QY + 4;  // kFused: the despeckled quads with the denoise halo
    __shared__ float4 A[AH][AW];
    __shared__ float4 B[MODE == kFused ? BH : 1][MODE == kFused ? BW : 1];
    const TileId tl = xcd_tile();
    const int qx0 = tl.x * QX, qy0 = tl.y * QY;
    const int qw = MODE == kFused ? m.im.w / 2 : in.w, qh = MODE == kFused ? m.im.h / 2 : in.h;
    const int tid = threadIdx.x + threadIdx.y * QX;
    // A[j][i] = quad (clamp(qx0 - HA + i), clamp(qy0 - HA + j))
    stage<AW * AH, 256>(
        tid,
        [&](int i) {
            const int gx = clampi(qx0 - HA + i % AW, 0, qw - 1), gy = clampi(qy0 - HA + i / AW, 0, qh - 1);
            return MODE == kFused ? mosaic_quad<U16>(m, gx, gy) : row_f4(in, gy)[gx];
        },
        [&](int i, float4 v) { A[i / AW][i % AW] = v; });
    __syncthreads();
    const QuadTileView<AW> va{A, qx0 - HA, qy0 - HA};
    const int x = qx0 + threadIdx.x, y = qy0 + threadIdx.y;
    if constexpr (MODE == kFused) {
        // B[j][i] = despeckled quad clamp(qx0 - 2 + i, qy0 - 2 + j): its 3x3
        // window lies in A (A holds the clamped quads around it)
        for (int i = tid; i < BW * BH; i += 256) {
            const int bx = clampi(qx0 - 2 + i % BW, 0, qw - 1), by = clampi(qy0 - 2 + i / BW, 0, qh - 1);
            B[i / BW][i % BW] = despeckle_quad(va, bx, by);
        }
        __syncthreads();
        if (x >= qw || y >= qh) return;
        const float4 vr = make_float4(var.get(0), var.get(1), var.get(2), var.get(3));
        const float4 r = denoise_quad(QuadTileView<BW>{B, qx0 - 2, qy0 - 2}, x, y, vr);
        // rawRGBAToBayer: the quad's two rows as float2 stores
        const float rv[4] = {r.x, r.y, r.z, r.w};
        float t[2][2] = {{0.0f, 0.0f}, {0.0f, 0.0f}};
#pragma unroll
        for (int c = 0; c < 4; c++) {
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int i = 0; i < 2; i++)
                    if (m.b.oy[c] == j && m.b.ox[c] == i) t[j][i] = rv[c];
        }
        reinterpret_cast<float2*>(wrow_f(out, 2 * y))[x] = make_float2(t[0][0], t[0][1]);
        reinterpret_cast<float2*>(wrow_f(out, 2 * y + 1))[x] = make_float2(t[1][0], t[1][1]);
        return;
    }
    if (x >= qw || y >= qh) return;
    if constexpr (MODE == kDespeckle) {
        wrow_f4(out, y)[x] = despeckle_quad(va, x, y);
    } else {
        const float4 vr = make_float4(var.get(0), var.get(1), var.get(2), var.get(3));
        wrow_f4(out, y)[x] = denoise_quad(va, x, y, vr);
    }
}

// bayerToRawRGBA / rawRGBAToBayer (cl/demosaic.cl:898-929): re-layouts
__global__ void __launch_bounds__(256) k_bayer_to_raw_rgba(Img raw, Img rgba, Bayer b) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= rgba.w || y >= rgba.h) return;
    float v[4];
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = row_f(raw, 2 * y + b.oy[c])[2 * x + b.ox[c]];
    wrow_f4(rgba, y)[x] = make_float4(v[0], v[1], v[2], v[3]);
}

__global__ void __launch_bounds__(256) k_raw_rgba_to_bayer(Img rgba, Img raw, Bayer b) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= rgba.w || y >= rgba.h) return;
    const float4 q = row_f4(rgba, y)[x];
    const float v[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int c = 0; c < 4; c++) wrow_f(raw, 2 * y + b.oy[c])[2 * x + b.ox[c]] = v[c];
}

void launch_bayer_to_raw_rgba(const Img& raw, const Img& rgba, const Bayer& b, hipStream_t st) {
    dim3 blk(64, 4), grd((rgba.w + 63) / 64, (rgba.h + 3) / 4);
    hipLaunchKernelGGL(k_bayer_to_raw_rgba, grd, blk, 0, st, raw, rgba, b);
}

void launch_raw_rgba_to_bayer(const Img& rgba, const Img& raw, const Bayer& b, hipStream_t st) {
    dim3 blk(64, 4), grd((rgba.w + 63) / 64, (rgba.h + 3) / 4);
    hipLaunchKernelGGL(k_raw_rgba_to_bayer, grd, blk, 0, st, rgba, raw, b);
}

void launch_despeckle_raw_rgba(const Img& in, const Img& out, hipStream_t st) {
    dim3 blk(QX, QY), grd((in.w + QX - 1) / QX, (in.h + QY - 1) / QY);
    hipLaunchKernelGGL((k_raw_rgba_filter<kDespeckle, false>), grd, blk, 0, st, in, MosaicArg{}, out, ParamRef{});
}

void launch_denoise_raw_rgba(const Img& in, const ParamRef& var, const Img& out, hipStream_t st) {
    dim3 blk(QX, QY), grd((in.w + QX - 1) / QX, (in.h + QY - 1) / QY);
    hipLaunchKernelGGL((k_raw_rgba_filter<kDenoise, false>), grd, blk, 0, st, in, MosaicArg{}, out, var);
}

void launch_raw_predenoise(const MosaicArg& m, const ParamRef& var, const Img& out, hipStream_t st) {
    const int qw = m.im.w / 2, qh = m.im.h / 2;
    dim3 blk(QX, QY), grd((qw + QX - 1) / QX, (qh + QY - 1) / QY);
    const Img none{nullptr, 0, 0, 0};
    if (m.u16)
        hipLaunchKernelGGL((k_raw_rgba_filter<kFused, true>), grd, blk, 0, st, none, m, out, var);
    else
        hipLaunchKernelGGL((k_raw_rgba_filter<kFused, false>), grd, blk, 0, st, none, m, out, var);
}

}  // namespace clim
