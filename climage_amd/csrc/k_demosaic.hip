// k_demosaic.hip - gfx950 kernels for the demosaic half of the RAW->RGB path:
// scaleRawData, rawNoiseStatistics, interpolateGreen, interpolateRedBlue
// (optionally fused with blendHighlightsImage and the camera->YCbCr
// transformImage), despeckleLumaMedianChromaImage, and the stand-alone
// pointwise operators.
//
// Reference: cl/demosaic.cl (cl/ = CLImage/app/src/main/cpp/OpenCL/); line
// numbers are cited per kernel.  All loads clamp to the image.
#include <type_traits>

#include "fit_acc.h"
#include "kernels.h"
#include "tiles.h"

namespace clim {

// ------------------------------------------------------------ scaleRawData
// cl/demosaic.cl:69-78: max(scaleMul[c] * (raw/65535 - black), 0).
// One thread per pixel (the reference works on quads; the per-site
// arithmetic is the same).
__global__ void __launch_bounds__(256) k_scale_raw(Img raw, Img out, Bayer b, float4 smul, float black) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const uint16_t v16 = reinterpret_cast<const uint16_t*>(raw.p + (size_t)y * raw.pitch)[x];
    wrow_f(out, y)[x] = scale_texel(v16, b.site(x, y), smul, black);
}

void launch_scale_raw(const Img& raw, const Img& out, const Bayer& b, const float smul[4], float black,
                      hipStream_t st) {
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    hipLaunchKernelGGL(k_scale_raw, grd, blk, 0, st, raw, out, b, make_float4(smul[0], smul[1], smul[2], smul[3]),
                       black);
}

// ------------------------------------------------------ rawNoiseStatistics
// cl/demosaic.cl:1288-1324.  Output grid (W/2) x (H/2); per quad the 5x5-quad
// sums of each CFA channel in (R, G, B, G2) order, then
// mean = sum/25, var = (sumSq - sum^2/25)/25.
// The quads of a 16x16-quad tile plus a 2-quad halo are staged in LDS as
// float4 (one quad per texel).
__global__ void __launch_bounds__(256) k_raw_stats(Img scaled, Bayer b, Img mean, Img var) {
    constexpr int T = 16, R = 2, S = T + 2 * R;
    __shared__ QuadTile<T> sm;
    const TileId tl = xcd_tile();
    const int qx0 = tl.x * T, qy0 = tl.y * T;
    const int tid = threadIdx.y * T + threadIdx.x;
    stage<S * S, 256>(
        tid,
        [&](int i) {
            const int qx = qx0 - R + i % S, qy = qy0 - R + i / S;
            return make_float4(ld1(scaled, 2 * qx + b.ox[0], 2 * qy + b.oy[0]),
                               ld1(scaled, 2 * qx + b.ox[1], 2 * qy + b.oy[1]),
                               ld1(scaled, 2 * qx + b.ox[2], 2 * qy + b.oy[2]),
                               ld1(scaled, 2 * qx + b.ox[3], 2 * qy + b.oy[3]));
        },
        [&](int i, float4 v) { sm.q[i / S][i % S] = v; });
    __syncthreads();
    sm.hsums(tid);
    __syncthreads();
    const int qx = qx0 + threadIdx.x, qy = qy0 + threadIdx.y;
    if (qx >= mean.w || qy >= mean.h) return;
    float m[4], v[4];
    sm.stats(threadIdx.x, threadIdx.y, m, v);
    wrow_f4(mean, qy)[qx] = make_float4(m[0], m[1], m[2], m[3]);
    wrow_f4(var, qy)[qx] = make_float4(v[0], v[1], v[2], v[3]);
}

void launch_raw_stats(const Img& scaled, const Bayer& b, const Img& mean, const Img& var, hipStream_t st) {
    dim3 blk(16, 16), grd((mean.w + 15) / 16, (mean.h + 15) / 16);
    hipLaunchKernelGGL(k_raw_stats, grd, blk, 0, st, scaled, b, mean, var);
}

// The loads of one quad and their conversion to scaled (R, G, B, G2): the two
// 32-bit row words of a uint16 mosaic, or the four texels of an FP32 one;
// coordinates are clamped per texel (MosaicSrc).
template <bool U16>
struct QuadFetch;
template <>
struct QuadFetch<true> {
    using Raw = uint2;
    RawQuadLoader ld;
    __device__ __forceinline__ Raw load(int qx, int qy) const { return ld.words(qx, qy); }
    __device__ __forceinline__ float4 value(Raw w, int qx, int qy) const { return ld.quad(w, qx, qy); }
};
template <>
struct QuadFetch<false> {
    using Raw = float4;
    MosaicSrc<false> src;
    __device__ __forceinline__ Raw load(int qx, int qy) const {
        const Bayer& b = src.b;
        return make_float4(src(2 * qx + b.ox[0], 2 * qy + b.oy[0]), src(2 * qx + b.ox[1], 2 * qy + b.oy[1]),
                           src(2 * qx + b.ox[2], 2 * qy + b.oy[2]), src(2 * qx + b.ox[3], 2 * qy + b.oy[3]));
    }
    __device__ __forceinline__ float4 value(Raw v, int, int) const { return v; }
};

// rawNoiseStatistics with the first regression pass of
// computeRawNoiseStatistics accumulated on the way: each of the grid's blocks
// takes a contiguous run of 16x16-quad tiles and keeps its regression sums
// across them (one partial per block); the quads in range are listed for the
// refit.
// Five waves per SIMD for the uint16 mosaic; the FP32 mosaic's wider loads
// need the registers of four.
template <bool U16>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(U16 ? 5 : 4)))
k_raw_stats_fit(FitJobs jobs, float4 smul, float black) {
    constexpr int T = 16, R = 2, S = T + 2 * R;
    using L = Layout<4, false, false>;
    __shared__ QuadTile<T> sm;
    __shared__ clim_fx128 tot[L::K];
    __shared__ int nlisted;
    static_assert(sizeof(QuadTile<T>) >= 1024 * sizeof(clim_fx128), "the tile storage serves as the reduction buffer");
    const int lb = xcd_block();
    const FitJob& jb = jobs.j[job_of_block(jobs, lb)];
    const int bi = lb - jb.block_base;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const Bayer& b = jb.b;
    const QuadFetch<U16> src{{jb.src, b, smul, black}};
    const int qw = jb.src.w / 2, qh = jb.src.h / 2;  // quads
    const int tiles_x = (qw + T - 1) / T;
    const long long ntiles = jb.ntiles;
    const int t0 = (int)(ntiles * bi / jb.nblocks), t1 = (int)(ntiles * (bi + 1) / jb.nblocks);
    // this block's runs of the sample lists (the first barrier of the loop
    // orders the counter's reset before its first use)
    float4* lm = reinterpret_cast<float4*>(jb.m.p) + (size_t)t0 * (T * T);
    float4* lv = reinterpret_cast<float4*>(jb.v.p) + (size_t)t0 * (T * T);
    if (tid == 0) nlisted = 0;
    Acc<4, false, false> acc;
    acc.zero();
    // the next tile's loads are issued before the current tile's compute
    Staged<S * S, 256, typename QuadFetch<U16>::Raw> pf;
    auto fetch = [&](int t) {
        const int qx0 = (t % tiles_x) * T - R, qy0 = (t / tiles_x) * T - R;
        pf.fetch(tid, [&](int i) { return src.load(qx0 + i % S, qy0 + i / S); });
    };
    if (t0 < t1) fetch(t0);
    for (int t = t0; t < t1; t++) {
        const int qx0 = (t % tiles_x) * T, qy0 = (t / tiles_x) * T;
        pf.put(tid, [&](int i, typename QuadFetch<U16>::Raw r) {
            sm.q[i / S][i % S] = src.value(r, qx0 - R + i % S, qy0 - R + i / S);
        });
        __syncthreads();
        if (t + 1 < t1) fetch(t + 1);
        sm.hsums(tid);
        __syncthreads();
        const int qx = qx0 + tx, qy = qy0 + ty;
        float m[4], v[4];
        sm.stats(tx, ty, m, v);
        const bool ok = qx < qw && qy < qh && Acc<4, false, false>::in_range(m, v);
        const int slot = list_slot(ok, &nlisted);
        if (ok) {
            lm[slot] = make_float4(m[0], m[1], m[2], m[3]);
            lv[slot] = make_float4(v[0], v[1], v[2], v[3]);
        }
        acc.add_tested(m, v, ok);
        __syncthreads();  // the next tile reuses the LDS
    }
    // the last tile is done (the loop ends with a barrier): its LDS is free
    if (tid == 0) jb.counts[bi] = (unsigned int)nlisted;
    store_block_partials(acc, reinterpret_cast<clim_fx128*>(&sm), tot, jb.partials, bi, kFitBlocksPerJob);
}

bool launch_raw_stats_fit(FitJobs& jobs, const float* smul, float black, hipStream_t st) {
    long long tiles[5];
    for (int j = 0; j < jobs.njobs; j++) {
        FitJob& jb = jobs.j[j];
        jb.ntiles = ((jb.src.w / 2 + 15) / 16) * ((jb.src.h / 2 + 15) / 16);
        jb.seg = 16 * 16;
        if ((long long)jb.ntiles * jb.seg > jb.cap) return false;
        tiles[j] = jb.ntiles;
    }
    const int grid = fit_share_blocks(jobs, tiles, kRawStatsGrid);
    if (smul)
        hipLaunchKernelGGL(k_raw_stats_fit<true>, dim3(grid), dim3(256), 0, st, jobs,
                           make_float4(smul[0], smul[1], smul[2], smul[3]), black);
    else
        hipLaunchKernelGGL(k_raw_stats_fit<false>, dim3(grid), dim3(256), 0, st, jobs, make_float4(0, 0, 0, 0), 0.0f);
    return true;
}

// --------------------------------------------------------- interpolateGreen
// cl/demosaic.cl:96-129 and :160-227.  Tile of 32x16 output pixels; the raw
// tile carries a 3-pixel halo, the |Sobel| responses a 2-pixel halo, so every
// Sobel response is computed once per block instead of 26 times per pixel.
namespace {
constexpr int GT_X = 32, GT_Y = 16;
constexpr int GR_X = GT_X + 6, GR_Y = GT_Y + 6;   // raw tile
constexpr int GS_X = GT_X + 4, GS_Y = GT_Y + 4;   // sobel tile
// The LDS tiles are unpadded (rows of 38 and 36 floats).  LDS serves 32
// lanes at a time from 32 four-byte banks.  The staging and Sobel passes
// walk the tiles in order, so 32 consecutive lanes meet 32 consecutive
// words.  In the interpolation pass a half-wave reads two tile rows, each
// at every second column (one non-green pixel per pair) with the column
// parity alternating by row: the two rows' words are 38 +- 1 or 36 +- 1
// apart, an odd number of banks, so one row meets the even banks and the
// other the odd ones.
}

struct GreenLds {
    float r[GR_Y][GR_X];
    float sgx[GS_Y][GS_X];
    float sgy[GS_Y][GS_X];
    float sgl[GS_Y][GS_X];
};

// one GT_X x GT_Y tile at (x0, y0) by the block's 256 threads (tid); returns
// with the LDS still being read by other threads (the caller's next use of
// sm needs a barrier first)
template <bool U16>
__device__ __forceinline__ void green_tile(const MosaicSrc<U16>& raw, const Img& green, int gxo, int gyo,
                                           const ParamRef& lv, int x0, int y0, int tid, GreenLds& sm) {
    auto& r = sm.r;
    auto& sgx = sm.sgx;
    auto& sgy = sm.sgy;
    auto& sgl = sm.sgl;
    // the raw tile in groups of four texels, columns x0 - 4 + 4 k (x0 is a
    // multiple of 4): one wide load per group; the tile keeps x0 - 3 .. x0 + 34
    constexpr int GG = (GR_X + 1 + 3) / 4;
    static_assert(GT_X % 4 == 0 && GR_X == GT_X + 6, "groups of four");
    stage<GG * GR_Y, 256>(
        tid, [&](int i) { return raw.row4(x0 - 4 + 4 * (i % GG), y0 - 3 + i / GG); },
        [&](int i, float4 v) {
            const int j = i / GG, c = 4 * (i % GG) - 1;
            if (c >= 0) r[j][c] = v.x;
            r[j][c + 1] = v.y;
            r[j][c + 2] = v.z;
            if (c + 3 < GR_X) r[j][c + 3] = v.w;
        });
    __syncthreads();
    // sobel(), cl/demosaic.cl:96-112, value / sqrt(4.5) then abs().  The
    // 9-tap sums in their order with the zero taps left out and the +-1, +-2
    // taps as exact adds: every partial sum is the same value (up to the sign
    // of a zero, which the abs() removes).
    // items (sx, sy): the first 32 columns with one row per half-wave (32
    // consecutive words per access), then the last GS_X - 32 columns
    static_assert(GS_X > 32 && 32 * GS_Y + (GS_X - 32) * GS_Y <= 4 * 256, "four rounds");
    for (int i = tid; i < GS_X * GS_Y; i += 256) {
        constexpr int kMain = 32 * GS_Y, kRest = GS_X - 32;
        const int sy = i < kMain ? i >> 5 : (i - kMain) / kRest;
        const int sx = i < kMain ? i & 31 : 32 + (i - kMain) % kRest;
        const float s00 = r[sy][sx], s01 = r[sy][sx + 1], s02 = r[sy][sx + 2];
        const float s10 = r[sy + 1][sx], s12 = r[sy + 1][sx + 2];
        const float s20 = r[sy + 2][sx], s21 = r[sy + 2][sx + 1], s22 = r[sy + 2][sx + 2];
        const float vx = ((((s00 - s02) + 2.0f * s10) - 2.0f * s12) + s20) - s22;
        const float vy = ((((s00 + 2.0f * s01) + s02) - s20) - 2.0f * s21) - s22;
        const float ax = fabsf(vx * 0.471404521f), ay = fabsf(vy * 0.471404521f);
        sgx[sy][sx] = ax;
        sgy[sy][sx] = ay;
        sgl[sy][sx] = len2(ax, ay);
    }
    __syncthreads();
    const float lumaVariance = mmax(lv.get(0), 1e-4f);
    // Thread t takes the horizontal pixel pair k = t % 16 of tile row t / 16:
    // it interpolates the pair's non-green pixel (so every lane of a wave runs
    // the same code) and copies the green one.
    const int ty = tid >> 4, k = tid & 15;
    const int y = y0 + ty;
    if (y >= green.h) return;
    const int cpar = ((y & 1) == (gyo & 1) ? gxo + 1 : gxo) & 1;  // x parity of the non-green site
    const int tx = 2 * k + cpar;                                    // x0 is even
    const int x = x0 + tx;
    const int Y = ty + 3;  // raw tile coordinates
    float out = 0.0f;
    if (x < green.w) {
        const int X = tx + 3;
        const float g_left = r[Y][X - 1], g_right = r[Y][X + 1];
        const float g_up = r[Y - 1][X], g_down = r[Y + 1][X];
        const float c_xy = r[Y][X];
        const float c_left = r[Y][X - 2], c_right = r[Y][X + 2];
        const float c_up = r[Y - 2][X], c_down = r[Y + 2][X];
        const float c2_ave = (r[Y - 1][X - 1] + r[Y - 1][X + 1] + r[Y + 1][X - 1] + r[Y + 1][X + 1]) / 4.0f;
        const float g_ave = (g_left + g_right + g_up + g_down) / 4.0f;
        const float lumaStdDev = sqrtf(lumaVariance * g_ave);

        // bilateralSobel(), cl/demosaic.cl:114-129
        const int SX = tx + 2, SY = ty + 2;  // sobel tile coordinates
        const float cl = sgl[SY][SX];
        // 1 - smoothstep(4 sigma, 16 sigma, 0.05 |dl|) by clim_green_tap_weight
        const float e0 = 4.0f * lumaStdDev, inv = 1.0f / (16.0f * lumaStdDev - 4.0f * lumaStdDev);
        const float c1 = 0.05f * inv, c0 = clim_green_tap_c0(e0, inv);
        float sx = 0.0f, sy = 0.0f, wsum = 0.0f;
#pragma unroll
        for (int j = -2; j <= 2; j++) {
#pragma unroll
            for (int i = -2; i <= 2; i++) {
                const float gx = sgx[SY + j][SX + i], gy = sgy[SY + j][SX + i];
                const float wt = clim_green_tap_weight(fabsf(sgl[SY + j][SX + i] - cl), c1, c0);
                sx = fmaf(wt, gx, sx);
                sy = fmaf(wt, gy, sy);
                wsum += wt;  // the x and y weight sums are equal
            }
        }
        // the two quotients by one reciprocal (the FP32 restatement's form)
        const float iw = 1.0f / wsum;
        const float gradx = sx * iw, grady = sy * iw;

        const float g_lf_x = (g_left + g_right) / 2.0f, g_lf_y = (g_up + g_down) / 2.0f;
        const float g_hf_x = (2.0f * c_xy - (c_left + c_right)) / 4.0f;
        const float g_hf_y = (2.0f * c_xy - (c_up + c_down)) / 4.0f;
        const float whiteness =
            clampf(mmin(c_xy, mmin(g_ave, c2_ave)) / mmax(c_xy, mmax(g_ave, c2_ave)), 0.0f, 1.0f);
        const float glen = len2(gradx, grady);
        const float lgt = smoothstep(lumaStdDev, 4.0f * lumaStdDev, glen);
        // glen / sqrt(2) as a multiplication (the FP32 restatement's form);
        // smoothstep(0, 0.3, .) with the span division as a reciprocal and one
        // fused correction, the same values (clim_smoothstep_rc)
        const float hf_gain =
            lgt * mmin(0.5f * whiteness + clim_smoothstep_rc(0.0f, 0.3f, 1.0f / 0.3f, glen * 0.707106781f), 1.0f);
        const float g_est_x = g_lf_x + hf_gain * g_hf_x;
        const float g_est_y = g_lf_y + hf_gain * g_hf_y;
        float direction = 2.0f * clim_atan2pif(grady, gradx);
        direction = direction < 0.5f
                        ? mixf(direction, 0.0f, 1.0f - smoothstep(0.3f * lgt, 0.45f * lgt, direction))
                        : mixf(direction, 1.0f, smoothstep((1.0f - 0.45f) * lgt, (1.0f - 0.3f) * lgt, direction));
        direction = mixf(0.5f, direction, lgt);
        out = mixf(g_est_y, g_est_x, direction);
    }
    // green sites are copied
    const int xe = x0 + 2 * k;  // even pixel of the pair
    const float copy = r[Y][2 * k + (cpar ^ 1) + 3];
    const float ev = cpar ? copy : out, od = cpar ? out : copy;
    float* row = wrow_f(green, y);
    if (xe + 1 < green.w && (green.pitch & 7) == 0) {
        *reinterpret_cast<float2*>(row + xe) = make_float2(ev, od);
    } else {
        if (xe < green.w) row[xe] = ev;
        if (xe + 1 < green.w) row[xe + 1] = od;
    }
}

template <bool U16>
__global__ void __launch_bounds__(256) k_green(MosaicSrc<U16> raw, Img green, int gxo, int gyo, ParamRef lv) {
    __shared__ GreenLds sm;
    const TileId tl = xcd_tile();
    green_tile<U16>(raw, green, gxo, gyo, lv, tl.x * GT_X, tl.y * GT_Y, threadIdx.y * 32 + threadIdx.x, sm);
}

void launch_green(const MosaicArg& raw, const Img& green, int bayer, const ParamRef& lv, hipStream_t st) {
    // raw_green site of the OpenCL table
    const int gx[4] = {0, 0, 0, 0}, gy[4] = {0, 0, 1, 1};
    dim3 blk(32, 8), grd((green.w + GT_X - 1) / GT_X, (green.h + GT_Y - 1) / GT_Y);
    if (raw.u16)
        hipLaunchKernelGGL(k_green<true>, grd, blk, 0, st, raw.src<true>(), green, gx[bayer], gy[bayer], lv);
    else
        hipLaunchKernelGGL(k_green<false>, grd, blk, 0, st, raw.src<false>(), green, gx[bayer], gy[bayer], lv);
}

// ------------------------------------------------------- interpolateRedBlue
// cl/demosaic.cl:229-318, optionally followed by blendHighlightsImage
// (cl/demosaic.cl:338-380, clip) and transformImage (cl/demosaic.cl:728-733)
// on the same pixel.  chroma variance = (B_R + B_B) / 2 of the raw noise
// model (cpp/raw_converter.cpp:127-128) when taken from the device.
struct Mat3 { float m[9]; };

// interpolateRedBlue of one 2x2 quad at (x0, y0) (even): the 4x4 windows of
// the mosaic and of the green plane around it are loaded once and the four
// sites computed in turn; with kFuse, blendHighlights and the camera->YCbCr
// transform follow on each pixel.  out[dy][dx] = (R, G, B) or (Y, Cb, Cr).
// Which site is R, G, B or G2 is the same for every quad, so every lane of a
// wave runs the same code.
// The quad's 4x4 windows: R[j][i], G[j][i] are the mosaic and green at
// (x0 - 1 + i, y0 - 1 + j), clamped to the image.
template <bool kFuse>
__device__ __forceinline__ void red_blue_from(const float (&R)[4][4], const float (&G)[4][4], const Bayer& b,
                                              int red_is_c1_site, float chromaVariance, float clip, const Mat3& m,
                                              float3 out[2][2]) {
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const int color = b.site(dx, dy);  // x0, y0 are even
            const int X = 1 + dx, Y = 1 + dy;
            const float g = G[Y][X];
            float red, blue;
            if (color == 0 || color == 2) {
                const float c1 = R[Y][X];
                const float c2_tl = G[Y - 1][X - 1] - R[Y - 1][X - 1];
                const float c2_tr = G[Y - 1][X + 1] - R[Y - 1][X + 1];
                const float c2_bl = G[Y + 1][X - 1] - R[Y + 1][X - 1];
                const float c2_br = G[Y + 1][X + 1] - R[Y + 1][X + 1];
                const float chromaStdDev = sqrtf(chromaVariance * g);
                const float dvx = fabsf(c2_tl - c2_br), dvy = fabsf(c2_tr - c2_bl);
                const float flatness = 1.0f - smoothstep(0.25f * chromaStdDev, chromaStdDev, len2(dvx, dvy));
                const float alpha = mixf(2.0f * clim_atan2pif(dvy, dvx), 0.5f, flatness);
                const float c2 = g - mixf((c2_tr + c2_bl) / 2.0f, (c2_tl + c2_br) / 2.0f, alpha);
                if (color == 0) { red = c1; blue = c2; } else { blue = c1; red = c2; }
            } else {
                const float c1_left = G[Y][X - 1] - R[Y][X - 1];
                const float c1_right = G[Y][X + 1] - R[Y][X + 1];
                const float c2_up = G[Y - 1][X] - R[Y - 1][X];
                const float c2_down = G[Y + 1][X] - R[Y + 1][X];
                const float c1 = g - (c1_left + c1_right) / 2.0f;
                const float c2 = g - (c2_up + c2_down) / 2.0f;
                if (color == red_is_c1_site) { red = c1; blue = c2; } else { blue = c1; red = c2; }
            }
            float px = red, py = g, pz = blue;
            if (kFuse) {
                blend_highlights_px(px, py, pz, clip);
                const float tx = m.m[0] * px + m.m[1] * py + m.m[2] * pz;
                const float ty = m.m[3] * px + m.m[4] * py + m.m[5] * pz;
                const float tz = m.m[6] * px + m.m[7] * py + m.m[8] * pz;
                px = tx; py = ty; pz = tz;
            }
            out[dy][dx] = make_float3(px, py, pz);
        }
    }
}

template <bool kFuse, bool U16>
__device__ __forceinline__ void red_blue_quad(const MosaicSrc<U16>& raw, const Img& green, int x0, int y0,
                                              const Bayer& b, int red_is_c1_site, float chromaVariance, float clip,
                                              const Mat3& m, float3 out[2][2]) {
    float R[4][4], G[4][4];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            R[j][i] = raw(x0 - 1 + i, y0 - 1 + j);
            G[j][i] = ld1(green, x0 - 1 + i, y0 - 1 + j);
        }
    red_blue_from<kFuse>(R, G, b, red_is_c1_site, chromaVariance, clip, m, out);
}

__device__ __forceinline__ float chroma_variance(const ParamRef& cv) {
    const float v = cv.dev ? (cv.dev[4] + cv.dev[6]) / 2.0f : cv.host[0];
    return mmax(v, 1e-6f);
}

// one quad per thread
template <bool kFuse, bool U16>
__global__ void __launch_bounds__(256) k_red_blue(MosaicSrc<U16> raw, Img green, Img out, Bayer b,
                                                  int red_is_c1_site, ParamRef cv, int rotate_180, float clip,
                                                  Mat3 m) {
    const TileId tl = xcd_tile();
    const int qx = tl.x * 64 + threadIdx.x, qy = tl.y * 4 + threadIdx.y;
    const int x0 = 2 * qx, y0 = 2 * qy;
    if (x0 >= out.w || y0 >= out.h) return;
    float3 q[2][2];
    red_blue_quad<kFuse, U16>(raw, green, x0, y0, b, red_is_c1_site, chroma_variance(cv), clip, m, q);
#pragma unroll
    for (int dy = 0; dy < 2; dy++) {
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const int x = x0 + dx, y = y0 + dy;
            if (x >= out.w || y >= out.h) continue;
            int ox = x, oy = y;
            if (rotate_180) {
                ox = out.w - x;
                oy = out.h - y;
                if (ox >= out.w || oy >= out.h) continue;
            }
            st4(out, ox, oy, make_float4(q[dy][dx].x, q[dy][dx].y, q[dy][dx].z, 0.0f));
        }
    }
}

void launch_red_blue(const MosaicArg& raw, const Img& green, const Img& out, int bayer, const Bayer& b,
                     const ParamRef& cv, int rotate_180, bool fuse, float clip, const float c2y[9],
                     hipStream_t st) {
    // cl/demosaic.cl:301: at green sites red is the horizontal difference on
    // raw_green for bggr/grbg, on raw_green2 otherwise
    const int red_is_c1 = (bayer == 3 || bayer == 0) ? 1 : 3;
    Mat3 m;
    for (int i = 0; i < 9; i++) m.m[i] = c2y ? c2y[i] : (i % 4 == 0 ? 1.0f : 0.0f);
    dim3 blk(64, 4), grd((out.w + 127) / 128, (out.h + 7) / 8);  // one quad per thread
    if (fuse && raw.u16)
        hipLaunchKernelGGL((k_red_blue<true, true>), grd, blk, 0, st, raw.src<true>(), green, out, b, red_is_c1, cv,
                           rotate_180, clip, m);
    else if (fuse)
        hipLaunchKernelGGL((k_red_blue<true, false>), grd, blk, 0, st, raw.src<false>(), green, out, b, red_is_c1, cv,
                           rotate_180, clip, m);
    else if (raw.u16)
        hipLaunchKernelGGL((k_red_blue<false, true>), grd, blk, 0, st, raw.src<true>(), green, out, b, red_is_c1, cv,
                           rotate_180, clip, m);
    else
        hipLaunchKernelGGL((k_red_blue<false, false>), grd, blk, 0, st, raw.src<false>(), green, out, b, red_is_c1,
                           cv, rotate_180, clip, m);
}

// Row 0 and column 0 to zero: the texels a rotated interpolateRedBlue never
// writes (its target of source (x, y) is (w - x, h - y), kept when inside)
__global__ void __launch_bounds__(256) k_zero_row0_col0(Img out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (i < out.w) st4(out, i, 0, z);
    if (i < out.h) st4(out, 0, i, z);
}

void launch_zero_row0_col0(const Img& out, hipStream_t st) {
    const int n = out.w > out.h ? out.w : out.h;
    hipLaunchKernelGGL(k_zero_row0_col0, dim3((n + 255) / 256), dim3(256), 0, st, out);
}

// ------------------------------------------- stand-alone pointwise operators
__global__ void __launch_bounds__(256) k_blend_highlights(Img in, Img out, float clip) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    float4 p = row_f4(in, y)[x];
    blend_highlights_px(p.x, p.y, p.z, clip);
    wrow_f4(out, y)[x] = make_float4(p.x, p.y, p.z, 0.0f);
}

__global__ void __launch_bounds__(256) k_transform(Img in, Img out, Mat3 m) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    const float4 p = row_f4(in, y)[x];
    const float tx = m.m[0] * p.x + m.m[1] * p.y + m.m[2] * p.z;
    const float ty = m.m[3] * p.x + m.m[4] * p.y + m.m[5] * p.z;
    const float tz = m.m[6] * p.x + m.m[7] * p.y + m.m[8] * p.z;
    wrow_f4(out, y)[x] = make_float4(tx, ty, tz, 0.0f);
}

void launch_blend_highlights(const Img& in, const Img& out, float clip, hipStream_t st) {
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    hipLaunchKernelGGL(k_blend_highlights, grd, blk, 0, st, in, out, clip);
}

void launch_transform(const Img& in, const Img& out, const float mat[9], hipStream_t st) {
    Mat3 m;
    for (int i = 0; i < 9; i++) m.m[i] = mat[i];
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    hipLaunchKernelGGL(k_transform, grd, blk, 0, st, in, out, m);
}

// ----------------------------------------- despeckleLumaMedianChromaImage
// cl/demosaic.cl:384-413 (forgetful-selection median of Cb and Cr) and
// :443-479 (luma clamped between its noise-adaptive second extremes), in FP32.
#define CLIM_S(a, b) { const float t_ = a; a = mmin(a, b); b = mmax(t_, b); }

__device__ __forceinline__ float median9(float v0, float v1, float v2, float v3, float v4, float v5, float v6,
                                         float v7, float v8) {
    float a0 = v0, a1 = v1, a2 = v2, a3 = v3, a4 = v4, a5 = v5;
    CLIM_S(a0, a1) CLIM_S(a2, a3) CLIM_S(a4, a5) CLIM_S(a0, a2) CLIM_S(a1, a3) CLIM_S(a0, a4) CLIM_S(a3, a5)
    a0 = v6;
    CLIM_S(a0, a1) CLIM_S(a2, a3) CLIM_S(a0, a2) CLIM_S(a1, a3) CLIM_S(a0, a4) CLIM_S(a3, a4)
    a0 = v7;
    CLIM_S(a0, a1) CLIM_S(a2, a3) CLIM_S(a0, a2) CLIM_S(a1, a3)
    a0 = v8;
    CLIM_S(a0, a1) CLIM_S(a0, a2) CLIM_S(a1, a2)
    return a1;
}

// The median of nine values without NaN, as the selection network above
// returns it: sort three rows of three, then the median of (the largest row
// minimum, the median of the row medians, the smallest row maximum).  Equal
// to median9 for every NaN-free input (checked exhaustively over {0..3}^9);
// v_min3 / v_med3 / v_max3 do each sort of three.
__device__ __forceinline__ float med3f(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }
__device__ __forceinline__ float median9_nonan(const float* v) {
    float lo[3], md[3], hi[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const float a = v[3 * r], b = v[3 * r + 1], c = v[3 * r + 2];
        lo[r] = fminf(fminf(a, b), c);
        md[r] = med3f(a, b, c);
        hi[r] = fmaxf(fmaxf(a, b), c);
    }
    return med3f(fmaxf(fmaxf(lo[0], lo[1]), lo[2]), med3f(md[0], md[1], md[2]), fminf(fminf(hi[0], hi[1]), hi[2]));
}

// One pixel; fetch(dx, dy) returns the YCbCr texel at offset (dx, dy),
// clamped to the image.  NONAN (the caller knows the neighbourhood holds no
// NaN): the luma trackers keep the two largest of {0, 0, Y...} and the two
// smallest of {65535, 65535, Y...} with max / min and a median of three (a
// running clamp), and the chroma medians come from median9_nonan: the same
// values as the reference's comparisons for every NaN-free input (a zero's
// sign aside), in about a third of the instructions.
template <bool NONAN = false, typename Fetch>
__device__ __forceinline__ float4 despeckle_px(Fetch fetch, float var_a, float var_b) {
    // read order of fast_median3x3
    const int ox[9] = {0, 1, 0, 1, 0, 1, -1, -1, -1};
    const int oy[9] = {-1, -1, 0, 0, 1, 1, 1, 0, -1};
    float sample = 0.0f, firstMax = 0.0f, secMax = 0.0f, firstMin = 65535.0f, secMin = 65535.0f;
    float cb[9], cr[9];
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const float4 p = fetch(ox[k], oy[k]);
        const float v = p.x;
        if (NONAN) {
            secMax = med3f(secMax, v, firstMax);
            firstMax = fmaxf(firstMax, v);
            secMin = med3f(firstMin, v, secMin);
            firstMin = fminf(firstMin, v);
        } else {
            secMax = v <= firstMax && v > secMax ? v : secMax;
            secMax = v > firstMax ? firstMax : secMax;
            firstMax = v > firstMax ? v : firstMax;
            secMin = v >= firstMin && v < secMin ? v : secMin;
            secMin = v < firstMin ? firstMin : secMin;
            firstMin = v < firstMin ? v : firstMin;
        }
        if (k == 2) sample = v;
        cb[k] = p.y;
        cr[k] = p.z;
    }
    const float mcb = NONAN ? median9_nonan(cb) : median9(cb[0], cb[1], cb[2], cb[3], cb[4], cb[5], cb[6], cb[7], cb[8]);
    const float mcr = NONAN ? median9_nonan(cr) : median9(cr[0], cr[1], cr[2], cr[3], cr[4], cr[5], cr[6], cr[7], cr[8]);
    const float sigma = sqrtf(var_a + var_b * sample);
    const float minVal = mixf(secMin, firstMin, smoothstep(sigma, 4.0f * sigma, secMin - firstMin));
    const float maxVal = mixf(secMax, firstMax, smoothstep(sigma, 4.0f * sigma, firstMax - secMax));
    sample = clampf(sample, minVal, maxVal);
    return make_float4(sample, mcb, mcr, 0.0f);
}

__global__ void __launch_bounds__(256) k_despeckle(Img in, Img out, float var_a, float var_b) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x;
    const int y = tl.y * 4 + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    st4(out, x, y, despeckle_px([&](int dx, int dy) { return ld4(in, x + dx, y + dy); }, var_a, var_b));
}

void launch_despeckle(const Img& in, const Img& out, float var_a, float var_b, hipStream_t st) {
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    hipLaunchKernelGGL(k_despeckle, grd, blk, 0, st, in, out, var_a, var_b);
}

// ------------------------------ interpolateRedBlue + ... + despeckleImage
// The pipeline's interpolateRedBlue (with blendHighlights and the
// camera->YCbCr transform) and despeckleImage in one kernel: a 64x24 output
// tile's YCbCr values plus a 2-pixel margin (34x14 quads) are computed into
// LDS, and the despeckle reads its 3x3 neighbourhoods there, at the clamped
// image coordinates the stand-alone kernels would read.  The level-0 YCbCr
// image between the two is never written.  The quads' 4x4 windows come from
// the scaled mosaic and the green plane of the region (70x30 texels, clamped
// coordinates) staged in LDS first, each texel loaded and scaled once.
namespace {
constexpr int RD_X = 64, RD_Y = 24;                 // output tile
constexpr int RD_RX = RD_X + 4, RD_RY = RD_Y + 4;   // YCbCr region (even origin)
constexpr int RD_QX = RD_RX / 2, RD_QY = RD_RY / 2;  // its quads
constexpr int RD_WX = RD_RX + 2, RD_WY = RD_RY + 2;  // mosaic / green window of the region
constexpr int RD_QPT = (RD_QX * RD_QY + 255) / 256;  // quads per thread
// A window row in LDS: its even and its odd columns as two planes of
// RD_PS floats, RD_WS floats per row.  Quad (qx, qy) reads column 2 qx + i
// of row 2 qy + j, i.e. word (2 qy + j) RD_WS + (i & 1) RD_PS + qx + i / 2;
// two rows are 162 words, 2 mod 32, so the bank is q mod 32 plus a
// constant (q = 34 qy + qx) and the 32 consecutive quads LDS serves at a
// time meet 32 distinct banks.  The staging pass writes consecutive
// columns, alternately to the two planes, 40 words apart (a 2-way overlap
// of the even and odd lanes' banks on the writes, once per texel).
// The YCbCr region is three planes of floats (alpha is 0 and not kept):
// 22.8 KB, so seven blocks fit a CU's LDS (the window takes 19.4 KB).
constexpr int RD_PS = 40, RD_WS = 81;
static_assert(RD_WX / 2 <= RD_PS && RD_PS + (RD_WX - 1) / 2 < RD_WS && RD_QX == 34 && (2 * RD_WS) % 32 == 2,
              "window planes");
__device__ __forceinline__ int rd_col(int i) { return (i & 1) * RD_PS + (i >> 1); }
}

union RbDespeckleLds {
    struct {
        float r[RD_WY][RD_WS], g[RD_WY][RD_WS];
    } win;
    float t[3][RD_RY][RD_RX];
    __device__ __forceinline__ float4 at(int y, int x) const {
        return make_float4(t[0][y][x], t[1][y][x], t[2][y][x], 0.0f);
    }
    __device__ __forceinline__ void set(int y, int x, float3 v) {
        t[0][y][x] = v.x;
        t[1][y][x] = v.y;
        t[2][y][x] = v.z;
    }
};

// one RD_X x RD_Y output tile at (x0, y0) by the block's 256 threads; as
// green_tile, returns with the LDS possibly still being read
template <bool U16, bool H>
__device__ __forceinline__ void rb_despeckle_tile(const MosaicSrc<U16>& raw, const Img& green, const Img& out,
                                                  const Bayer& b, int red_is_c1_site, const ParamRef& cv, float clip,
                                                  const Mat3& m, float var_a, float var_b, int x0, int y0, int tid,
                                                  RbDespeckleLds& sm) {
    const int rx0 = x0 - 2, ry0 = y0 - 2;
    const float cvar = chroma_variance(cv);
    // the region's mosaic (scaled) and green, texel (rx0 - 1 + i, ry0 - 1 + j),
    // in groups of four texels at columns rx0 - 2 + 4 k (x0 is a multiple of
    // 4): one wide load of each image per group
    struct Group {
        float4 r, g;
    };
    constexpr int WG = (RD_WX + 1 + 3) / 4;
    static_assert(RD_X % 4 == 0, "groups of four");
    stage<WG * RD_WY, 256>(
        tid,
        [&](int k) {
            const int xg = rx0 - 2 + 4 * (k % WG), y = ry0 - 1 + k / WG;
            return Group{raw.row4(xg, y), ld1x4(green, xg, y)};
        },
        [&](int k, const Group& v) {
            const int j = k / WG, c = 4 * (k % WG) - 1;
            const float rv[4] = {v.r.x, v.r.y, v.r.z, v.r.w}, gv[4] = {v.g.x, v.g.y, v.g.z, v.g.w};
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (c + i < 0 || c + i >= RD_WX) continue;
                sm.win.r[j][rd_col(c + i)] = rv[i];
                sm.win.g[j][rd_col(c + i)] = gv[i];
            }
        });
    __syncthreads();
    float3 o[RD_QPT][2][2];
    int has_nan = 0;
#pragma unroll
    for (int n = 0; n < RD_QPT; n++) {
        const int q = tid + 256 * n;
        if (q >= RD_QX * RD_QY) break;
        const int qx = q % RD_QX, qy = q / RD_QX;
        float R[4][4], G[4][4];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                R[j][i] = sm.win.r[2 * qy + j][rd_col(2 * qx + i)];
                G[j][i] = sm.win.g[2 * qy + j][rd_col(2 * qx + i)];
            }
        red_blue_from<true>(R, G, b, red_is_c1_site, cvar, clip, m, o[n]);
#pragma unroll
        for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++)
                has_nan |= (int)(isnan(o[n][dy][dx].x) || isnan(o[n][dy][dx].y) || isnan(o[n][dy][dx].z));
    }
    // the windows are read (the YCbCr region takes their place); a region
    // without NaN takes the despeckle's NaN-free form
    const bool nonan = !__syncthreads_or(has_nan);
#pragma unroll
    for (int n = 0; n < RD_QPT; n++) {
        const int q = tid + 256 * n;
        if (q >= RD_QX * RD_QY) break;
        const int qx = q % RD_QX, qy = q / RD_QX;
#pragma unroll
        for (int dy = 0; dy < 2; dy++)
#pragma unroll
            for (int dx = 0; dx < 2; dx++)
                sm.set(2 * qy + dy, 2 * qx + dx, o[n][dy][dx]);
    }
    __syncthreads();
    // every 3x3 neighbourhood of an interior tile lies inside the image: no clamps
    const bool interior = x0 >= 1 && y0 >= 1 && x0 + RD_X + 1 <= out.w && y0 + RD_Y + 1 <= out.h;
    if (interior && nonan) {
        for (int i = tid; i < RD_X * RD_Y; i += 256) {
            const int lx = i % RD_X, ly = i / RD_X;
            st4_<H>(out, x0 + lx, y0 + ly,
                    despeckle_px<true>([&](int dx, int dy) { return sm.at(ly + 2 + dy, lx + 2 + dx); }, var_a, var_b));
        }
        return;
    }
    for (int i = tid; i < RD_X * RD_Y; i += 256) {
        const int x = x0 + i % RD_X, y = y0 + i / RD_X;
        if (x >= out.w || y >= out.h) continue;
        auto fetch = [&](int dx, int dy) {
            const int gx = clampi(x + dx, 0, out.w - 1), gy = clampi(y + dy, 0, out.h - 1);
            return sm.at(gy - ry0, gx - rx0);
        };
        st4_<H>(out, x, y, nonan ? despeckle_px<true>(fetch, var_a, var_b) : despeckle_px<false>(fetch, var_a, var_b));
    }
}

template <bool U16, bool H>
__global__ void __launch_bounds__(256) k_rb_despeckle(MosaicSrc<U16> raw, Img green, Img out, Bayer b,
                                                      int red_is_c1_site, ParamRef cv, float clip, Mat3 m,
                                                      float var_a, float var_b) {
    __shared__ RbDespeckleLds sm;
    const TileId tl = xcd_tile();
    rb_despeckle_tile<U16, H>(raw, green, out, b, red_is_c1_site, cv, clip, m, var_a, var_b, tl.x * RD_X, tl.y * RD_Y,
                           threadIdx.x, sm);
}

// ----------------------------------------------------- burst (cross-frame)
// The green and red/blue + despeckle stages of every frame of a burst in one
// launch each.  The frames' tiles form one list (frame-major, row-major
// within a frame) cut into 8 contiguous segments, one per XCD; a persistent
// block on XCD x (launch index mod 8) takes `grab` consecutive tiles of
// segment x at a time from that segment's counter until it runs dry, so
// neighbouring tiles (whose halos overlap) share an L2 while the blocks of
// an XCD balance among themselves.  No block waits for another: a block that
// starts after its segment is done leaves at once.  The last block out
// clears the counters, so the next launch (or graph replay) starts from zero.
__device__ __forceinline__ int burst_frame_of(const BurstFrame* frames, int kind, int t) {
    int f = 0;
    while (t >= frames[f].tile_end[kind]) f++;
    return f;
}

__device__ __forceinline__ void burst_leave(BurstQueue* q) {
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&q->done, 1u) == gridDim.x - 1) {
            __threadfence();
            for (int i = 0; i < 8; i++) atomicExch(&q->next[i], 0u);
            atomicExch(&q->done, 0u);
        }
    }
}

template <int KIND, int WAVES, bool H>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES)))
k_demosaic_burst(const BurstFrame* __restrict__ frames, int total, BurstQueue* q, int grab) {
    __shared__ typename std::conditional<KIND == 0, GreenLds, RbDespeckleLds>::type sm;
    __shared__ int s_t;
    const int xcd = blockIdx.x & 7;
    const int lo = (int)((long long)total * xcd / 8), hi = (int)((long long)total * (xcd + 1) / 8);
    for (;;) {
        if (threadIdx.x == 0) s_t = lo + (int)atomicAdd(&q->next[xcd], (unsigned int)grab);
        __syncthreads();
        const int t0 = __builtin_amdgcn_readfirstlane(s_t);  // uniform: the frame table is read by scalar loads
        if (t0 >= hi) break;
        const int t1 = min(t0 + grab, hi);
        for (int t = t0; t < t1; t++) {
            const int f = burst_frame_of(frames, KIND, t);
            const BurstFrame& fr = frames[f];
            const int lt = t - (f ? frames[f - 1].tile_end[KIND] : 0);
            const int ty = lt / fr.tiles_x[KIND], tx = lt - ty * fr.tiles_x[KIND];
            if (KIND == 0) {
                green_tile<true>(fr.raw, fr.green, fr.gxo, fr.gyo, fr.lv, tx * GT_X, ty * GT_Y, threadIdx.x,
                                 reinterpret_cast<GreenLds&>(sm));
            } else {
                Mat3 m;
#pragma unroll
                for (int i = 0; i < 9; i++) m.m[i] = fr.c2y[i];
                rb_despeckle_tile<true, H>(fr.raw, fr.green, fr.ycc, fr.b, fr.red_is_c1, fr.cv, fr.clip, m, fr.var_a,
                                        fr.var_b, tx * RD_X, ty * RD_Y, threadIdx.x,
                                        reinterpret_cast<RbDespeckleLds&>(sm));
            }
            __syncthreads();  // the tile's LDS is free
        }
    }
    burst_leave(q);
}

void burst_tiles(int w, int h, int kind, int* tiles_x, int* tiles) {
    const int tx = kind == 0 ? GT_X : RD_X, ty = kind == 0 ? GT_Y : RD_Y;
    *tiles_x = (w + tx - 1) / tx;
    *tiles = *tiles_x * ((h + ty - 1) / ty);
}

void launch_demosaic_burst(int kind, int waves, bool half, const BurstFrame* frames, int total_tiles, BurstQueue* q,
                           int blocks, int grab, hipStream_t st) {
    // the green stage writes FP32 green planes whatever the level storage
    if (kind == 0 && waves >= 8)
        hipLaunchKernelGGL((k_demosaic_burst<0, 8, false>), dim3(blocks), dim3(256), 0, st, frames, total_tiles, q,
                           grab);
    else if (kind == 0 && waves >= 6)
        hipLaunchKernelGGL((k_demosaic_burst<0, 6, false>), dim3(blocks), dim3(256), 0, st, frames, total_tiles, q,
                           grab);
    else if (kind == 0)
        hipLaunchKernelGGL((k_demosaic_burst<0, 4, false>), dim3(blocks), dim3(256), 0, st, frames, total_tiles, q,
                           grab);
    else if (half)
        hipLaunchKernelGGL((k_demosaic_burst<1, 4, true>), dim3(blocks), dim3(256), 0, st, frames, total_tiles, q,
                           grab);
    else
        hipLaunchKernelGGL((k_demosaic_burst<1, 4, false>), dim3(blocks), dim3(256), 0, st, frames, total_tiles, q,
                           grab);
}

void launch_rb_despeckle(const MosaicArg& raw, const Img& green, const Img& out, int bayer, const Bayer& b,
                         const ParamRef& cv, float clip, const float c2y[9], float var_a, float var_b,
                         hipStream_t st) {
    const int red_is_c1 = (bayer == 3 || bayer == 0) ? 1 : 3;  // as in launch_red_blue
    Mat3 m;
    for (int i = 0; i < 9; i++) m.m[i] = c2y[i];
    dim3 blk(256), grd((out.w + RD_X - 1) / RD_X, (out.h + RD_Y - 1) / RD_Y);
    with_storage(out, [&](auto H) {
        if (raw.u16)
            hipLaunchKernelGGL((k_rb_despeckle<true, H()>), grd, blk, 0, st, raw.src<true>(), green, out, b,
                               red_is_c1, cv, clip, m, var_a, var_b);
        else
            hipLaunchKernelGGL((k_rb_despeckle<false, H()>), grd, blk, 0, st, raw.src<false>(), green, out, b,
                               red_is_c1, cv, clip, m, var_a, var_b);
    });
}

}  // namespace clim
