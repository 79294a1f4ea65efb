// k_pyramid.hip - gfx950 kernels for the pyramidal denoise and the tone map:
// downsampleImage, noiseStatistics, denoiseImage (hard sigma filter),
// denoiseImageGuided, reassembleImage and convertTosRGB (optionally preceded
// by the YCbCr->camera transformImage on the same pixel).
//
// Reference: cl/demosaic.cl (cl/ = CLImage/app/src/main/cpp/OpenCL/).
#include "fit_acc.h"
#include "kernels.h"
#include "tiles.h"
#include "tone.h"

namespace clim {

// CL_FILTER_LINEAR | CLK_ADDRESS_CLAMP_TO_EDGE with normalised coordinates:
// u = x*W - 0.5, i0 = floor(u), weight frac(u) rounded to 8 fractional bits.
// The sampling half, given the texel (i0, j0) and the rounded weights (a, b):
template <typename Fetch>
__device__ __forceinline__ float4 bilinear4_at(Fetch ld, int i0, int j0, float a, float b) {
    const float4 t00 = ld(i0, j0), t10 = ld(i0 + 1, j0);
    const float4 t01 = ld(i0, j0 + 1), t11 = ld(i0 + 1, j0 + 1);
    const float w00 = (1.0f - a) * (1.0f - b), w10 = a * (1.0f - b);
    const float w01 = (1.0f - a) * b, w11 = a * b;
    float4 r;
    r.x = w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x;
    r.y = w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y;
    r.z = w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z;
    r.w = w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w;
    return r;
}
template <bool H = false>
__device__ __forceinline__ float4 bilinear4_at(const Img& im, int i0, int j0, float a, float b) {
    return bilinear4_at([&](int i, int j) { return ld4_<H>(im, i, j); }, i0, j0, a, b);
}

template <bool H = false>
__device__ __forceinline__ float4 bilinear4(const Img& im, float un, float vn) {
    const float u = un * (float)im.w - 0.5f;
    const float v = vn * (float)im.h - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    float a = u - fu, b = v - fv;
    a = floorf(a * 256.0f + 0.5f) * 0.00390625f;
    b = floorf(b * 256.0f + 0.5f) * 0.00390625f;
    return bilinear4_at<H>(im, (int)fu, (int)fv, a, b);
}

// Exact 2x ratios.  For an output of at most kDown2xMax texels per axis
// downsampled from exactly twice that, the float coordinate arithmetic of
// downsampleImage puts every tap at texel 2x-1 or 2x+1 with weight exactly
// 1/2; for an output of at most kUp2xMax texels per axis sampling an image of
// exactly half that (reassembleImage), even x lands on texel x/2-1 with
// weight 3/4, odd x on (x-1)/2 with 1/4.  (Checked exhaustively for every
// size up to these limits; the general path serves larger images.)  The
// kernels then skip the coordinate arithmetic and call the same sampling code.
constexpr int kDown2xMax = 6144, kUp2xMax = 16384;

// ------------------------------------------------------------ downsampleImage
// cl/demosaic.cl:853-865: four linear taps at +-1/2 output texel, averaged.
template <bool EXACT2X, bool H>
__global__ void __launch_bounds__(256) k_downsample(Img in, Img out) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x;
    const int y = tl.y * 4 + threadIdx.y;
    if (x >= out.w || y >= out.h) return;
    float4 t0, t1, t2, t3;
    if (EXACT2X) {
        t0 = bilinear4_at<H>(in, 2 * x - 1, 2 * y - 1, 0.5f, 0.5f);
        t1 = bilinear4_at<H>(in, 2 * x + 1, 2 * y - 1, 0.5f, 0.5f);
        t2 = bilinear4_at<H>(in, 2 * x - 1, 2 * y + 1, 0.5f, 0.5f);
        t3 = bilinear4_at<H>(in, 2 * x + 1, 2 * y + 1, 0.5f, 0.5f);
    } else {
        const float nx = 1.0f / (float)out.w, ny = 1.0f / (float)out.h;
        const float px = ((float)x + 0.5f) * nx, py = ((float)y + 0.5f) * ny;
        const float sx = 0.5f * nx, sy = 0.5f * ny;
        t0 = bilinear4<H>(in, px - sx, py - sy);
        t1 = bilinear4<H>(in, px + sx, py - sy);
        t2 = bilinear4<H>(in, px - sx, py + sy);
        t3 = bilinear4<H>(in, px + sx, py + sy);
    }
    float4 acc = t0;
    acc.x += t1.x; acc.y += t1.y; acc.z += t1.z; acc.w += t1.w;
    acc.x += t2.x; acc.y += t2.y; acc.z += t2.z; acc.w += t2.w;
    acc.x += t3.x; acc.y += t3.y; acc.z += t3.z; acc.w += t3.w;
    st4_<H>(out, x, y, make_float4(0.25f * acc.x, 0.25f * acc.y, 0.25f * acc.z, 0.25f * acc.w));
}

bool downsample_exact2x(const Img& in, const Img& out) {
    return in.w == 2 * out.w && in.h == 2 * out.h && out.w <= kDown2xMax && out.h <= kDown2xMax;
}

void launch_downsample(const Img& in, const Img& out, hipStream_t st) {
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    // a level and the next share their storage
    with_storage(in, [&](auto H) {
        if (downsample_exact2x(in, out))
            hipLaunchKernelGGL((k_downsample<true, H()>), grd, blk, 0, st, in, out);
        else
            hipLaunchKernelGGL((k_downsample<false, H()>), grd, blk, 0, st, in, out);
    });
}

// 5x5 stencils over an RGBA level: a 32x16 output tile plus a 2-texel halo in
// LDS (xyz only; alpha never enters these filters).
namespace {
constexpr int TX = 32, TY = 16, HR = 2;
}

struct StatsTile {
    LevelTile<TX, TY> t;
    StatsRows<TX, TY> r;
};

// ------------------------------------------------------------ noiseStatistics
// cl/demosaic.cl:1267-1286: (mean Y, var Y, var Cb, var Cr) over 5x5.
__global__ void __launch_bounds__(256) k_noise_stats(Img in, Img out) {
    __shared__ StatsTile sm;
    const TileId tl = xcd_tile();
    const int x0 = tl.x * TX, y0 = tl.y * TY;
    const int tid = threadIdx.y * 32 + threadIdx.x;
    sm.t.load(in, x0, y0, tid);
    __syncthreads();
    sm.r.compute(sm.t, tid);
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < 2; rr++) {
        const int tx = threadIdx.x, ty = threadIdx.y + 8 * rr;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= out.w || y >= out.h) continue;
        wrow_f4(out, y)[x] = sm.r.stats(tx, ty);
    }
}

void launch_noise_stats(const Img& in, const Img& out, hipStream_t st) {
    dim3 blk(32, 8), grd((out.w + TX - 1) / TX, (out.h + TY - 1) / TY);
    hipLaunchKernelGGL(k_noise_stats, grd, blk, 0, st, in, out);
}

// downsampleImage (exact 2x, k_downsample<true>) of the 16x8 half-size
// texels a 32x16 level tile covers, in two steps.  Half-size texel (X, Y)
// sums the bilinear taps (weights 1/4) at texels (2X -/+ 1, 2Y -/+ 1); tap
// (2X - 1, 2Y - 1) sits at tile column 2 lx + 1, row 2 ly + 1 (lx, ly: the
// texel's place in the tile's 16x8 block) and neighbouring texels share
// taps, so the tile's 17x9 distinct taps are computed once (down_taps: all
// four channels, the arithmetic of bilinear4_at) and each texel adds its
// four in k_downsample's order (down_write).
constexpr int kTapW = TX / 2 + 1, kTapH = TY / 2 + 1;

// the statistics kernel's LDS: the interleaved level tile, its horizontal
// sums and (DOWN) the next level's taps, channels (x, y) and (z, w)
template <bool DOWN>
struct StatsFitTile {
    LevelTile4<TX, TY> t;
    StatsRows<TX, TY, true> r;
    float2 taps[DOWN ? 2 : 1][DOWN ? kTapH : 1][DOWN ? kTapW : 1];
};

__device__ __forceinline__ void down_taps(const LevelTile4<TX, TY>& t, float2 (*tp)[kTapH][kTapW], int tid) {
    // taps go to the threads with two rows of horizontal sums (tid >= 128) first
    const int k = 255 - tid;
    if (k >= kTapW * kTapH) return;
    const int i = k % kTapW, j = k / kTapW, c = 2 * i + 1, r = 2 * j + 1;
    const float4 a = t.p[r][c], b = t.p[r][c + 1], d = t.p[r + 1][c], e = t.p[r + 1][c + 1];
    tp[0][j][i] = make_float2(0.25f * a.x + 0.25f * b.x + 0.25f * d.x + 0.25f * e.x,
                              0.25f * a.y + 0.25f * b.y + 0.25f * d.y + 0.25f * e.y);
    tp[1][j][i] = make_float2(0.25f * a.z + 0.25f * b.z + 0.25f * d.z + 0.25f * e.z,
                              0.25f * a.w + 0.25f * b.w + 0.25f * d.w + 0.25f * e.w);
}

// thread tid writes channels 2p, 2p + 1 (p = tid / 128) of texel tid % 128
// (16x8, row-major)
template <bool H>
__device__ __forceinline__ void down_write(const float2 (*tp)[kTapH][kTapW], const Img& out, int x0, int y0,
                                           int tid) {
    const int p = tid >> 7, o = tid & 127, lx = o & 15, ly = o >> 4;
    const int X = (x0 >> 1) + lx, Y = (y0 >> 1) + ly;
    if (X >= out.w || Y >= out.h) return;
    const float2 t0 = tp[p][ly][lx], t1 = tp[p][ly][lx + 1], t2 = tp[p][ly + 1][lx], t3 = tp[p][ly + 1][lx + 1];
    float ax = t0.x, ay = t0.y;
    ax += t1.x; ay += t1.y;
    ax += t2.x; ay += t2.y;
    ax += t3.x; ay += t3.y;
    if (H) {
        half2_t hv;
        hv.x = (_Float16)(0.25f * ax);
        hv.y = (_Float16)(0.25f * ay);
        reinterpret_cast<half2_t*>(out.p + (size_t)Y * out.pitch)[2 * X + p] = hv;
    } else {
        reinterpret_cast<float2*>(wrow_f4(out, Y) + X)[p] = make_float2(0.25f * ax, 0.25f * ay);
    }
}

// noiseStatistics of pyramid levels (one job per level) with the first
// regression pass of computeNoiseStatistics accumulated on the way: each of
// the grid's blocks takes a contiguous run of one level's 32x16 tiles and
// keeps its regression sums across them (one partial per block); the
// statistics in range are listed for the refit.
// DOWN (one job): each tile also writes the next level's texels it covers
// (downsampleImage), so that level is made without reading this one again.
template <bool DOWN, bool H>
__global__ void __launch_bounds__(256) k_noise_stats_fit(FitJobs jobs, Img down) {
    using L = Layout<3, true, false>;
    using Tile = LevelTile4<TX, TY>;
    __shared__ StatsFitTile<DOWN> sm;
    __shared__ clim_fx128 tot[L::K];
    __shared__ int nlisted;
    static_assert(sizeof(StatsFitTile<DOWN>) >= 1024 * sizeof(clim_fx128),
                  "the tile storage serves as the reduction buffer");
    const int lb = xcd_block();
    const FitJob& jb = jobs.j[job_of_block(jobs, lb)];
    const int bi = lb - jb.block_base;
    const int tid = threadIdx.x, tx = tid & 31, ty0 = tid >> 5;
    const Img& in = jb.src;
    const int tiles_x = (in.w + TX - 1) / TX;
    const long long ntiles = jb.ntiles;
    const int t0 = (int)(ntiles * bi / jb.nblocks), t1 = (int)(ntiles * (bi + 1) / jb.nblocks);
    // this block's run of the sample list (the first barrier of the loop
    // orders the counter's reset before its first use)
    float4* list = reinterpret_cast<float4*>(jb.m.p) + (size_t)t0 * (TX * TY);
    if (tid == 0) nlisted = 0;
    Acc<3, true, false> acc;
    acc.zero();
    // the next tile's loads are issued before the current tile's compute
    typename Tile::Regs pf;
    if (t0 < t1) Tile::template fetch<H>(pf, in, (t0 % tiles_x) * TX, (t0 / tiles_x) * TY, tid);
    for (int t = t0; t < t1; t++) {
        const int x0 = (t % tiles_x) * TX, y0 = (t / tiles_x) * TY;
        sm.t.put(pf, tid);
        __syncthreads();
        if (t + 1 < t1) Tile::template fetch<H>(pf, in, ((t + 1) % tiles_x) * TX, ((t + 1) / tiles_x) * TY, tid);
        sm.r.compute(sm.t, tid);
        if constexpr (DOWN) down_taps(sm.t, sm.taps, tid);
        __syncthreads();
        if constexpr (DOWN) down_write<H>(sm.taps, down, x0, y0, tid);
        // this thread's outputs: rows 2 (tid / 32) and 2 (tid / 32) + 1 of column tid % 32
        float4 s4[2];
        bool ok[2];
        int slot[2];
        sm.r.stats2(tx, 2 * ty0, s4[0], s4[1]);
#pragma unroll
        for (int rr = 0; rr < 2; rr++) {
            const int x = x0 + tx, y = y0 + 2 * ty0 + rr;
            const float m[1] = {s4[rr].x};
            const float v[3] = {s4[rr].y, s4[rr].z, s4[rr].w};
            ok[rr] = x < in.w && y < in.h && Acc<3, true, false>::in_range(m, v);
        }
        list_slots2(ok[0], ok[1], &nlisted, slot[0], slot[1]);
#pragma unroll
        for (int rr = 0; rr < 2; rr++) {
            if (ok[rr]) list[slot[rr]] = s4[rr];
            const float m[1] = {s4[rr].x};
            const float v[3] = {s4[rr].y, s4[rr].z, s4[rr].w};
            acc.add_tested(m, v, ok[rr]);
        }
        __syncthreads();  // the next tile reuses the LDS
    }
    // the last tile is done (the loop ends with a barrier): its LDS is free
    if (tid == 0) jb.counts[bi] = (unsigned int)nlisted;
    store_block_partials(acc, reinterpret_cast<clim_fx128*>(&sm), tot, jb.partials, bi, kFitBlocksPerJob);
}

void fit_list_dims(bool raw, int w, int h, int* aw, int* ah) {
    // raw: 16x16-quad tiles of the (w/2) x (h/2) quads; levels: 32x16 tiles
    if (raw) {
        *aw = (w / 2 + 15) / 16 * 16;
        *ah = (h / 2 + 15) / 16 * 16;
    } else {
        *aw = (w + TX - 1) / TX * TX;
        *ah = (h + TY - 1) / TY * TY;
    }
}

bool launch_noise_stats_fit(FitJobs& jobs, int j0, int j1, const Img* down, hipStream_t st) {
    FitJobs sub{};
    sub.njobs = j1 - j0;
    long long tiles[5];
    for (int j = j0; j < j1; j++) {
        FitJob& jb = jobs.j[j];
        jb.ntiles = ((jb.src.w + TX - 1) / TX) * ((jb.src.h + TY - 1) / TY);
        jb.seg = TX * TY;
        if ((long long)jb.ntiles * jb.seg > jb.cap) return false;
        sub.j[j - j0] = jb;
        tiles[j - j0] = jb.ntiles;
    }
    const int grid = fit_share_blocks(sub, tiles, kLevelStatsGrid);
    for (int j = j0; j < j1; j++) jobs.j[j] = sub.j[j - j0];
    // the jobs' levels share their storage (and the next level's)
    with_storage(sub.j[0].src, [&](auto H) {
        if (down)
            hipLaunchKernelGGL((k_noise_stats_fit<true, H()>), dim3(grid), dim3(256), 0, st, sub, *down);
        else
            hipLaunchKernelGGL((k_noise_stats_fit<false, H()>), dim3(grid), dim3(256), 0, st, sub, Img{});
    });
    return true;
}

// --------------------------------------------------------------- denoiseImage
// cl/demosaic.cl:735-765.  nlf: A_Y, A_Cb, A_Cr, B_Y, B_Cb, B_Cr.  The range
// test length((s - c) / sigma) < edge is sum u_c^2 < edge^2 with
// u_c = fma(s_c, r_c, -(c_c r_c)), r_c = 1 / sigma_c (clim_range_len2; the FP32
// oracle evaluates the same expressions).
// SAME_EDGE: gradientBoost <= 1 and the luma and chroma edges are equal, so
// every tap's luma and chroma weights are equal (one test, one weight sum).
//
// One output texel's state: 1 / sigma per channel, the scaled centre, the
// squared edges, the weighted sums and the weight counts (the weights are 0
// or 1).
struct BilCentre {
    float r0, r1, r2, cr0, cr1, cr2, e0, e1;
    float f0, f1, f2;
    int c0, c1;
    float nH, E, n;  // SAME_EDGE: the step of bilateral_tap, its weight count
};

// the centre at tile texel (X, Y) (halo included); MAP: the strength m may
// be non-zero (without it m is 0 and the edges are the same for every texel)
template <bool SAME_EDGE, bool MAP>
__device__ __forceinline__ BilCentre bilateral_centre(const LevelTile4<TX, TY>& t, int X, int Y, float va0, float va1,
                                                      float va2, float vb0, float vb1, float vb2, float chromaBoost,
                                                      float gradientBoost, float m) {
    const float4 cp = t.p[Y][X];
    const float cY = cp.x, cCb = cp.y, cCr = cp.z;
    float lumaBoost = 1.0f;
    if (!SAME_EDGE && gradientBoost > 1.0f) {
        const float sig0 = sqrtf(va0 + vb0 * cY);
        // sobel() of the luma channel, cl/demosaic.cl:96-112
        const float kx[3][3] = {{1, 0, -1}, {2, 0, -2}, {1, 0, -1}};
        const float ky[3][3] = {{1, 2, 1}, {0, 0, 0}, {-1, -2, -1}};
        float vx = 0.0f, vy = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const float s = t.p[Y - 1 + j][X - 1 + i].x;
                vx += kx[j][i] * s;
                vy += ky[j][i] * s;
            }
        vx = vx * 0.471404521f;
        vy = vy * 0.471404521f;
        const float is0 = 1.0f / sig0;
        lumaBoost = gradientBoost - (gradientBoost - 1.0f) * smoothstep(0.5f, 2.0f, len2(vx * is0, vy * is0));
    }
    BilCentre c;
    float e0 = lumaBoost * 1.41421356f, e1 = chromaBoost * 1.41421356f;
    e0 = e0 * e0;
    e1 = e1 * e1;
    if (MAP && m > 0.0f) {
        // learned strength: the range sigma times m, i.e. the squared edges times m^2
        const float m2 = m * m;
        e0 = e0 * m2;
        e1 = e1 * m2;
    }
    c.e0 = e0;
    c.e1 = e1;
    // SAME_EDGE's weight as one fused step (bilateral_tap): H = 2^(50 - e),
    // e the exponent of e0, so that E = e0 H is in [2^50, 2^51) exactly
    // (e0 is a positive normal float: 2 m^2 with m in [2^-1.5, 2^1.5] for the
    // learned map, 2 without it)
    c.nH = -clim_u2f((304u - (clim_f2u(e0) >> 23)) << 23);
    c.E = e0 * -c.nH;
    c.n = 0.0f;
    // 1 / sigma per channel (clim_inv_sigma); the squared distance by
    // clim_range_len2.  The Newton steps run on every lane and the special
    // cases (a variance that is not a positive finite float) once per wave,
    // only when some lane has one: the same values as clim_inv_sigma.
    const float s0 = va0 + vb0 * cY, s1 = va1 + vb1 * cY, s2 = va2 + vb2 * cY;
    c.r0 = clim_rsqrt_newton(s0);
    c.r1 = clim_rsqrt_newton(s1);
    c.r2 = clim_rsqrt_newton(s2);
    constexpr int kPosFinite = 128 | 256;  // +subnormal, +normal
    if (__builtin_expect(!(__builtin_amdgcn_classf(s0, kPosFinite) && __builtin_amdgcn_classf(s1, kPosFinite) &&
                           __builtin_amdgcn_classf(s2, kPosFinite)),
                         0)) {
        c.r0 = clim_inv_sigma(s0);
        c.r1 = clim_inv_sigma(s1);
        c.r2 = clim_inv_sigma(s2);
    }
    c.cr0 = cY * c.r0;
    c.cr1 = cCb * c.r1;
    c.cr2 = cCr * c.r2;
    c.f0 = c.f1 = c.f2 = 0.0f;
    c.c0 = c.c1 = 0;
    return c;
}

// one tap (a, b, cc) of the centre's 5x5 window, in window order
template <bool SAME_EDGE>
__device__ __forceinline__ void bilateral_tap(BilCentre& c, float a, float b, float cc) {
    const float l = clim_range_len2(a, b, cc, c.r0, c.r1, c.r2, c.cr0, c.cr1, c.cr2);
    // f += w * s with w in {0, 1}: the product is exact, so the fused form
    // rounds once on the same value (the same bits)
    if (SAME_EDGE) {
        // w = (l < e0) as clamp(fma(l, -H, e0 H), 0, 1), one instruction (the
        // clamp is the FMA's output modifier): for l < e0 the exact value
        // H (e0 - l) >= H ulp(e0) / 2 = 2^26 rounds to >= 1; for l >= e0 it is
        // <= 0; a NaN clamps to 0 (clim_step_below_scaled)
        const float w = __builtin_amdgcn_fmed3f(fmaf(l, c.nH, c.E), 0.0f, 1.0f);
        c.f0 = fmaf(w, a, c.f0);
        c.f1 = fmaf(w, b, c.f1);
        c.f2 = fmaf(w, cc, c.f2);
        c.n += w;
        return;
    }
    const bool in0 = l < c.e0, in1 = l < c.e1;
    const float w0 = in0 ? 1.0f : 0.0f, w1 = in1 ? 1.0f : 0.0f;
    c.f0 = fmaf(w0, a, c.f0);
    c.c0 += in0;
    c.f1 = fmaf(w1, b, c.f1);
    c.f2 = fmaf(w1, cc, c.f2);
    c.c1 += in1;
}

// the weighted means: sum / count by clim_div_count (the IEEE quotient's
// bits; rcp[n] = 1.0f / n); the Cr count equals the Cb one (same threshold,
// same weights)
template <bool SAME_EDGE>
__device__ __forceinline__ float4 bilateral_mean(const BilCentre& c, const float* rcp) {
    const int k0 = SAME_EDGE ? (int)c.n : c.c0, k1 = SAME_EDGE ? k0 : c.c1;
    const float n0 = SAME_EDGE ? c.n : (float)k0, n1 = SAME_EDGE ? n0 : (float)k1;
    const float q0 = rcp[k0], q1 = SAME_EDGE ? q0 : rcp[k1];
    // the fused form on every lane; clim_div_count's other cases once per
    // wave, only when some lane's product is not a normal float
    float4 r = make_float4(clim_div_count_normal(c.f0, n0, q0), clim_div_count_normal(c.f1, n1, q1),
                           clim_div_count_normal(c.f2, n1, q1), 0.0f);
    constexpr int kNormal = 8 | 256;  // -normal, +normal
    if (__builtin_expect(!(__builtin_amdgcn_classf(c.f0 * q0, kNormal) && __builtin_amdgcn_classf(c.f1 * q1, kNormal) &&
                           __builtin_amdgcn_classf(c.f2 * q1, kNormal)),
                         0)) {
        r.x = clim_div_count(c.f0, n0, q0);
        r.y = clim_div_count(c.f1, n1, q1);
        r.z = clim_div_count(c.f2, n1, q1);
    }
    return r;
}

// Two vertically adjacent outputs (tile rows ty, ty + 1) of one thread: their
// 5x5 windows share four rows, so the six rows of taps are read from LDS once
// and each tap goes to the one or two windows it belongs to, every window in
// its own row-major order (the sums are those of the one-texel form).
template <bool SAME_EDGE>
__device__ __forceinline__ void bilateral_pair(const LevelTile4<TX, TY>& t, int tx, int ty, BilCentre& A,
                                               BilCentre& B) {
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const float4 v = t.p[ty][tx + i];
        bilateral_tap<SAME_EDGE>(A, v.x, v.y, v.z);
    }
#pragma unroll 1
    for (int j = 1; j < 5; j++) {
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const float4 v = t.p[ty + j][tx + i];
            bilateral_tap<SAME_EDGE>(A, v.x, v.y, v.z);
            bilateral_tap<SAME_EDGE>(B, v.x, v.y, v.z);
        }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
        const float4 v = t.p[ty + 5][tx + i];
        bilateral_tap<SAME_EDGE>(B, v.x, v.y, v.z);
    }
}

// The outputs of one 32x16 tile (x0, y0) of denoiseImage (a w x h level)
// whose texels are in LDS (t, and the strength map texels in mt);
// rcp[n] = 1.0f / n.  Each result goes to emit(x, y, value).
template <bool SAME_EDGE, bool MAP, typename Emit>
__device__ __forceinline__ void bilateral_compute_(const LevelTile4<TX, TY>& t, const MapStage& ms, const float* mt,
                                                   const float* rcp, int w, int h, int x0, int y0,
                                                   const ParamRef& nlf, float chromaBoost, float gradientBoost,
                                                   const StrengthRef& sr, int level, Emit emit) {
    const float va0 = nlf.get(0), va1 = nlf.get(1), va2 = nlf.get(2);
    const float vb0 = nlf.get(3), vb1 = nlf.get(4), vb2 = nlf.get(5);
    // thread (tx, ty) finishes rows 2 ty and 2 ty + 1 of column tx
    const int tx = threadIdx.x, ty = 2 * threadIdx.y;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= w || y >= h) return;
    const float ma = MAP ? strength_at_staged(sr, ms, mt, level, x, y) : 0.0f;
    const float mb = MAP ? strength_at_staged(sr, ms, mt, level, x, y + 1) : 0.0f;
    BilCentre A = bilateral_centre<SAME_EDGE, MAP>(t, tx + HR, ty + HR, va0, va1, va2, vb0, vb1, vb2, chromaBoost,
                                                   gradientBoost, ma);
    BilCentre B = bilateral_centre<SAME_EDGE, MAP>(t, tx + HR, ty + 1 + HR, va0, va1, va2, vb0, vb1, vb2,
                                                   chromaBoost, gradientBoost, mb);
    bilateral_pair<SAME_EDGE>(t, tx, ty, A, B);
    const float4 ra = bilateral_mean<SAME_EDGE>(A, rcp), rb = bilateral_mean<SAME_EDGE>(B, rcp);
    emit(x, y, ra);
    if (y + 1 < h) emit(x, y + 1, rb);
}
template <bool SAME_EDGE, typename Emit>
__device__ __forceinline__ void bilateral_compute(const LevelTile4<TX, TY>& t, const MapStage& ms, const float* mt,
                                                  const float* rcp, int w, int h, int x0, int y0,
                                                  const ParamRef& nlf, float chromaBoost, float gradientBoost,
                                                  const StrengthRef& sr, int level, Emit emit) {
    // without a strength map for the level (or a constant strength) every
    // texel's m is 0 and the edge setup folds to constants
    if (!(sr.constant > 0.0f) && (!sr.map || !sr.open[level]))
        bilateral_compute_<SAME_EDGE, false>(t, ms, mt, rcp, w, h, x0, y0, nlf, chromaBoost, gradientBoost, sr, level,
                                             emit);
    else
        bilateral_compute_<SAME_EDGE, true>(t, ms, mt, rcp, w, h, x0, y0, nlf, chromaBoost, gradientBoost, sr, level,
                                            emit);
}

// stages one 32x16 tile (x0, y0) of denoiseImage: the level's texels with
// their halo, the strength map texels, the count reciprocals (the caller's
// barrier follows)
template <bool H>
__device__ __forceinline__ MapStage bilateral_stage(LevelTile4<TX, TY>& t, float* mt, float* rcp, const Img& in,
                                                    int x0, int y0, const StrengthRef& sr, int level, int tid) {
    t.template load<H>(in, x0, y0, tid);
    const MapStage ms = stage_strength<TX, TY>(sr, level, x0, y0, mt, tid);
    if (tid < 26) rcp[tid] = 1.0f / (float)tid;
    return ms;
}

// one 32x16 tile (x0, y0) of denoiseImage
template <bool SAME_EDGE, bool H>
__device__ __forceinline__ void bilateral_tile(LevelTile4<TX, TY>& t, float* mt, float* rcp, const Img& in,
                                               const Img& out, int x0, int y0, const ParamRef& nlf, float chromaBoost,
                                               float gradientBoost, const StrengthRef& sr, int level) {
    const int tid = threadIdx.y * 32 + threadIdx.x;
    const MapStage ms = bilateral_stage<H>(t, mt, rcp, in, x0, y0, sr, level, tid);
    __syncthreads();
    bilateral_compute<SAME_EDGE>(t, ms, mt, rcp, out.w, out.h, x0, y0, nlf, chromaBoost, gradientBoost, sr, level,
                                 [&](int x, int y, float4 v) { st4_<H>(out, x, y, v); });
}

template <bool SAME_EDGE, bool H>
__global__ void __launch_bounds__(256) k_bilateral(Img in, Img out, ParamRef nlf, float chromaBoost,
                                                   float gradientBoost, StrengthRef sr, int level) {
    __shared__ LevelTile4<TX, TY> t;
    __shared__ float mt[kMapStageMax];
    __shared__ float rcp[26];
    const TileId tl = xcd_tile();
    bilateral_tile<SAME_EDGE, H>(t, mt, rcp, in, out, tl.x * TX, tl.y * TY, nlf, chromaBoost, gradientBoost, sr,
                                 level);
}

// denoiseImage of several pyramid levels in one launch (the levels' denoising
// reads only the noisy levels): a 1-D grid of the levels' tiles, the finest
// first, in XCD-contiguous runs.
template <bool SAME_EDGE>
__global__ void __launch_bounds__(256) k_bilateral_levels(BilateralLevels lv) {
    __shared__ LevelTile4<TX, TY> t;
    __shared__ float mt[kMapStageMax];
    __shared__ float rcp[26];
    const int b = xcd_block();
    int l = 0;
    while (l + 1 < lv.n && b >= lv.base[l + 1]) l++;
    const int i = b - lv.base[l], tiles_x = (lv.out[l].w + TX - 1) / TX;
    bilateral_tile<SAME_EDGE, false>(t, mt, rcp, lv.in[l], lv.out[l], (i % tiles_x) * TX, (i / tiles_x) * TY,
                                     lv.nlf[l], lv.chromaBoost[l], lv.gradientBoost[l], lv.sr, lv.level[l]);
}

void launch_bilateral(const Img& in, const Img& out, const ParamRef& nlf, float chromaBoost, float gradientBoost,
                      hipStream_t st, const StrengthRef* srp, int level) {
    dim3 blk(32, 8), grd((out.w + TX - 1) / TX, (out.h + TY - 1) / TY);
    const StrengthRef sr = srp ? *srp : StrengthRef{nullptr, 0, 0, nullptr, 0.0f};
    // the edges as the kernel computes them (lumaBoost = 1 without gradient boost)
    float e0 = 1.0f * 1.41421356f, e1 = chromaBoost * 1.41421356f;
    e0 = e0 * e0;
    e1 = e1 * e1;
    with_storage(in, [&](auto H) {
        if (!(gradientBoost > 1.0f) && e0 == e1)
            hipLaunchKernelGGL((k_bilateral<true, H()>), grd, blk, 0, st, in, out, nlf, chromaBoost, gradientBoost, sr,
                               level);
        else
            hipLaunchKernelGGL((k_bilateral<false, H()>), grd, blk, 0, st, in, out, nlf, chromaBoost, gradientBoost,
                               sr, level);
    });
}

template <bool SAME_EDGE>
__global__ void __launch_bounds__(256) k_bilateral_levels_mixed(BilateralLevels lv);

// the one-test form when launch_bilateral would take it for every level
void launch_bilateral_levels(BilateralLevels& lv, hipStream_t st) {
    bool same = true;
    int base = 0;
    for (int l = 0; l < lv.n; l++) {
        float e0 = 1.0f * 1.41421356f, e1 = lv.chromaBoost[l] * 1.41421356f;
        e0 = e0 * e0;
        e1 = e1 * e1;
        same = same && !(lv.gradientBoost[l] > 1.0f) && e0 == e1;
        lv.base[l] = base;
        base += ((lv.out[l].w + TX - 1) / TX) * ((lv.out[l].h + TY - 1) / TY);
    }
    // FP16 level images: the mixed-precision form (a pyramid's levels share
    // their storage; the FP32 form reads FP32 levels only)
    const bool mixed = lv.in[0].half != 0;
    if (same && mixed)
        hipLaunchKernelGGL(k_bilateral_levels_mixed<true>, dim3(base), dim3(32, 8), 0, st, lv);
    else if (mixed)
        hipLaunchKernelGGL(k_bilateral_levels_mixed<false>, dim3(base), dim3(32, 8), 0, st, lv);
    else if (same)
        hipLaunchKernelGGL(k_bilateral_levels<true>, dim3(base), dim3(32, 8), 0, st, lv);
    else
        hipLaunchKernelGGL(k_bilateral_levels<false>, dim3(base), dim3(32, 8), 0, st, lv);
}

// --------------------------------------- denoiseImage, mixed precision form
// The same filter on FP16 level images with the range test and the weights in
// packed FP16 and the weighted sums in FP32.  Two vertically adjacent taps go
// through each step together: q[c][r][x] holds channel c of texels (x, r) and
// (x, r + 1) as one FP16 pair, so a 5x5 window is two pair rows and a row of
// single taps (second half zero) of five taps each.  Per tap pair:
//   u_c = (s_c - c_c) r_c            v_pk_add_f16, v_pk_mul_f16 per channel
//                                    (the difference first: FP16 s_c r_c -
//                                    c_c r_c would carry the rounding of
//                                    c_c r_c, up to 2^-11 c_c / sigma_c)
//   l - e = u_Y^2 + u_Cb^2 + u_Cr^2 - e   three v_pk_fma_f16 from -e
//   w = clamp((e - l) 2^12 2^12, 0, 1)    two v_pk_mul_f16, the second clamping
//   f_c += s_c.lo w.lo + s_c.hi w.hi      v_dot2_f32_f16: FP32 sums (and count)
// w is exactly 1 when l - e < 0 and 0 otherwise (and for NaN): a negative
// FP16 value is at most -2^-24, and 2^24 times its magnitude is at least 1
// or overflows to +inf, both of which clamp to 1.  (With separate luma and
// chroma edges, l is formed first and each edge subtracted.)  r_c = 1 /
// sigma_c is computed in FP32 per centre pixel and rounded to FP16 (finite
// values above the FP16 range saturate; inf and NaN pass, giving no weights
// as the FP32 form does); the edges e = (boost sqrt 2)^2 (m^2 with the
// learned strength) are rounded to FP16; the sums are divided through the
// hardware reciprocal, the result being stored as FP16.
struct PairTile {
    static constexpr int SX = TX + 4, SY = TY + 4;
    // qv[r][x] = the (Y, Cb, Cr) pairs (texel (x, r), texel (x, r + 1)), one
    // 16-byte LDS read per tap pair; the last row's second halves are zero
    uint4 qv[SY][SX];
    __device__ __forceinline__ static uint32_t pack(float a, float b) {
        half2_t h;
        h.x = (_Float16)a;
        h.y = (_Float16)b;
        return __builtin_bit_cast(uint32_t, h);
    }
    // Thread t < 36 x 7 takes column t % SX and rows 3 (t / SX) .. + 2: it
    // loads the four texels of those rows and the next (zero past the tile)
    // and writes the three pairs, so each texel is loaded once or twice.
    static constexpr int kRun = 3, kRuns = (SY + kRun - 1) / kRun;
    static_assert(SX * kRuns <= 256, "one run per thread");
    __device__ __forceinline__ void load(const Img& in, int x0, int y0, int tid) {
        if (tid >= SX * kRuns) return;
        const int x = tid % SX, r0 = (tid / SX) * kRun;
        float4 v[kRun + 1];
#pragma unroll
        for (int k = 0; k <= kRun; k++)
            v[k] = r0 + k < SY ? ld4_<true>(in, x0 - 2 + x, y0 - 2 + r0 + k) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < kRun; k++)
            if (r0 + k < SY)
                qv[r0 + k][x] = make_uint4(pack(v[k].x, v[k + 1].x), pack(v[k].y, v[k + 1].y),
                                           pack(v[k].z, v[k + 1].z), 0u);
    }
    // texel (x, r) of channel c
    __device__ __forceinline__ _Float16 at(int c, int r, int x) const {
        const uint4 v = qv[r][x];
        const uint32_t w = c == 0 ? v.x : c == 1 ? v.y : v.z;
        return __builtin_bit_cast(half2_t, w).x;
    }
};

// (w.x, 0)
__device__ __forceinline__ half2_t only_first(half2_t w) {
    return __builtin_bit_cast(half2_t, __builtin_bit_cast(uint32_t, w) & 0xffffu);
}

__device__ __forceinline__ _Float16 to_half_sat(float r) {
    // finite values beyond the FP16 range saturate; inf and NaN pass
    return (_Float16)((r <= 65504.0f || !(r <= 3.0e38f)) ? r : 65504.0f);
}

template <bool SAME_EDGE>
__device__ __forceinline__ float4 bilateral_px_mixed(const PairTile& t, int tx, int ty, float va0, float va1,
                                                     float va2, float vb0, float vb1, float vb2, float chromaBoost,
                                                     float gradientBoost, float m) {
    const int X = tx + HR, Y = ty + HR;
    const float cY = (float)t.at(0, Y, X);
    float lumaBoost = 1.0f;
    if (!SAME_EDGE && gradientBoost > 1.0f) {
        const float sig0 = sqrtf(va0 + vb0 * cY);
        const float kx[3][3] = {{1, 0, -1}, {2, 0, -2}, {1, 0, -1}};
        const float ky[3][3] = {{1, 2, 1}, {0, 0, 0}, {-1, -2, -1}};
        float vx = 0.0f, vy = 0.0f;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int i = 0; i < 3; i++) {
                const float s = (float)t.at(0, Y - 1 + j, X - 1 + i);
                vx += kx[j][i] * s;
                vy += ky[j][i] * s;
            }
        vx = vx * 0.471404521f;
        vy = vy * 0.471404521f;
        const float is0 = 1.0f / sig0;
        lumaBoost = gradientBoost - (gradientBoost - 1.0f) * smoothstep(0.5f, 2.0f, len2(vx * is0, vy * is0));
    }
    float e0 = lumaBoost * 1.41421356f, e1 = chromaBoost * 1.41421356f;
    e0 = e0 * e0;
    e1 = e1 * e1;
    if (m > 0.0f) {
        const float m2 = m * m;
        e0 = e0 * m2;
        e1 = e1 * m2;
    }
    const float r0 = clim_inv_sigma(va0 + vb0 * cY), r1 = clim_inv_sigma(va1 + vb1 * cY),
                r2 = clim_inv_sigma(va2 + vb2 * cY);
    const half2_t rY = {to_half_sat(r0), to_half_sat(r0)}, rCb = {to_half_sat(r1), to_half_sat(r1)},
                  rCr = {to_half_sat(r2), to_half_sat(r2)};
    const half2_t cY2 = {t.at(0, Y, X), t.at(0, Y, X)}, cCb2 = {t.at(1, Y, X), t.at(1, Y, X)},
                  cCr2 = {t.at(2, Y, X), t.at(2, Y, X)};
    const _Float16 e0h = to_half_sat(e0), e1h = to_half_sat(e1);
    const half2_t nE0 = {-e0h, -e0h}, E0 = {e0h, e0h}, E1 = {e1h, e1h};
    const half2_t big = {(_Float16)4096.0f, (_Float16)4096.0f}, nbig = {(_Float16)-4096.0f, (_Float16)-4096.0f},
                  zero = {(_Float16)0.0f, (_Float16)0.0f}, one = {(_Float16)1.0f, (_Float16)1.0f};
    float f0 = 0.0f, f1 = 0.0f, f2 = 0.0f, c0 = 0.0f, c1 = 0.0f;
#pragma unroll
    for (int jp = 0; jp < 3; jp++) {
        // the fifth row has no partner: its pairs' second weights are cleared
#pragma unroll
        for (int i = 0; i < 5; i++) {
            const uint4 v = t.qv[ty + 2 * jp][tx + i];
            const half2_t sY = __builtin_bit_cast(half2_t, v.x), sCb = __builtin_bit_cast(half2_t, v.y),
                          sCr = __builtin_bit_cast(half2_t, v.z);
            const half2_t uY = (sY - cY2) * rY;
            const half2_t uCb = (sCb - cCb2) * rCb;
            const half2_t uCr = (sCr - cCr2) * rCr;
            if (SAME_EDGE) {
                // l - e accumulated from -e; its sign gives the weight
                const half2_t le = __builtin_elementwise_fma(
                    uCr, uCr, __builtin_elementwise_fma(uCb, uCb, __builtin_elementwise_fma(uY, uY, nE0)));
                half2_t w0 = __builtin_elementwise_min(__builtin_elementwise_max(le * nbig * big, zero), one);
                if (jp == 2) w0 = only_first(w0);
                f0 = __builtin_amdgcn_fdot2(sY, w0, f0, false);
                c0 = __builtin_amdgcn_fdot2(w0, one, c0, false);
                f1 = __builtin_amdgcn_fdot2(sCb, w0, f1, false);
                f2 = __builtin_amdgcn_fdot2(sCr, w0, f2, false);
            } else {
                const half2_t l =
                    __builtin_elementwise_fma(uCr, uCr, __builtin_elementwise_fma(uCb, uCb, uY * uY));
                half2_t w0 = __builtin_elementwise_min(__builtin_elementwise_max((E0 - l) * big * big, zero), one);
                half2_t w1 = __builtin_elementwise_min(__builtin_elementwise_max((E1 - l) * big * big, zero), one);
                if (jp == 2) {
                    w0 = only_first(w0);
                    w1 = only_first(w1);
                }
                f0 = __builtin_amdgcn_fdot2(sY, w0, f0, false);
                c0 = __builtin_amdgcn_fdot2(w0, one, c0, false);
                f1 = __builtin_amdgcn_fdot2(sCb, w1, f1, false);
                f2 = __builtin_amdgcn_fdot2(sCr, w1, f2, false);
                c1 = __builtin_amdgcn_fdot2(w1, one, c1, false);
            }
        }
    }
    // the quotients by the weight counts (1..25) through the reciprocal: the
    // result is stored as FP16
    const float k0 = __builtin_amdgcn_rcpf(c0), k1 = SAME_EDGE ? k0 : __builtin_amdgcn_rcpf(c1);
    return make_float4(f0 * k0, f1 * k1, f2 * k1, 0.0f);
}

template <bool SAME_EDGE>
__global__ void __launch_bounds__(256) k_bilateral_levels_mixed(BilateralLevels lv) {
    __shared__ PairTile t;
    __shared__ float mt[kMapStageMax];
    const int b = xcd_block();
    int l = 0;
    while (l + 1 < lv.n && b >= lv.base[l + 1]) l++;
    const int i = b - lv.base[l], tiles_x = (lv.out[l].w + TX - 1) / TX;
    const int x0 = (i % tiles_x) * TX, y0 = (i / tiles_x) * TY;
    const int tid = threadIdx.y * 32 + threadIdx.x;
    const Img& out = lv.out[l];
    const MapStage ms = stage_strength<TX, TY>(lv.sr, lv.level[l], x0, y0, mt, tid);
    t.load(lv.in[l], x0, y0, tid);
    __syncthreads();
    const ParamRef& nlf = lv.nlf[l];
    const float va0 = nlf.get(0), va1 = nlf.get(1), va2 = nlf.get(2);
    const float vb0 = nlf.get(3), vb1 = nlf.get(4), vb2 = nlf.get(5);
#pragma unroll 1
    for (int rr = 0; rr < 2; rr++) {
        const int tx = threadIdx.x, ty = threadIdx.y + 8 * rr;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= out.w || y >= out.h) continue;
        const float m = strength_at_staged(lv.sr, ms, mt, lv.level[l], x, y);
        st4_<true>(out, x, y,
                   bilateral_px_mixed<SAME_EDGE>(t, tx, ty, va0, va1, va2, vb0, vb1, vb2, lv.chromaBoost[l],
                                                 lv.gradientBoost[l], m));
    }
}

// --------------------------------------------------------- denoiseImageGuided
// cl/demosaic.cl:767-851: one-pass 5x5 self-guided colour guided filter.  The
// nine box sums (I and its six products) use the separable order.
struct GuidedTile {
    LevelTile<TX, TY> t;
    float h[9][TY + 4][TX + 1];
};

template <bool H>
__global__ void __launch_bounds__(256) k_guided_fast(Img in, Img out, ParamRef nlf, StrengthRef sr, int level) {
    __shared__ GuidedTile sm;
    __shared__ float mt[kMapStageMax];
    const TileId tl = xcd_tile();
    const int x0 = tl.x * TX, y0 = tl.y * TY;
    const int tid = threadIdx.y * 32 + threadIdx.x;
    sm.t.load<H>(in, x0, y0, tid);
    const MapStage ms = stage_strength<TX, TY>(sr, level, x0, y0, mt, tid);
    __syncthreads();
    {
        constexpr int N = (TY + 4) * TX, IT = (N + 255) / 256;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * 256;
            if (i >= N) break;
            const int r = i / TX, c = i % TX;
            float hs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int d = 0; d < 5; d++) {
                const float a = sm.t.x[r][c + d], b = sm.t.y[r][c + d], e = sm.t.z[r][c + d];
                const float v[9] = {a, b, e, a * a, a * b, a * e, b * b, b * e, e * e};
#pragma unroll
                for (int q = 0; q < 9; q++) hs[q] += v[q];
            }
#pragma unroll
            for (int q = 0; q < 9; q++) sm.h[q][r][c] = hs[q];
        }
    }
    __syncthreads();
    const float va0 = nlf.get(0), va1 = nlf.get(1), va2 = nlf.get(2);
    const float vb0 = nlf.get(3), vb1 = nlf.get(4), vb2 = nlf.get(5);
    const float norm = 1.0f / 25.0f;
#pragma unroll 1
    for (int rr = 0; rr < 2; rr++) {
        const int tx = threadIdx.x, ty = threadIdx.y + 8 * rr;
        const int x = x0 + tx, y = y0 + ty;
        if (x >= out.w || y >= out.h) continue;
        const int X = tx + HR, Y = ty + HR;
        const float ix = sm.t.x[Y][X], iy = sm.t.y[Y][X], iz = sm.t.z[Y][X];
        float eps0 = va0 + vb0 * ix, eps1 = va1 + vb1 * ix, eps2 = va2 + vb2 * ix;
        const float m = strength_at_staged(sr, ms, mt, level, x, y);
        if (m > 0.0f) {
            // learned strength: epsilon times m
            eps0 = eps0 * m;
            eps1 = eps1 * m;
            eps2 = eps2 * m;
        }
        float acc[9];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            float a = 0.0f;
#pragma unroll
            for (int j = 0; j < 5; j++) a += sm.h[q][ty + j][tx];
            acc[q] = a;
        }
        float m0 = acc[0], m1 = acc[1], m2 = acc[2];
        float rr_ = acc[3], rg = acc[4], rb = acc[5], gg = acc[6], gb = acc[7], bb = acc[8];
        m0 *= norm; m1 *= norm; m2 *= norm;
        rr_ *= norm; rg *= norm; rb *= norm; gg *= norm; gb *= norm; bb *= norm;
        const float vrr = rr_ - m0 * m0, vrg = rg - m0 * m1, vrb = rb - m0 * m2;
        const float vgg = gg - m1 * m1, vgb = gb - m1 * m2, vbb = bb - m2 * m2;
        const float vrr_e = vrr + eps0, vgg_e = vgg + eps1, vbb_e = vbb + eps2;
        float invrr = vgg_e * vbb_e - vgb * vgb;
        float invrg = vgb * vrb - vrg * vbb_e;
        float invrb = vrg * vgb - vgg_e * vrb;
        float invgg = vrr_e * vbb_e - vrb * vrb;
        float invgb = vrb * vrg - vrr_e * vgb;
        float invbb = vrr_e * vgg_e - vrg * vrg;
        const float invCovDet = 1.0f / (invrr * vrr_e + invrg * vrg + invrb * vrb);
        invrr *= invCovDet; invrg *= invCovDet; invrb *= invCovDet;
        invgg *= invCovDet; invgb *= invCovDet; invbb *= invCovDet;
        const float cr[3] = {vrr, vrg, vrb}, cg[3] = {vrg, vgg, vgb}, cb[3] = {vrb, vgb, vbb};
        const float mI[3] = {m0, m1, m2};
        float ar[3], ag[3], ab[3], q[3];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            ar[c] = invrr * cr[c] + invrg * cg[c] + invrb * cb[c];
            ag[c] = invrg * cr[c] + invgg * cg[c] + invgb * cb[c];
            ab[c] = invrb * cr[c] + invgb * cg[c] + invbb * cb[c];
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float b = mI[c] - ar[c] * m0 - ag[c] * m1 - ab[c] * m2;
            q[c] = ar[c] * ix + ag[c] * iy + ab[c] * iz + b;
        }
        st4_<H>(out, x, y, make_float4(q[0], q[1], q[2], 0.0f));
    }
}

void launch_guided_fast(const Img& in, const Img& out, const ParamRef& nlf, hipStream_t st, const StrengthRef* srp,
                        int level) {
    dim3 blk(32, 8), grd((out.w + TX - 1) / TX, (out.h + TY - 1) / TY);
    const StrengthRef sr = srp ? *srp : StrengthRef{nullptr, 0, 0, nullptr, 0.0f};
    with_storage(in, [&](auto H) {
        hipLaunchKernelGGL((k_guided_fast<H()>), grd, blk, 0, st, in, out, nlf, sr, level);
    });
}

// -------------------------------------------------------------- convertTosRGB
// (the per-pixel tone map is in tone.h)
// convertTosRGB; LTMB: with the localToneMapping mask (R32F, same size)
template <bool LTMB>
__global__ void __launch_bounds__(256) k_tone(Img in, Img mask, Img out8, Img outf, ToneArgs a) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= out8.w || y >= out8.h) return;
    float v[3];
    tone_pre(ldt4(in, x, y), a, v);
    tone_store_v<LTMB>(out8, outf, x, y, v, a, LTMB ? row_f(mask, y)[x] : 1.0f);
}

// ------------------------------------------------------ localToneMappingMask
// cl/demosaic.cl:1024-1107.  The guide (an RGBA image read through .x;
// pyramid level 3 in the pipeline) is sampled by the linear sampler at nine
// taps per mask texel: offsets -1.5, 0.5, 2.0 (+ 0.5) texels OF THE MASK per
// axis, box weights 2/5, 2/5, 1/5 per axis.  The "variance" is
// sum((w s)^2) - mean^2 as in the reference; min(L, 1) follows the OpenCL
// definition (1 if 1 < L, else L: NaN stays NaN).  The per-tap division by
// var + eps is a multiplication by (luma - mean) / (var + eps), and
// length(sqrt(2 rgb)) = sqrt(2r + 2g + 2b), NaN if a component is negative.
struct LtmArgs {
    Img guide;
    float eps, shadows, highlights, detail;
    float T[9];  // YCbCr -> camera RGB (the stand-alone operator)
};

// One axis of the taps: the sampler's texel and 8-bit weight per offset.
// Non-decreasing in x and in the offset (float rounding is monotone), so a
// block's taps lie between those of its first pixel's first offset and its
// last pixel's last offset.
__device__ __forceinline__ void ltm_axis(int x, float nx, int gsize, int i0[3], float a[3]) {
    const float px = (float)x * nx;
    const float off[3] = {-1.0f, 1.0f, 2.5f};  // offset + 0.5
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float u = (px + off[k] * nx) * (float)gsize - 0.5f;
        const float fu = floorf(u);
        a[k] = floorf((u - fu) * 256.0f + 0.5f) * 0.00390625f;
        i0[k] = (int)fu;
    }
}

constexpr int kLtmLW = 40, kLtmLH = 8;  // LDS guide footprint of a 64x4 block (ratios down to 1/2)

struct LtmGuideLds {
    const float* s;
    int i0, j0;
    __device__ __forceinline__ float at(int i, int j) const { return s[(j - j0) * kLtmLW + (i - i0)]; }
};
struct LtmGuideMem {
    Img g;
    __device__ __forceinline__ float at(int i, int j) const {
        return ld4(g, i, j).x;
    }
};

struct LtmFoot {
    int i0, j0;
    bool lds;
};

// The taps (ltm_axis) of a 64x4 block's columns and rows, computed once per
// column and row instead of once per pixel.
struct LtmAxes {
    int i[64][3], j[4][3];
    float a[64][3], b[4][3];
};

// For a 64x4 block at (x0, y0): the guide texels (clamped to the edge) under
// the taps of the pixels [x0, x1] x [y0, y1], staged in LDS when they fit
// (block-uniform), and the block's column and row taps.
__device__ __forceinline__ LtmFoot ltm_stage(float* sh, LtmAxes* ax, int x0, int x1, int y0, int y1, float nx,
                                             float ny, const Img& g) {
    int ia[3], ib[3], ja[3], jb[3];
    float t[3];
    ltm_axis(x0, nx, g.w, ia, t);
    ltm_axis(x1, nx, g.w, ib, t);
    ltm_axis(y0, ny, g.h, ja, t);
    ltm_axis(y1, ny, g.h, jb, t);
    LtmFoot f;
    f.i0 = ia[0];
    f.j0 = ja[0];
    const int nc = ib[2] + 2 - ia[0], nr = jb[2] + 2 - ja[0];
    f.lds = nc <= kLtmLW && nr <= kLtmLH;
    const int tid = threadIdx.x + threadIdx.y * blockDim.x, nt = blockDim.x * blockDim.y;
    if (tid < 64)
        ltm_axis(x0 + tid, nx, g.w, ax->i[tid], ax->a[tid]);
    else if (tid < 68)
        ltm_axis(y0 + tid - 64, ny, g.h, ax->j[tid - 64], ax->b[tid - 64]);
    if (f.lds) {
        const LtmGuideMem m{g};
        for (int k = tid; k < nc * nr; k += nt) {
            const int r = k / nc, c = k - r * nc;
            sh[r * kLtmLW + c] = m.at(f.i0 + c, f.j0 + r);
        }
    }
    __syncthreads();
    return f;
}

// pixel (lx, ly) of the block whose taps ax holds
template <class G>
__device__ __forceinline__ float ltm_mask_px(const G& g, const LtmAxes& ax, int lx, int ly, float luma,
                                             const float rgb[3], const LtmArgs& a) {
    int ii[3], jj[3];
    float fa[3], fb[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        ii[k] = ax.i[lx][k];
        fa[k] = ax.a[lx][k];
        jj[k] = ax.j[ly][k];
        fb[k] = ax.b[ly][k];
    }
    const float W[3][3] = {{0.16f, 0.16f, 0.08f}, {0.16f, 0.16f, 0.08f}, {0.08f, 0.08f, 0.04f}};  // [y][x]
    float s[9], sum = 0.0f, sumSq = 0.0f;
#pragma unroll
    for (int j = 0; j < 3; j++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float A = fa[i], B = fb[j];
            const float w00 = (1.0f - A) * (1.0f - B), w10 = A * (1.0f - B);
            const float w01 = (1.0f - A) * B, w11 = A * B;
            const float v = w00 * g.at(ii[i], jj[j]) + w10 * g.at(ii[i] + 1, jj[j]) +
                            w01 * g.at(ii[i], jj[j] + 1) + w11 * g.at(ii[i] + 1, jj[j] + 1);
            s[j * 3 + i] = v;
            const float ws = W[j][i] * v;
            sum += ws;
            sumSq += ws * ws;
        }
    }
    const float mean = sum, var = sumSq - sum * sum;
    const float q = (luma - mean) / (var + a.eps);
    float filtered = 0.0f, norm = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        const float sw = W[k / 3][k % 3] * (1.0f + (s[k] - mean) * q);
        filtered += sw * s[k];
        norm += sw;
    }
    const float illuminance = filtered / norm;
    const float reflectance = luma / illuminance;
    const float L = (rgb[0] >= 0.0f && rgb[1] >= 0.0f && rgb[2] >= 0.0f)
                        ? sqrtf(2.0f * rgb[0] + 2.0f * rgb[1] + 2.0f * rgb[2])
                        : __int_as_float(0x7fc00000);
    const float hc = 1.0f < L ? 1.0f : L;
    const float tc = mixf(a.shadows, a.highlights, hc);
    return clim_powf(illuminance, 1.0f / tc) * clim_powf(reflectance, a.detail) / luma;
}

__device__ __forceinline__ float ltm_mask_at(const LtmFoot& f, const float* sh, const LtmAxes& ax, int lx, int ly,
                                             float luma, const float rgb[3], const LtmArgs& a) {
    if (f.lds) return ltm_mask_px(LtmGuideLds{sh, f.i0, f.j0}, ax, lx, ly, luma, rgb, a);
    return ltm_mask_px(LtmGuideMem{a.guide}, ax, lx, ly, luma, rgb, a);
}

__global__ void __launch_bounds__(256) k_ltm_mask(Img in, Img out, LtmArgs a) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x;
    const int y = tl.y * 4 + threadIdx.y;
    const float nx = 1.0f / (float)out.w, ny = 1.0f / (float)out.h;
    __shared__ float sh[kLtmLW * kLtmLH];
    __shared__ LtmAxes ax;
    const LtmFoot f = ltm_stage(sh, &ax, tl.x * 64, min(tl.x * 64 + 63, out.w - 1), tl.y * 4,
                                min(tl.y * 4 + 3, out.h - 1), nx, ny, a.guide);
    if (x >= out.w || y >= out.h) return;
    const float4 p = ldt4(in, x, y);
    const float rgb[3] = {a.T[0] * p.x + a.T[1] * p.y + a.T[2] * p.z, a.T[3] * p.x + a.T[4] * p.y + a.T[5] * p.z,
                          a.T[6] * p.x + a.T[7] * p.y + a.T[8] * p.z};
    wrow_f(out, y)[x] = ltm_mask_at(f, sh, ax, threadIdx.x, threadIdx.y, p.x, rgb, a);
}

static LtmArgs make_ltm_args(const Img& guide, const float ltm[4], const float* T) {
    LtmArgs a;
    a.guide = guide;
    a.eps = ltm[0];
    a.shadows = ltm[1];
    a.highlights = ltm[2];
    a.detail = ltm[3];
    for (int i = 0; i < 9; i++) a.T[i] = T ? T[i] : 0.0f;
    return a;
}

void launch_ltm_mask(const Img& in, const Img& guide, const float ltm[4], const float T[9], const Img& out,
                     hipStream_t st) {
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    hipLaunchKernelGGL(k_ltm_mask, grd, blk, 0, st, in, out, make_ltm_args(guide, ltm, T));
}

void launch_tone(const Img& in, const Img& mask, const Img& out8, const Img& outf, const float* pre,
                 const float cam[9], float contrast, float saturation, float slope, float exposureBias, float blacks,
                 hipStream_t st) {
    const ToneArgs a = make_tone_args(pre, cam, contrast, saturation, slope, exposureBias, blacks);
    dim3 blk(64, 4), grd((out8.w + 63) / 64, (out8.h + 3) / 4);
    if (mask.p)
        hipLaunchKernelGGL(k_tone<true>, grd, blk, 0, st, in, mask, out8, outf, a);
    else
        hipLaunchKernelGGL(k_tone<false>, grd, blk, 0, st, in, mask, out8, outf, a);
}

// ------------------------------------------------------------ reassembleImage
// reassembleImage (cl/demosaic.cl:867-896) of texel (x, y) whose denoised
// value is p0 (nx, ny: 1 / the level's size): p0 - (noisy - reassembled) of
// the next level sampled at the texel, the luma blended towards the
// reassembled next level by the detail sharpening.  EXACT2X: half_n1 /
// half_d1 (i, j) give the next level's noisy and reassembled texels (clamped
// to the image); otherwise n1 and d1 are sampled.
template <bool EXACT2X, bool H, typename N1, typename D1>
__device__ __forceinline__ float4 reassemble_px(float4 p0, int x, int y, float nx, float ny, float sharpening0,
                                                const ParamRef& nlf, const Img& n1, const Img& d1, N1 half_n1,
                                                D1 half_d1) {
    const float px = ((float)x + 0.5f) * nx, py = ((float)y + 0.5f) * ny;
    float4 a1, b1;
    // exact 2x: the taps of the samples at px, py (i0, j0; weights a, b); those
    // at px +- nx, py +- ny (the sharpening gradient) sit at texels x/2 and
    // x/2 - 1 (y/2, y/2 - 1), rounded down, with weight 1 - a (1 - b)
    const int i0 = (x & 1) ? (x - 1) >> 1 : (x >> 1) - 1, j0 = (y & 1) ? (y - 1) >> 1 : (y >> 1) - 1;
    const float a = (x & 1) ? 0.25f : 0.75f, b = (y & 1) ? 0.25f : 0.75f;
    if (EXACT2X) {
        a1 = bilinear4_at(half_n1, i0, j0, a, b);
        b1 = bilinear4_at(half_d1, i0, j0, a, b);
    } else {
        a1 = bilinear4<H>(n1, px, py);
        b1 = bilinear4<H>(d1, px, py);
    }
    float4 d;
    d.x = p0.x - (a1.x - b1.x);
    d.y = p0.y - (a1.y - b1.y);
    d.z = p0.z - (a1.z - b1.z);
    d.w = p0.w - (a1.w - b1.w);
    float sharpening = sharpening0;
    if (sharpening > 1.0f) {
        float dx, dy;
        if (EXACT2X) {
            const float ax = 1.0f - a, by = 1.0f - b;
            const int xh = x >> 1, yh = y >> 1;
            dx = (bilinear4_at(half_d1, xh, j0, ax, b).x - bilinear4_at(half_d1, xh - 1, j0, ax, b).x) / 2.0f;
            dy = (bilinear4_at(half_d1, i0, yh, a, by).x - bilinear4_at(half_d1, i0, yh - 1, a, by).x) / 2.0f;
        } else {
            dx = (bilinear4<H>(d1, px + nx, py).x - bilinear4<H>(d1, px - nx, py).x) / 2.0f;
            dy = (bilinear4<H>(d1, px, py + ny).x - bilinear4<H>(d1, px, py - ny).x) / 2.0f;
        }
        const float threshold = 0.25f * sqrtf(nlf.get(0) + nlf.get(3) * b1.x);
        const float detail = smoothstep(0.25f * threshold, threshold, len2(dx, dy)) *
                             (1.0f - smoothstep(0.95f, 1.0f, d.x)) * (0.6f + 0.4f * smoothstep(0.0f, 0.1f, d.x));
        sharpening = 1.0f + (sharpening - 1.0f) * detail;
    }
    d.x = mixf(b1.x, d.x, sharpening);
    return d;
}

// reassembleImage, cl/demosaic.cl:867-896.  nlf: (A_Y, B_Y) at nlf.get(0), nlf.get(3).
// out may alias d0 (each texel is read, then written, by one thread).
// TONE: the level-0 pass of the pipeline, which hands its result straight to
// convertTosRGB (out8 / outf) instead of writing it.  LTM (with TONE): the
// localToneMapping mask of the result (guide: pyramid level 3) is computed
// in the same thread and applied by the tone map (cpp/raw_converter.cpp:164-174).
template <bool EXACT2X, bool TONE, bool LTM, bool H>
__global__ void __launch_bounds__(256) k_reassemble(Img d0, Img n1, Img d1, Img out, float sharpening0,
                                                    ParamRef nlf, Img out8, Img outf, ToneArgs ta, LtmArgs la) {
    const TileId tl = xcd_tile();
    const int x = tl.x * 64 + threadIdx.x;
    const int y = tl.y * 4 + threadIdx.y;
    const float nx = 1.0f / (float)out.w, ny = 1.0f / (float)out.h;
    __shared__ float gsh[LTM ? kLtmLW * kLtmLH : 1];
    __shared__ LtmAxes gax[1];  // (dropped by the compiler without LTM)
    LtmFoot foot{0, 0, false};
    if (LTM)
        foot = ltm_stage(gsh, gax, tl.x * 64, min(tl.x * 64 + 63, out.w - 1), tl.y * 4, min(tl.y * 4 + 3, out.h - 1),
                         nx, ny, la.guide);
    // exact 2x, without the tone map: the half-size texels the block's 64x4
    // pixels read (columns x0/2 - 1 .. x0/2 + 32, rows y0/2 - 1 .. y0/2 + 2,
    // clamped to the image) of both half-size images, staged once.  (The
    // level-0 pass with the tone map measured slower staged: its own VALU
    // work hides the loads, and the barrier costs more than the clamps.)
    constexpr bool STAGE = EXACT2X && !TONE;
    constexpr int RW = 34, RH = 4;
    __shared__ float4 hs[STAGE ? 2 : 1][STAGE ? RH : 1][STAGE ? RW : 1];
    const int hx0 = tl.x * 32 - 1, hy0 = tl.y * 2 - 1;
    if (STAGE) {
        const int tid = threadIdx.y * 64 + threadIdx.x;
        for (int k = tid; k < 2 * RH * RW; k += 256) {
            const int m = k / (RH * RW), r = (k / RW) % RH, c = k % RW;
            hs[m][r][c] = ld4_<H>(m ? d1 : n1, hx0 + c, hy0 + r);
        }
        __syncthreads();
    }
    auto half_n1 = [&](int i, int j) { return STAGE ? hs[0][j - hy0][i - hx0] : ld4_<H>(n1, i, j); };
    auto half_d1 = [&](int i, int j) { return STAGE ? hs[STAGE ? 1 : 0][j - hy0][i - hx0] : ld4_<H>(d1, i, j); };
    if (x >= out.w || y >= out.h) return;
    const float4 d = reassemble_px<EXACT2X, H>(ldt4_<H>(d0, x, y), x, y, nx, ny, sharpening0, nlf, n1, d1, half_n1,
                                               half_d1);
    if (TONE && LTM) {
        float v[3];
        tone_pre(d, ta, v);  // the camera RGB of the mask is the tone map's input
        const float boost = ltm_mask_at(foot, gsh, gax[0], threadIdx.x, threadIdx.y, d.x, v, la);
        tone_store_v<true>(out8, outf, x, y, v, ta, boost);
    } else if (TONE) {
        tone_store(out8, outf, x, y, d, ta);
    } else {
        st4_<H>(out, x, y, d);
    }
}

void launch_reassemble(const Img& d0, const Img& n1, const Img& d1, const Img& out, float sharpening,
                       const ParamRef& nlf, hipStream_t st, const ToneLaunch* tone) {
    dim3 blk(64, 4), grd((out.w + 63) / 64, (out.h + 3) / 4);
    const bool exact = out.w == 2 * n1.w && out.h == 2 * n1.h && n1.w == d1.w && n1.h == d1.h &&
                       out.w <= kUp2xMax && out.h <= kUp2xMax;
    ToneArgs ta{};
    LtmArgs la{};
    Img o8{nullptr, 0, 0, 0}, of{nullptr, 0, 0, 0};
    const bool ltm = tone && tone->ltm_guide.p;
    if (tone) {
        ta = make_tone_args(tone->pre, tone->cam, tone->contrast, tone->saturation, tone->slope, tone->exposureBias,
                            tone->blacks);
        o8 = tone->out8;
        of = tone->outf;
        if (ltm) la = make_ltm_args(tone->ltm_guide, tone->ltm, nullptr);
    }
    // the three level images (and out) share their storage
    with_storage(d0, [&](auto H) {
#define CLIM_REASSEMBLE(E, T, L)                                                                                  \
    hipLaunchKernelGGL((k_reassemble<E, T, L, H()>), grd, blk, 0, st, d0, n1, d1, out, sharpening, nlf, o8, of, ta, \
                       la)
        if (exact && ltm)
            CLIM_REASSEMBLE(true, true, true);
        else if (exact && tone)
            CLIM_REASSEMBLE(true, true, false);
        else if (exact)
            CLIM_REASSEMBLE(true, false, false);
        else if (ltm)
            CLIM_REASSEMBLE(false, true, true);
        else if (tone)
            CLIM_REASSEMBLE(false, true, false);
        else
            CLIM_REASSEMBLE(false, false, false);
#undef CLIM_REASSEMBLE
    });
}

// ------------------------------------------ denoiseImage + reassembleImage
// One FP32 pyramid level denoised and reassembled in one launch: the 32x16
// tile of k_bilateral with the next level's noisy and reassembled texels
// under it (columns x0/2 - 1 .. x0/2 + 16, rows y0/2 - 1 .. y0/2 + 8,
// clamped to the image) staged before the one barrier; each denoised texel is
// reassembled as k_reassemble<true, false, false, false> does and written.
// The values are those of the two kernels in turn (the denoised FP32 texel
// goes from the filter to the reassembly in registers instead of through
// HBM), and out may not alias the level.
namespace {
constexpr int BR_X = TX / 2 + 2, BR_Y = TY / 2 + 2;
}

template <bool SAME_EDGE>
__global__ void __launch_bounds__(256) k_bilateral_reassemble(Img in, Img n1, Img d1, Img out, ParamRef nlf,
                                                              float chromaBoost, float gradientBoost,
                                                              float sharpening, StrengthRef sr, int level) {
    __shared__ LevelTile4<TX, TY> t;
    __shared__ float mt[kMapStageMax];
    __shared__ float rcp[26];
    __shared__ float4 hs[2][BR_Y][BR_X];
    const TileId tl = xcd_tile();
    const int x0 = tl.x * TX, y0 = tl.y * TY;
    const int tid = threadIdx.y * 32 + threadIdx.x;
    const int hx0 = x0 / 2 - 1, hy0 = y0 / 2 - 1;
    const MapStage ms = bilateral_stage<false>(t, mt, rcp, in, x0, y0, sr, level, tid);
    for (int k = tid; k < 2 * BR_Y * BR_X; k += 256) {
        const int m = k / (BR_Y * BR_X), r = (k / BR_X) % BR_Y, c = k % BR_X;
        hs[m][r][c] = ld4_<false>(m ? d1 : n1, hx0 + c, hy0 + r);
    }
    __syncthreads();
    const float nx = 1.0f / (float)out.w, ny = 1.0f / (float)out.h;
    auto half_n1 = [&](int i, int j) { return hs[0][j - hy0][i - hx0]; };
    auto half_d1 = [&](int i, int j) { return hs[1][j - hy0][i - hx0]; };
    bilateral_compute<SAME_EDGE>(t, ms, mt, rcp, out.w, out.h, x0, y0, nlf, chromaBoost, gradientBoost, sr, level,
                                 [&](int x, int y, float4 v) {
                                     st4_<false>(out, x, y,
                                                 reassemble_px<true, false>(v, x, y, nx, ny, sharpening, nlf, n1, d1,
                                                                            half_n1, half_d1));
                                 });
}

bool bilateral_reassemble_fits(const Img& in, const Img& n1, const Img& d1, const Img& out) {
    return !in.half && !n1.half && !d1.half && !out.half && out.w == in.w && out.h == in.h && in.w == 2 * n1.w &&
           in.h == 2 * n1.h && n1.w == d1.w && n1.h == d1.h && in.w <= kUp2xMax && in.h <= kUp2xMax;
}

void launch_bilateral_reassemble(const Img& in, const Img& n1, const Img& d1, const Img& out, const ParamRef& nlf,
                                 float chromaBoost, float gradientBoost, float sharpening, const StrengthRef& sr,
                                 int level, hipStream_t st) {
    dim3 blk(32, 8), grd((in.w + TX - 1) / TX, (in.h + TY - 1) / TY);
    // the edges as the kernel computes them (lumaBoost = 1 without gradient boost)
    float e0 = 1.0f * 1.41421356f, e1 = chromaBoost * 1.41421356f;
    e0 = e0 * e0;
    e1 = e1 * e1;
    if (!(gradientBoost > 1.0f) && e0 == e1)
        hipLaunchKernelGGL(k_bilateral_reassemble<true>, grd, blk, 0, st, in, n1, d1, out, nlf, chromaBoost,
                           gradientBoost, sharpening, sr, level);
    else
        hipLaunchKernelGGL(k_bilateral_reassemble<false>, grd, blk, 0, st, in, n1, d1, out, nlf, chromaBoost,
                           gradientBoost, sharpening, sr, level);
}

}  // namespace clim
