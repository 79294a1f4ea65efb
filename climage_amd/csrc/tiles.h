// tiles.h - LDS tiles with 5x5 window statistics, shared by the statistics,
// regression and filter kernels.
//
// Window sums follow the path's definition: the vertical sum of five
// horizontal 5-tap sums, each a sequential sum starting from zero,
//   S = 0 + H(-2) + H(-1) + H(0) + H(1) + H(2),  H(j) = 0 + x(j,-2) + ... + x(j,2),
// so a horizontal sum is computed once per tile row and reused by the five
// output rows that need it.
#pragma once

#include "device_common.h"

namespace clim {

// ------------------------------------------------------------ RGBA levels
// TXxTY outputs, 256 threads, a 2-texel halo; only x, y, z enter the filters.
template <int TX, int TY>
struct LevelTile {
    static constexpr int SX = TX + 4, SY = TY + 4;
    float x[SY][SX + 1], y[SY][SX + 1], z[SY][SX + 1];

    // the tile's texels in registers (fetch, then put)
    using Regs = Staged<SX * SY, 256, float4>;
    // H: the level is stored as FP16
    template <bool H = false>
    __device__ __forceinline__ static void fetch(Regs& r, const Img& in, int x0, int y0, int tid) {
        r.fetch(tid, [&](int i) { return ld4_<H>(in, x0 - 2 + i % SX, y0 - 2 + i / SX); });
    }
    __device__ __forceinline__ void put(const Regs& r, int tid) {
        r.put(tid, [&](int i, float4 p) {
            x[i / SX][i % SX] = p.x;
            y[i / SX][i % SX] = p.y;
            z[i / SX][i % SX] = p.z;
        });
    }
    template <bool H = false>
    __device__ __forceinline__ void load(const Img& in, int x0, int y0, int tid) {
        Regs r;
        fetch<H>(r, in, x0, y0, tid);
        put(r, tid);
    }
};

// The same tile with the texels interleaved: one 16-byte LDS read per tap
// and constant offsets from one address for a whole 5x5 window.
template <int TX, int TY>
struct LevelTile4 {
    static constexpr int SX = TX + 4, SY = TY + 4;
    float4 p[SY][SX];
    template <bool H = false>
    __device__ __forceinline__ void load(const Img& in, int x0, int y0, int tid) {
        // a tile whose halo lies inside the image reads without clamps (the
        // same texels)
        if (x0 >= 2 && y0 >= 2 && x0 + SX - 2 <= in.w && y0 + SY - 2 <= in.h)
            stage<SX * SY, 256>(
                tid, [&](int i) { return ldt4_<H>(in, x0 - 2 + i % SX, y0 - 2 + i / SX); },
                [&](int i, float4 v) { p[i / SX][i % SX] = v; });
        else
            stage<SX * SY, 256>(
                tid, [&](int i) { return ld4_<H>(in, x0 - 2 + i % SX, y0 - 2 + i / SX); },
                [&](int i, float4 v) { p[i / SX][i % SX] = v; });
    }
    // the tile's texels in registers (fetch, then put)
    using Regs = Staged<SX * SY, 256, float4>;
    template <bool H = false>
    __device__ __forceinline__ static void fetch(Regs& r, const Img& in, int x0, int y0, int tid) {
        r.fetch(tid, [&](int i) { return ld4_<H>(in, x0 - 2 + i % SX, y0 - 2 + i / SX); });
    }
    __device__ __forceinline__ void put(const Regs& r, int tid) {
        r.put(tid, [&](int i, float4 v) { p[i / SX][i % SX] = v; });
    }
};

// Horizontal 5-tap sums of the level tile for all SY rows: h[0..2] the
// channel sums, h[3..5] the sums of squares (noiseStatistics, cl/demosaic.cl:1267-1286).
// PAIRS (TX = 32): each thread finishes two vertically adjacent outputs
// (stats2), which share four of their five rows of sums.  A wave then reads
// rows two apart in its two halves, so the rows are 48 floats long (two
// rows = 96 words, half the 64 banks apart), and the horizontal pass writes
// rows two apart per wave too (row_of); without PAIRS the rows are unpadded
// (one output row per half-wave, rows one apart = 32 words).
template <int TX, int TY, bool PAIRS = false>
struct StatsRows {
    static constexpr int SY = TY + 4;
    static constexpr int RS = PAIRS ? TX + TX / 2 : TX;
    static_assert(!PAIRS || (TX == 32 && SY % 4 == 0), "paired rows: 32-wide tiles, rows in fours");
    float h[6][SY][RS];

    // the tile row of work item i (one row per half-wave)
    __device__ __forceinline__ static int row_of(int i) {
        const int hw = i / TX;
        return PAIRS ? (hw & ~3) | ((hw & 1) << 1) | ((hw >> 1) & 1) : hw;
    }

    __device__ __forceinline__ void compute(const LevelTile<TX, TY>& t, int tid) {
        constexpr int N = SY * TX, IT = (N + 255) / 256;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * 256;
            if (i >= N) break;
            const int r = row_of(i), c = i % TX;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
#pragma unroll
            for (int d = 0; d < 5; d++) {
                const float a = t.x[r][c + d], b = t.y[r][c + d], e = t.z[r][c + d];
                s0 += a; q0 += a * a;
                s1 += b; q1 += b * b;
                s2 += e; q2 += e * e;
            }
            h[0][r][c] = s0; h[1][r][c] = s1; h[2][r][c] = s2;
            h[3][r][c] = q0; h[4][r][c] = q1; h[5][r][c] = q2;
        }
    }

    // the same sums from the interleaved tile (one LDS read per texel)
    __device__ __forceinline__ void compute(const LevelTile4<TX, TY>& t, int tid) {
        constexpr int N = SY * TX, IT = (N + 255) / 256;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * 256;
            if (i >= N) break;
            const int r = row_of(i), c = i % TX;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, q0 = 0.0f, q1 = 0.0f, q2 = 0.0f;
#pragma unroll
            for (int d = 0; d < 5; d++) {
                const float4 v = t.p[r][c + d];
                s0 += v.x; q0 += v.x * v.x;
                s1 += v.y; q1 += v.y * v.y;
                s2 += v.z; q2 += v.z * v.z;
            }
            h[0][r][c] = s0; h[1][r][c] = s1; h[2][r][c] = s2;
            h[3][r][c] = q0; h[4][r][c] = q1; h[5][r][c] = q2;
        }
    }

    // (mean Y, var Y, var Cb, var Cr) from the six window sums
    __device__ __forceinline__ static float4 finish(const float (&s)[6]) {
        // x / 25.0f by clim_div25_nz: the sums start from +0, so no operand is -0
        return make_float4(clim_div25_nz(s[0]), clim_div25_nz(s[3] - clim_div25_nz(s[0] * s[0])),
                           clim_div25_nz(s[4] - clim_div25_nz(s[1] * s[1])),
                           clim_div25_nz(s[5] - clim_div25_nz(s[2] * s[2])));
    }

    // (mean Y, var Y, var Cb, var Cr) of output (lx, ly)
    __device__ __forceinline__ float4 stats(int lx, int ly) const {
        float s[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            float a = 0.0f;
#pragma unroll
            for (int j = 0; j < 5; j++) a += h[k][ly + j][lx];
            s[k] = a;
        }
        return finish(s);
    }

    // outputs (lx, ly) and (lx, ly + 1) from six rows of sums, each output's
    // vertical sum in its own order (0 + rows ly .. ly + 4, 0 + rows ly + 1 .. ly + 5)
    __device__ __forceinline__ void stats2(int lx, int ly, float4& o0, float4& o1) const {
        float s0[6], s1[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            float r[6];
#pragma unroll
            for (int j = 0; j < 6; j++) r[j] = h[k][ly + j][lx];
            float a = 0.0f, b = 0.0f;
#pragma unroll
            for (int j = 0; j < 5; j++) {
                a += r[j];
                b += r[j + 1];
            }
            s0[k] = a;
            s1[k] = b;
        }
        o0 = finish(s0);
        o1 = finish(s1);
    }
};

// ------------------------------------------------------------ raw quads
// RTxRT output quads, 256 threads, 2-quad halo; texels are (R, G, B, G2).
template <int RT>
struct QuadTile {
    static constexpr int S = RT + 4;
    float4 q[S][S];
    float h[8][S][RT];  // horizontal sums: 4 channel sums, 4 sums of squares (unpadded rows, fewer bank conflicts)

    __device__ __forceinline__ void hsums(int tid) {
        constexpr int N = S * RT, IT = (N + 255) / 256;
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * 256;
            if (i >= N) break;
            const int r = i / RT, c = i % RT;
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, q0 = 0.0f, q1 = 0.0f, q2 = 0.0f, q3 = 0.0f;
#pragma unroll
            for (int d = 0; d < 5; d++) {
                const float4 v = q[r][c + d];
                s0 += v.x; q0 += v.x * v.x;
                s1 += v.y; q1 += v.y * v.y;
                s2 += v.z; q2 += v.z * v.z;
                s3 += v.w; q3 += v.w * v.w;
            }
            h[0][r][c] = s0; h[1][r][c] = s1; h[2][r][c] = s2; h[3][r][c] = s3;
            h[4][r][c] = q0; h[5][r][c] = q1; h[6][r][c] = q2; h[7][r][c] = q3;
        }
    }

    // rawNoiseStatistics (cl/demosaic.cl:1302-1324) of quad (lx, ly): mean, var
    __device__ __forceinline__ void stats(int lx, int ly, float* m, float* v) const {
        float s[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            float a = 0.0f;
#pragma unroll
            for (int j = 0; j < 5; j++) a += h[k][ly + j][lx];
            s[k] = a;
        }
#pragma unroll
        for (int c = 0; c < 4; c++) {
            // = x / 25.0f: the sums start from +0, so no operand is -0 (clim_div25_nz)
            m[c] = clim_div25_nz(s[c]);
            v[c] = clim_div25_nz(s[4 + c] - clim_div25_nz(s[c] * s[c]));
        }
    }
};

}  // namespace clim
