// k_nlf.hip - on-device noise-level-function regression.
//
// computeRawNoiseStatistics / computeNoiseStatistics
// (cpp/pyramidal_denoise.cpp:161-256, :259-362) write a statistics image, map
// it to the host and run three FP64 passes over it.  Here each pass is one
// launch that recomputes the 5x5 statistics of its tile in LDS (bit-identical
// to rawNoiseStatistics / noiseStatistics) and accumulates the regression
// sums directly, so no statistics image is written or read:
//   pass 0  sums (s_x, s_y, s_xx, s_xy, N) over in-range samples
//           -> A, B = least squares of var = A + B*mean (B, A floored at 1e-8)
//   pass 1  sum of squared residuals -> err2 = sqrt(sum / N)
//   pass 2  sums over inliers (residual^2 <= err2 in every channel) -> A, B
// Terms are accumulated in 128-bit fixed point (2^-90 resolution) with
// integer conversions (floor(t * 2^90) of the exact value), so the sums do not
// depend on the reduction order and equal the CPU oracle's.  The last block of
// each job combines the per-block partials and solves in FP64.  One launch
// serves several jobs (all pyramid levels).
#include <type_traits>

#include "kernels.h"
#include "tiles.h"

namespace clim {

namespace {
constexpr int RT = 16, RR = 2, RS = RT + 2 * RR;                  // raw: 16x16 quads + 2-quad halo
constexpr int LTX = 32, LTY = 16, LR = 2, LSX = LTX + 2 * LR, LSY = LTY + 2 * LR;  // level tile
}

// accumulator layout: per-channel means (raw) or one shared mean (level)
template <int NCH, bool SHARED_M>
struct Layout {
    static constexpr int K = SHARED_M ? 2 + 2 * NCH + 1 : 4 * NCH + 1;
    __device__ static constexpr int sx(int c) { return SHARED_M ? 0 : 4 * c; }
    __device__ static constexpr int sxx(int c) { return SHARED_M ? 1 : 4 * c + 2; }
    __device__ static constexpr int sy(int c) { return SHARED_M ? 2 + 2 * c : 4 * c + 1; }
    __device__ static constexpr int sxy(int c) { return SHARED_M ? 3 + 2 * c : 4 * c + 3; }
    static constexpr int n = K - 1;
};

__device__ __forceinline__ clim_fx128 wave_sum(clim_fx128 a) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) a = clim_fx_add(a, fx_shfl_xor(a, m));
    return a;
}

// block-wide sums of K fixed-point values; valid in thread 0
template <int K>
__device__ __forceinline__ void block_sum(clim_fx128* v, clim_fx128 (*lds)[K]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const clim_fx128 s = wave_sum(v[k]);
        if (lane == 0) lds[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; k++) {
            clim_fx128 s = lds[0][k];
            for (int wv = 1; wv < 4; wv++) s = clim_fx_add(s, lds[wv][k]);
            v[k] = s;
        }
    }
    __syncthreads();
}

template <int NCH, bool SHARED_M>
struct Acc {
    using L = Layout<NCH, SHARED_M>;
    clim_fx128 a[L::K];
    unsigned int n;
    double A[NCH], B[NCH], e2[NCH];

    __device__ void zero() {
#pragma unroll
        for (int k = 0; k < L::K; k++) a[k].hi = a[k].lo = 0;
        n = 0;
    }

    // one sample: means m (m[0] only when SHARED_M), variances v.  The range
    // test (mean in [0.001, 0.5], variance in [0, 0.001] as doubles,
    // cpp/pyramidal_denoise.cpp:151-159) is done on the float values against
    // kMeanMin = the smallest float >= 0.001 and kVarMax = the largest float
    // <= 0.001, which is the same test for float inputs.
    __device__ __forceinline__ void add(int pass, const float* m, const float* v) {
        constexpr float kMeanMin = 0x1.0624dep-10f, kVarMax = 0x1.0624dcp-10f;
        bool ok = true;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const float mm = m[SHARED_M ? 0 : c], vv = v[c];
            if (mm < kMeanMin || mm > 0.5f) ok = false;
            if (vv < 0.0f || vv > kVarMax) ok = false;
        }
        if (!ok) return;
        if (pass == 1) {
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const double diff = (A[c] + B[c] * (double)m[SHARED_M ? 0 : c]) - (double)v[c];
                a[c] = clim_fx_add(a[c], clim_fx_from_double_bits(diff * diff));
            }
            return;
        }
        if (pass == 2) {
            bool inlier = true;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const double diff = (A[c] + B[c] * (double)m[SHARED_M ? 0 : c]) - (double)v[c];
                if (!(diff * diff <= e2[c])) inlier = false;
            }
            if (!inlier) return;
        }
        // exact integer terms (clim_q_mean, clim_q_var)
        uint32_t mq[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            if (SHARED_M && c > 0) {
                mq[c] = mq[0];
                continue;
            }
            mq[c] = clim_q_mean(m[c]);
            a[L::sx(c)] = clim_fx_add(a[L::sx(c)], clim_fx_u64(mq[c]));
            a[L::sxx(c)] = clim_fx_add(a[L::sxx(c)], clim_fx_u64((uint64_t)mq[c] * mq[c]));
        }
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const uint64_t vq = clim_q_var(v[c]);
            a[L::sy(c)] = clim_fx_add(a[L::sy(c)], clim_fx_u64(vq));
            a[L::sxy(c)] = clim_fx_add(a[L::sxy(c)], clim_fx_mul32x64(mq[c], vq));
        }
        n++;
    }
};

template <int NCH, bool SHARED_M>
__device__ void fit_solve(const clim_fx128* tot, unsigned long long n, double* A, double* B) {
    using L = Layout<NCH, SHARED_M>;
    const double N = (double)n;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        const double s_x = clim_fx_to_double_scaled(tot[L::sx(c)], 0x1p-33);
        const double s_y = clim_fx_to_double_scaled(tot[L::sy(c)], 0x1p-64);
        const double s_xx = clim_fx_to_double_scaled(tot[L::sxx(c)], 0x1p-66);
        const double s_xy = clim_fx_to_double_scaled(tot[L::sxy(c)], 0x1p-97);
        B[c] = fmax((N * s_xy - s_x * s_y) / (N * s_xx - s_x * s_x), 1e-8);
        A[c] = fmax((s_y - B[c] * s_x) / N, 1e-8);
    }
}

struct LevelStatsTile {
    LevelTile<32, 16> t;
    StatsRows<32, 16> r;
};

template <int KIND>
__global__ void __launch_bounds__(256) k_fit(FitJobs jobs, int pass) {
    constexpr bool RAW = KIND != FIT_LEVEL;
    constexpr int NCH = RAW ? 4 : 3;
    constexpr bool SHARED_M = !RAW;
    using L = Layout<NCH, SHARED_M>;
    constexpr int K = L::K;
    __shared__ typename std::conditional<RAW, QuadTile<RT>, LevelStatsTile>::type sm;
    __shared__ clim_fx128 lds[4][K];
    __shared__ unsigned int s_last;

    int ji = 0;
    while (ji + 1 < jobs.njobs && (int)blockIdx.x >= jobs.j[ji + 1].block_base) ji++;
    const FitJob& jb = jobs.j[ji];
    FitState* st = jb.st;
    if (pass > 0 && st->status == 1) return;  // no sample in range: fallback already written
    const int bi = blockIdx.x - jb.block_base;
    const int tid = threadIdx.x;

    Acc<NCH, SHARED_M> acc;
    acc.zero();
    if (pass > 0) {
#pragma unroll
        for (int c = 0; c < NCH; c++) { acc.A[c] = st->A[c]; acc.B[c] = st->B[c]; acc.e2[c] = st->err2[c]; }
    }
    // The texels of the next tile are loaded into registers while the current
    // tile is reduced, so a block never waits on memory with nothing to do.
    constexpr int N = RAW ? RS * RS : LSX * LSY, IT = (N + 255) / 256;
    using Pre = typename std::conditional<KIND == FIT_RAW_U16, uint2, float4>::type;
    Pre pre[IT];
    const RawQuadLoader rl{jb.src, jb.b, jb.smul, jb.black};
    const int ntiles = jb.tiles_x * jb.tiles_y;
    auto origin = [&](int t, int& ox, int& oy) {
        ox = (t % jb.tiles_x) * (RAW ? RT : LTX) - (RAW ? RR : LR);
        oy = (t / jb.tiles_x) * (RAW ? RT : LTY) - (RAW ? RR : LR);
    };
    auto fetch = [&](int t) {
        int ox, oy;
        origin(t, ox, oy);
#pragma unroll
        for (int k = 0; k < IT; k++) {
            const int i = tid + k * 256;
            if (i >= N) break;
            if constexpr (KIND == FIT_RAW_U16) {
                pre[k] = rl.words(ox + i % RS, oy + i / RS);
            } else if constexpr (KIND == FIT_RAW_F32) {
                const int qx = ox + i % RS, qy = oy + i / RS;
                pre[k] = make_float4(ld1(jb.src, 2 * qx + jb.b.ox[0], 2 * qy + jb.b.oy[0]),
                                     ld1(jb.src, 2 * qx + jb.b.ox[1], 2 * qy + jb.b.oy[1]),
                                     ld1(jb.src, 2 * qx + jb.b.ox[2], 2 * qy + jb.b.oy[2]),
                                     ld1(jb.src, 2 * qx + jb.b.ox[3], 2 * qy + jb.b.oy[3]));
            } else {
                pre[k] = ld4(jb.src, ox + i % LSX, oy + i / LSX);
            }
        }
    };
    if (bi < ntiles) fetch(bi);
    for (int t = bi; t < ntiles; t += jb.nblocks) {
        const int tile_x = t % jb.tiles_x, tile_y = t / jb.tiles_x;
        __syncthreads();  // the previous tile's LDS reads are done
        {
            int ox, oy;
            origin(t, ox, oy);
#pragma unroll
            for (int k = 0; k < IT; k++) {
                const int i = tid + k * 256;
                if (i >= N) break;
                if constexpr (KIND == FIT_RAW_U16) {
                    sm.q[i / RS][i % RS] = rl.quad(pre[k], ox + i % RS, oy + i / RS);
                } else if constexpr (KIND == FIT_RAW_F32) {
                    sm.q[i / RS][i % RS] = pre[k];
                } else {
                    sm.t.x[i / LSX][i % LSX] = pre[k].x;
                    sm.t.y[i / LSX][i % LSX] = pre[k].y;
                    sm.t.z[i / LSX][i % LSX] = pre[k].z;
                }
            }
        }
        __syncthreads();
        if (t + jb.nblocks < ntiles) fetch(t + jb.nblocks);
        if constexpr (RAW) {
            // rawNoiseStatistics, cl/demosaic.cl:1288-1324
            const int qx0 = tile_x * RT, qy0 = tile_y * RT;
            sm.hsums(tid);
            __syncthreads();
            const int lx = tid % RT, ly = tid / RT;
            if (qx0 + lx >= jb.src.w / 2 || qy0 + ly >= jb.src.h / 2) continue;
            float m[4], v[4];
            sm.stats(lx, ly, m, v);
            acc.add(pass, m, v);
        } else {
            // noiseStatistics, cl/demosaic.cl:1267-1286
            const int x0 = tile_x * LTX, y0 = tile_y * LTY;
            sm.r.compute(sm.t, tid);
            __syncthreads();
#pragma unroll
            for (int rr = 0; rr < 2; rr++) {
                const int lx = tid % LTX, ly = tid / LTX + 8 * rr;
                if (x0 + lx >= jb.src.w || y0 + ly >= jb.src.h) continue;
                const float4 st4 = sm.r.stats(lx, ly);
                const float m[1] = {st4.x};
                const float v[3] = {st4.y, st4.z, st4.w};
                acc.add(pass, m, v);
            }
        }
    }
    acc.a[L::n].hi = 0;
    acc.a[L::n].lo = acc.n;

    block_sum<K>(acc.a, lds);
    if (tid == 0) {
        clim_fx128* part = jb.partials + (size_t)bi * K;
        for (int k = 0; k < K; k++) part[k] = acc.a[k];
        __threadfence();
        const unsigned int prev = atomicAdd(&st->counter, 1u);
        s_last = prev == (unsigned int)jb.nblocks - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();

    // last block of the job: combine the partials of every block
    clim_fx128 tot[K];
#pragma unroll
    for (int k = 0; k < K; k++) tot[k].hi = tot[k].lo = 0;
    for (int b = tid; b < jb.nblocks; b += 256) {
        const clim_fx128* pb = jb.partials + (size_t)b * K;
#pragma unroll
        for (int k = 0; k < K; k++) {
            clim_fx128 v;
            v.hi = __hip_atomic_load(&pb[k].hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v.lo = __hip_atomic_load(&pb[k].lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            tot[k] = clim_fx_add(tot[k], v);
        }
    }
    block_sum<K>(tot, lds);
    if (tid != 0) return;

    const unsigned long long cnt = tot[L::n].lo;
    float* out = st->out;
    if (pass == 0) {
        st->n0 = cnt;
        if (cnt == 0) {
            st->status = 1;
            for (int k = 0; k < 2 * NCH; k++) out[k] = jb.nlf_in[k];
        } else {
            st->status = 0;
            fit_solve<NCH, SHARED_M>(tot, cnt, st->A, st->B);
        }
    } else if (pass == 1) {
        for (int c = 0; c < NCH; c++) st->err2[c] = sqrt(clim_fx_to_double(tot[c]) / (double)st->n0);
    } else {
        if (cnt > 0) {
            fit_solve<NCH, SHARED_M>(tot, cnt, st->A, st->B);
            st->status = 0;
        } else {
            st->status = 2;
        }
        for (int c = 0; c < NCH; c++) {
            out[c] = (float)st->A[c];
            out[NCH + c] = (float)st->B[c];
        }
    }
    if (pass == 2 || st->status == 1) {
        for (int k = 0; k < 2 * NCH; k++) st->cal[k] = out[k] * jb.mult[k];
    }
    st->counter = 0;
}

// Fills tiles/blocks of the jobs (block_base, nblocks) and launches the three passes.
void launch_fit(FitJobs& jobs, int kind, hipStream_t st) {
    // blocks are shared out in proportion to the tiles of each job, about two
    // rounds of resident blocks in all (each job keeps at most
    // kFitBlocksPerJob partials)
    constexpr int kBudget = 2048;
    long long total = 0;
    for (int j = 0; j < jobs.njobs; j++) {
        FitJob& jb = jobs.j[j];
        if (kind == FIT_LEVEL) {
            jb.tiles_x = (jb.src.w + LTX - 1) / LTX;
            jb.tiles_y = (jb.src.h + LTY - 1) / LTY;
        } else {
            jb.tiles_x = (jb.src.w / 2 + RT - 1) / RT;
            jb.tiles_y = (jb.src.h / 2 + RT - 1) / RT;
        }
        total += (long long)jb.tiles_x * jb.tiles_y;
    }
    int base = 0;
    for (int j = 0; j < jobs.njobs; j++) {
        FitJob& jb = jobs.j[j];
        const long long tiles = (long long)jb.tiles_x * jb.tiles_y;
        long long nb = (tiles * kBudget + total - 1) / total;
        if (nb > tiles) nb = tiles;
        if (nb > kFitBlocksPerJob) nb = kFitBlocksPerJob;
        if (nb < 1) nb = 1;
        jb.nblocks = (int)nb;
        jb.block_base = base;
        base += jb.nblocks;
    }
    for (int pass = 0; pass < 3; pass++) {
        if (kind == FIT_RAW_U16)
            hipLaunchKernelGGL(k_fit<FIT_RAW_U16>, dim3(base), dim3(256), 0, st, jobs, pass);
        else if (kind == FIT_RAW_F32)
            hipLaunchKernelGGL(k_fit<FIT_RAW_F32>, dim3(base), dim3(256), 0, st, jobs, pass);
        else
            hipLaunchKernelGGL(k_fit<FIT_LEVEL>, dim3(base), dim3(256), 0, st, jobs, pass);
    }
}

}  // namespace clim
