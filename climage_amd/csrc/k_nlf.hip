// k_nlf.hip - on-device noise-level-function regression.
//
// computeRawNoiseStatistics / computeNoiseStatistics
// (cpp/pyramidal_denoise.cpp:161-256, :259-362) map the statistics image to
// the host and run three FP64 passes over it (sums, squared residuals, sums
// over inliers).  Here two launches stream over the statistics image in
// device memory:
//   first   exact sums of the grid terms (clim_q32_mean / clim_q32_var):
//           s_x, s_y, s_xx, s_xy and s_vv per channel, N, over in-range
//           samples -> A, B by least squares (B, A floored at 1e-8), and the
//           RMS residual err2 from the moments (clim_fit_rss_q), which
//           replaces the reference's residual pass
//   refit   the same sums over inliers (residual^2 <= err2 in every
//           channel) -> A, B
// The grid terms are 32-bit integers, their products 64-bit, and each thread
// keeps them in 64- and 96-bit accumulators; blocks combine 128-bit sums, so
// the totals do not depend on the reduction order and equal the CPU oracle's.
// Each pass is two launches: the streaming kernel, whose blocks write their
// partial sums, and a finalize kernel (one block per job) that combines them
// and solves in FP64.  The kernel boundary makes the partials visible; a
// device-scope fence per block would write back and invalidate the L2 of its
// XCD every time.  One launch serves several jobs (all pyramid levels); each
// block streams an even share of one job's rows.
#include "kernels.h"

namespace clim {

// per-thread sum of 64-bit terms (at most 2^32 of them)
struct U96 {
    uint64_t lo;
    uint32_t hi;
    __device__ __forceinline__ void add(uint64_t x) {
        lo += x;
        hi += lo < x ? 1u : 0u;
    }
    __device__ __forceinline__ clim_fx128 fx() const {
        clim_fx128 r;
        r.hi = hi;
        r.lo = lo;
        return r;
    }
};

// Accumulator layout of the block partials: per channel sx, sxx, sy, sxy
// (and svv in the first pass), then N.  Level fits share one mean, so sx and
// sxx come once, first.
template <int NCH, bool SHARED_M, bool REFIT>
struct Layout {
    static constexpr int P = REFIT ? 2 : 3;  // per-channel variance sums: sy, sxy (, svv)
    static constexpr int Q = P + 2;          // per-channel sums with own means
    static constexpr int K = (SHARED_M ? 2 + NCH * P : NCH * Q) + 1;
    __device__ static constexpr int sx(int c) { return SHARED_M ? 0 : c * Q; }
    __device__ static constexpr int sxx(int c) { return SHARED_M ? 1 : c * Q + 1; }
    __device__ static constexpr int sy(int c) { return SHARED_M ? 2 + c * P : c * Q + 2; }
    __device__ static constexpr int sxy(int c) { return SHARED_M ? 3 + c * P : c * Q + 3; }
    __device__ static constexpr int svv(int c) { return SHARED_M ? 4 + c * P : c * Q + 4; }
    static constexpr int n = K - 1;
};

__device__ __forceinline__ clim_fx128 wave_sum(clim_fx128 a) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) a = clim_fx_add(a, fx_shfl_xor(a, m));
    return a;
}

// Block-wide sums of K fixed-point values per thread: wave sums into
// lds[wave][k], then thread k < K adds the four waves into lds[0][k].
template <int K>
__device__ __forceinline__ void block_sum(const clim_fx128* v, clim_fx128 (*lds)[K]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        const clim_fx128 s = wave_sum(v[k]);
        if (lane == 0) lds[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        clim_fx128 t = lds[0][threadIdx.x];
        for (int wv = 1; wv < 4; wv++) t = clim_fx_add(t, lds[wv][threadIdx.x]);
        lds[0][threadIdx.x] = t;
    }
    __syncthreads();
}

// Per-thread sums of one pass.
template <int NCH, bool SHARED_M, bool REFIT>
struct Acc {
    static constexpr int NM = SHARED_M ? 1 : NCH;  // mean channels
    uint64_t sx[NM], sy[NCH];
    U96 sxx[NM], sxy[NCH], svv[REFIT ? 1 : NCH];
    unsigned int n;
    double A[NCH], B[NCH], e2[NCH];

    __device__ void zero() {
#pragma unroll
        for (int c = 0; c < NM; c++) { sx[c] = 0; sxx[c] = U96{0, 0}; }
#pragma unroll
        for (int c = 0; c < NCH; c++) { sy[c] = 0; sxy[c] = U96{0, 0}; }
#pragma unroll
        for (int c = 0; c < (REFIT ? 1 : NCH); c++) svv[c] = U96{0, 0};
        n = 0;
    }

    // one sample: means m (m[0] only when SHARED_M), variances v.  The range
    // test (mean in [0.001, 0.5], variance in [0, 0.001] as doubles,
    // cpp/pyramidal_denoise.cpp:151-159) is done on the float values against
    // kMeanMin = the smallest float >= 0.001 and kVarMax = the largest float
    // <= 0.001, which is the same test for float inputs.  Terms of samples
    // that fail it (or, when refitting, are outliers) are zero, so every lane
    // of a wave runs the same instructions.
    __device__ __forceinline__ void add(const float* m, const float* v, bool inside) {
        constexpr float kMeanMin = 0x1.0624dep-10f, kVarMax = 0x1.0624dcp-10f;
        bool ok = inside;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const float mm = m[SHARED_M ? 0 : c], vv = v[c];
            ok = ok && mm >= kMeanMin && mm <= 0.5f && vv >= 0.0f && vv <= kVarMax;
        }
        if (REFIT) {
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const double diff = (A[c] + B[c] * (double)m[SHARED_M ? 0 : c]) - (double)v[c];
                ok = ok && diff * diff <= e2[c];
            }
        }
        uint32_t mq[NM];
#pragma unroll
        for (int c = 0; c < NM; c++) {
            mq[c] = ok ? clim_q32_mean(m[c]) : 0u;
            sx[c] += mq[c];
            sxx[c].add((uint64_t)mq[c] * mq[c]);
        }
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const uint32_t vq = ok ? clim_q32_var(v[c]) : 0u;
            sy[c] += vq;
            sxy[c].add((uint64_t)mq[SHARED_M ? 0 : c] * vq);
            if (!REFIT) svv[c].add((uint64_t)vq * vq);
        }
        n += ok ? 1u : 0u;
    }

    // the sums in the partial layout
    __device__ __forceinline__ void store(clim_fx128* out) const {
        using L = Layout<NCH, SHARED_M, REFIT>;
#pragma unroll
        for (int c = 0; c < NM; c++) {
            out[L::sx(c)] = clim_fx_u64(sx[c]);
            out[L::sxx(c)] = sxx[c].fx();
        }
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            out[L::sy(c)] = clim_fx_u64(sy[c]);
            out[L::sxy(c)] = sxy[c].fx();
            if (!REFIT) out[L::svv(c)] = svv[c].fx();
        }
        out[L::n] = clim_fx_u64(n);
    }
};

// RAW: mean and var images of (R, G, B, G2) quads; level: one statistics
// image (mean Y, var Y, var Cb, var Cr)
template <bool RAW, bool REFIT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_fit(FitJobs jobs) {
    constexpr int NCH = RAW ? 4 : 3;
    constexpr bool SHARED_M = !RAW;
    using L = Layout<NCH, SHARED_M, REFIT>;
    using AccT = Acc<NCH, SHARED_M, REFIT>;
    constexpr int K = L::K;
    constexpr int U = RAW ? 2 : 4;  // samples in flight per thread
    __shared__ clim_fx128 lds[4][K];

    int ji = 0;
    while (ji + 1 < jobs.njobs && (int)blockIdx.x >= jobs.j[ji + 1].block_base) ji++;
    const FitJob& jb = jobs.j[ji];
    FitState* st = jb.st;
    if (REFIT && st->status == 1) return;  // no sample in range: fallback already written
    const int bi = blockIdx.x - jb.block_base;
    const int tid = threadIdx.x;

    AccT acc;
    acc.zero();
    if (REFIT) {
#pragma unroll
        for (int c = 0; c < NCH; c++) { acc.A[c] = st->A[c]; acc.B[c] = st->B[c]; acc.e2[c] = st->err2[c]; }
    }
    // This block's share of the job: chunks [c0, c1) of 256*U consecutive
    // samples of a row (cpr chunks per row), U samples per thread per chunk.
    constexpr int CH = 256 * U;
    const int w = jb.m.w, cpr = (w + CH - 1) / CH;
    const long long nchunks = (long long)jb.m.h * cpr;
    const int c0 = (int)(nchunks * bi / jb.nblocks), c1 = (int)(nchunks * (bi + 1) / jb.nblocks);
    int y = c0 / cpr, xb = (c0 - y * cpr) * CH;
    for (int c = c0; c < c1; c++) {
        const float4* mrow = row_f4(jb.m, y);
        const float4* vrow = row_f4(jb.v, y);
        // unconditional loads of clamped columns (no branch, so all U are in
        // flight together); columns past the row end are masked
        float4 pm[U], pv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int x = min(xb + tid + 256 * u, w - 1);
            pm[u] = mrow[x];
            if (RAW) pv[u] = vrow[x];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool inside = xb + tid + 256 * u < w;
            if (RAW) {
                const float m[4] = {pm[u].x, pm[u].y, pm[u].z, pm[u].w};
                const float v[4] = {pv[u].x, pv[u].y, pv[u].z, pv[u].w};
                acc.add(m, v, inside);
            } else {
                const float m[1] = {pm[u].x};
                const float v[3] = {pm[u].y, pm[u].z, pm[u].w};
                acc.add(m, v, inside);
            }
        }
        xb += CH;
        if (xb >= w) {
            xb = 0;
            y++;
        }
    }
    clim_fx128 sums[K];
    acc.store(sums);

    block_sum<K>(sums, lds);
    // partials are stored k-major: partials[k * kFitBlocksPerJob + block]
    if (tid < K) jb.partials[(size_t)tid * kFitBlocksPerJob + bi] = lds[0][tid];
}

// One block per job: combines the partials of the job's blocks (plain loads
// of clamped indexes, masked, so that none waits for another) and solves.
template <bool RAW, bool REFIT>
__global__ void __launch_bounds__(256) k_fit_finalize(FitJobs jobs) {
    constexpr int NCH = RAW ? 4 : 3;
    constexpr bool SHARED_M = !RAW;
    using L = Layout<NCH, SHARED_M, REFIT>;
    constexpr int K = L::K;
    static_assert(kFitBlocksPerJob <= 4 * 256, "four partials per thread");
    __shared__ clim_fx128 lds[4][K];
    const FitJob& jb = jobs.j[blockIdx.x];
    FitState* st = jb.st;
    if (REFIT && st->status == 1) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll 3
    for (int k = 0; k < K; k++) {
        const clim_fx128* pk = jb.partials + (size_t)k * kFitBlocksPerJob;
        clim_fx128 s = {0, 0};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int b = tid + 256 * j;
            clim_fx128 v = pk[min(b, jb.nblocks - 1)];
            v.hi = b < jb.nblocks ? v.hi : 0;
            v.lo = b < jb.nblocks ? v.lo : 0;
            s = clim_fx_add(s, v);
        }
        s = wave_sum(s);
        if (lane == 0) lds[wave][k] = s;
    }
    __syncthreads();
    if (tid < K) {
        clim_fx128 t = lds[0][tid];
        for (int wv = 1; wv < 4; wv++) t = clim_fx_add(t, lds[wv][tid]);
        lds[0][tid] = t;
    }
    __syncthreads();
    if (tid != 0) return;
    const clim_fx128* tot = lds[0];
    const unsigned long long cnt = tot[L::n].lo;
    auto sums_of = [&](int c) {
        clim_qsums q;
        q.sx = tot[L::sx(c)];
        q.sxx = tot[L::sxx(c)];
        q.sy = tot[L::sy(c)];
        q.sxy = tot[L::sxy(c)];
        q.svv = REFIT ? clim_fx_u64(0) : tot[L::svv(c)];
        return q;
    };
    float* out = st->out;
    if (!REFIT) {
        st->n0 = cnt;
        if (cnt == 0) {
            st->status = 1;
            for (int k = 0; k < 2 * NCH; k++) out[k] = jb.nlf_in[k];
        } else {
            st->status = 0;
            for (int c = 0; c < NCH; c++) {
                const clim_qsums q = sums_of(c);
                clim_fit_solve_q(&q, cnt, &st->A[c], &st->B[c]);
                st->err2[c] = sqrt(clim_fit_rss_q(&q, cnt, st->A[c], st->B[c]) / (double)cnt);
            }
        }
    } else {
        if (cnt > 0) {
            for (int c = 0; c < NCH; c++) {
                const clim_qsums q = sums_of(c);
                clim_fit_solve_q(&q, cnt, &st->A[c], &st->B[c]);
            }
            st->status = 0;
        } else {
            st->status = 2;
        }
        for (int c = 0; c < NCH; c++) {
            out[c] = (float)st->A[c];
            out[NCH + c] = (float)st->B[c];
        }
    }
    if (REFIT || st->status == 1) {
        for (int k = 0; k < 2 * NCH; k++) st->cal[k] = out[k] * jb.mult[k];
    }
}

// Shares out the blocks: kFitGrid in all (every block resident at once), in
// proportion to the samples of each job, at least one and at most one per
// chunk of samples or kFitBlocksPerJob; then launches the two passes, each
// followed by its finalize kernel.
void launch_fit(FitJobs& jobs, bool raw, hipStream_t st) {
    constexpr int kFitGrid = 1024;
    const int ch = 256 * (raw ? 2 : 4);
    long long total = 0;
    for (int j = 0; j < jobs.njobs; j++) total += (long long)jobs.j[j].m.w * jobs.j[j].m.h;
    int base = 0;
    for (int j = 0; j < jobs.njobs; j++) {
        FitJob& jb = jobs.j[j];
        const long long n = (long long)jb.m.w * jb.m.h;
        const long long chunks = (long long)jb.m.h * ((jb.m.w + ch - 1) / ch);
        long long nb = (n * kFitGrid + total / 2) / total;
        if (nb > chunks) nb = chunks;
        if (nb > kFitBlocksPerJob) nb = kFitBlocksPerJob;
        if (nb < 1) nb = 1;
        jb.nblocks = (int)nb;
        jb.block_base = base;
        base += jb.nblocks;
    }
    const dim3 grid(base), fin(jobs.njobs), blk(256);
    if (raw) {
        hipLaunchKernelGGL((k_fit<true, false>), grid, blk, 0, st, jobs);
        hipLaunchKernelGGL((k_fit_finalize<true, false>), fin, blk, 0, st, jobs);
        hipLaunchKernelGGL((k_fit<true, true>), grid, blk, 0, st, jobs);
        hipLaunchKernelGGL((k_fit_finalize<true, true>), fin, blk, 0, st, jobs);
    } else {
        hipLaunchKernelGGL((k_fit<false, false>), grid, blk, 0, st, jobs);
        hipLaunchKernelGGL((k_fit_finalize<false, false>), fin, blk, 0, st, jobs);
        hipLaunchKernelGGL((k_fit<false, true>), grid, blk, 0, st, jobs);
        hipLaunchKernelGGL((k_fit_finalize<false, true>), fin, blk, 0, st, jobs);
    }
}

}  // namespace clim
