// k_nlf.hip - on-device noise-level-function regression.
//
// computeRawNoiseStatistics / computeNoiseStatistics
// (cpp/pyramidal_denoise.cpp:161-256, :259-362) map the statistics image to
// the host and run three FP64 passes over it (sums, squared residuals, sums
// over inliers).  Here there are two passes in device memory:
//   first   exact sums of the grid terms (clim_q32_mean / clim_q32_var):
//           s_x, s_y, s_xx, s_xy and s_vv per channel, N, over in-range
//           samples -> A, B by least squares (B, A floored at 1e-8), and the
//           RMS residual err2 from the moments (clim_fit_rss_q), which
//           replaces the reference's residual pass.  The statistics kernels
//           accumulate these sums as they compute the statistics
//           (k_raw_stats_fit, k_noise_stats_fit).
//   refit   the same sums over inliers (residual^2 <= err2 in every
//           channel) -> A, B, streaming over the samples the statistics
//           kernels listed (those in range; k_refit).
// The grid terms are 32-bit integers, their products 64-bit, and each thread
// keeps them in 64- and 96-bit accumulators; blocks combine 128-bit sums, so
// the totals do not depend on the reduction order and equal the CPU oracle's.
// Each pass ends with a reduce kernel (one block per job and sum) that adds
// the blocks' partial sums; its last block solves every job in FP64 (one
// thread per job).  The kernel boundary makes the statistics kernels'
// partials visible (a device-scope fence per statistics block would write
// back the L2 of its XCD, full of the statistics images, every time); the
// reduce blocks' fences cover a few totals only.  One launch serves several jobs
// (all pyramid levels); each block takes an even share of one job.
#include "fit_acc.h"
#include "kernels.h"

namespace clim {

// The refit: the regression sums over the listed samples (those in range)
// that are inliers of the first fit.  Block bi of a job takes the run the
// statistics kernel's block bi listed (the same block shares), U samples per
// thread in flight (plain loads of clamped slots, masked).  RAW: the mean and
// var lists of (R, G, B, G2) quads; level: one list of (mean Y, var Y,
// var Cb, var Cr).
template <bool RAW>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_refit(FitJobs jobs) {
    constexpr int NCH = RAW ? 4 : 3;
    constexpr bool SHARED_M = !RAW;
    using L = Layout<NCH, SHARED_M, true>;
    using AccT = Acc<NCH, SHARED_M, true>;
    constexpr int K = L::K;
    constexpr int U = RAW ? 2 : 4;
    __shared__ clim_fx128 buf[1024], tot[K];

    const FitJob& jb = jobs.j[job_of_block(jobs, blockIdx.x)];
    FitState* st = jb.st;
    if (st->status == 1) return;  // no sample in range: fallback already written
    const int bi = blockIdx.x - jb.block_base;
    const int tid = threadIdx.x;

    AccT acc;
    acc.zero();
#pragma unroll
    for (int c = 0; c < NCH; c++) { acc.A[c] = st->A[c]; acc.B[c] = st->B[c]; acc.e2[c] = st->err2[c]; }
    const long long t0 = (long long)jb.ntiles * bi / jb.nblocks;
    const float4* lm = reinterpret_cast<const float4*>(jb.m.p) + (size_t)t0 * jb.seg;
    const float4* lv = reinterpret_cast<const float4*>(jb.v.p) + (size_t)t0 * jb.seg;
    const int n = (int)jb.counts[bi];
    // the next U samples' loads are issued before the current ones are added
    float4 pm[U], pv[U];
    auto fetch = [&](int i0) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = min(i0 + tid + 256 * u, n - 1);
            pm[u] = lm[i];
            if (RAW) pv[u] = lv[i];
        }
    };
    if (n > 0) fetch(0);
    for (int i0 = 0; i0 < n; i0 += 256 * U) {
        float4 cm[U], cv[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            cm[u] = pm[u];
            cv[u] = pv[u];
        }
        if (i0 + 256 * U < n) fetch(i0 + 256 * U);
#pragma unroll
        for (int u = 0; u < U; u++) {
            // every listed sample passed the range test when it was listed
            const bool mine = i0 + tid + 256 * u < n;
            if (RAW) {
                const float m[4] = {cm[u].x, cm[u].y, cm[u].z, cm[u].w};
                const float v[4] = {cv[u].x, cv[u].y, cv[u].z, cv[u].w};
                acc.add_tested(m, v, mine);
            } else {
                const float m[1] = {cm[u].x};
                const float v[3] = {cm[u].y, cm[u].z, cm[u].w};
                acc.add_tested(m, v, mine);
            }
        }
    }
    store_block_partials(acc, buf, tot, jb.partials, bi, kFitBlocksPerJob);
}

// The least-squares solve of channel c of one job from the job's totals T
// (staged in LDS) and, for the refit, the first pass's status.  Each channel
// is solved by its own thread; channel 0's thread also writes the job's
// status and sample count.
template <bool RAW, bool REFIT>
__device__ void fit_solve_channel(const FitJob& jb, const clim_fx128* T, int status, int c) {
    constexpr int NCH = RAW ? 4 : 3;
    constexpr bool SHARED_M = !RAW;
    using L = Layout<NCH, SHARED_M, REFIT>;
    FitState* st = jb.st;
    if (REFIT && status == 1) return;
    const unsigned long long cnt = T[L::n].lo;
    clim_qsums q;
    q.sx = T[L::sx(c)];
    q.sxx = T[L::sxx(c)];
    q.sy = T[L::sy(c)];
    q.sxy = T[L::sxy(c)];
    q.svv = REFIT ? clim_fx_u64(0) : T[L::svv(c)];
    float* out = st->out;
    if (!REFIT) {
        if (c == 0) {
            st->n0 = cnt;
            st->status = cnt == 0 ? 1 : 0;
        }
        if (cnt == 0) {
            // no sample in range: the model passed in, calibrated
            out[c] = jb.nlf_in[c];
            out[NCH + c] = jb.nlf_in[NCH + c];
            st->cal[c] = out[c] * jb.mult[c];
            st->cal[NCH + c] = out[NCH + c] * jb.mult[NCH + c];
        } else {
            double A, B;
            clim_fit_solve_q(&q, cnt, &A, &B);
            st->A[c] = A;
            st->B[c] = B;
            st->err2[c] = sqrt(clim_fit_rss_q(&q, cnt, A, B) / (double)cnt);
        }
    } else {
        // no inliers: the first pass's model stays
        double A = st->A[c], B = st->B[c];
        if (cnt > 0) {
            clim_fit_solve_q(&q, cnt, &A, &B);
            st->A[c] = A;
            st->B[c] = B;
        }
        if (c == 0) st->status = cnt > 0 ? 0 : 2;
        out[c] = (float)A;
        out[NCH + c] = (float)B;
        st->cal[c] = out[c] * jb.mult[c];
        st->cal[NCH + c] = out[NCH + c] * jb.mult[NCH + c];
    }
}

// One block per (job, sum): adds the job's block partials of that sum (eight
// per thread, plain loads of clamped indexes, masked, all in flight at once)
// into the job's FitState::tot.  The last block of the launch to finish
// (a device-scope counter in the first job's FitState, released and acquired
// around the atomic increment, reset by that block) then solves every job,
// one thread per job.
template <bool RAW, bool REFIT>
__global__ void __launch_bounds__(256) k_fit_reduce(FitJobs jobs) {
    constexpr int K = Layout<RAW ? 4 : 3, !RAW, REFIT>::K;
    constexpr int PT = kFitBlocksPerJob / 256;  // partials per thread
    static_assert(kFitBlocksPerJob == PT * 256, "whole partials per thread");
    static_assert(K <= kFitSums, "FitState::tot holds every sum");
    __shared__ clim_fx128 wsum[4];
    __shared__ int last;
    const int j = blockIdx.x / K, k = blockIdx.x % K;
    const FitJob& jb = jobs.j[j];
    FitState* st = jb.st;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // no sample in range: the fallback is already written, nothing to add
    if (!(REFIT && st->status == 1)) {
        const clim_fx128* pk = jb.partials + (size_t)k * kFitBlocksPerJob;
        clim_fx128 v[PT];
#pragma unroll
        for (int i = 0; i < PT; i++) v[i] = pk[min(tid + 256 * i, jb.nblocks - 1)];
        clim_fx128 s = {0, 0};
#pragma unroll
        for (int i = 0; i < PT; i++) {
            const bool mine = tid + 256 * i < jb.nblocks;
            v[i].hi = mine ? v[i].hi : 0;
            v[i].lo = mine ? v[i].lo : 0;
            s = clim_fx_add(s, v[i]);
        }
        s = wave_sum_dpp(s);
        if (lane == 63) wsum[wave] = s;
        __syncthreads();
        if (tid == 0) st->tot[k] = clim_fx_add(clim_fx_add(wsum[0], wsum[1]), clim_fx_add(wsum[2], wsum[3]));
    }
    if (tid == 0) {
        __threadfence();  // this block's total before the count
        unsigned int* done = &jobs.j[0].st->reduced;
        last = atomicAdd(done, 1u) == (unsigned int)(jobs.njobs * K) - 1u;
        if (last) *done = 0u;  // every other block has counted
    }
    __syncthreads();
    if (!last) return;
    __threadfence();  // every block's total after the count
    // every job's totals (and first-pass status) in LDS with one round of
    // loads, then one thread per (job, channel)
    constexpr int NCH = RAW ? 4 : 3;
    __shared__ clim_fx128 T[5][K];
    __shared__ int status[5];
    for (int i = tid; i < jobs.njobs * K; i += 256) T[i / K][i % K] = jobs.j[i / K].st->tot[i % K];
    if (REFIT && tid < jobs.njobs) status[tid] = jobs.j[tid].st->status;
    __syncthreads();
    if (tid < jobs.njobs * NCH)
        fit_solve_channel<RAW, REFIT>(jobs.j[tid / NCH], T[tid / NCH], REFIT ? status[tid / NCH] : 0, tid % NCH);
}

// the combination of a pass's block partials and the solve
template <bool RAW, bool REFIT>
static void launch_finalize(const FitJobs& jobs, hipStream_t st) {
    constexpr int K = Layout<RAW ? 4 : 3, !RAW, REFIT>::K;
    hipLaunchKernelGGL((k_fit_reduce<RAW, REFIT>), dim3(jobs.njobs * K), dim3(256), 0, st, jobs);
}

int fit_share_blocks(FitJobs& jobs, const long long* units, int grid) {
    long long total = 0;
    for (int j = 0; j < jobs.njobs; j++) total += units[j];
    int base = 0;
    for (int j = 0; j < jobs.njobs; j++) {
        long long nb = total > 0 ? (units[j] * grid + total / 2) / total : 1;
        if (nb > units[j]) nb = units[j];
        if (nb > kFitBlocksPerJob) nb = kFitBlocksPerJob;
        if (nb < 1) nb = 1;
        jobs.j[j].nblocks = (int)nb;
        jobs.j[j].block_base = base;
        base += (int)nb;
    }
    return base;
}

// The finalisation of the first pass (whose partials the statistics kernels
// wrote with the jobs' block shares), the refit over the sample lists with
// the same shares (block_base renumbered for one launch over every job), and
// its finalisation.
void launch_fit_rest(FitJobs& jobs, bool raw, hipStream_t st) {
    const dim3 blk(256);
    if (raw)
        launch_finalize<true, false>(jobs, st);
    else
        launch_finalize<false, false>(jobs, st);
    int base = 0;
    for (int j = 0; j < jobs.njobs; j++) {
        jobs.j[j].block_base = base;
        base += jobs.j[j].nblocks;
    }
    if (raw) {
        hipLaunchKernelGGL(k_refit<true>, dim3(base), blk, 0, st, jobs);
        launch_finalize<true, true>(jobs, st);
    } else {
        hipLaunchKernelGGL(k_refit<false>, dim3(base), blk, 0, st, jobs);
        launch_finalize<false, true>(jobs, st);
    }
}

}  // namespace clim
