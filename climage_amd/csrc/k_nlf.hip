// k_nlf.hip - on-device noise-level-function regression.
//
// computeRawNoiseStatistics / computeNoiseStatistics
// (cpp/pyramidal_denoise.cpp:161-256, :259-362) map the statistics image to
// the host and run three FP64 passes over it.  Here the three passes are three
// launches over device memory, with no host round trip:
//   pass 0  sums (s_x, s_y, s_xx, s_xy, N) over in-range samples
//           -> A, B = least squares of var = A + B*mean
//   pass 1  sum of squared residuals -> err2 = sqrt(sum / N)
//   pass 2  sums over inliers (residual^2 <= err2 in every channel) -> A, B
// Terms are accumulated in 128-bit fixed point (2^-90 resolution), so the
// sums are exact for a given set of terms and independent of the reduction
// order.  The last block of each job to finish combines the per-block
// partials and solves in FP64.  Several jobs (pyramid levels) run in one
// launch, one grid row per job.
#include "kernels.h"

namespace clim {

template <int NCH>
__device__ __forceinline__ bool fit_load(const FitJob& jb, long long i, double* mm, double* vv) {
    const int x = (int)(i % jb.w), y = (int)(i / jb.w);
    const float* m = reinterpret_cast<const float*>(jb.m + (size_t)y * jb.pitch_m) + x * 4 + jb.m_elem;
    const float* v = reinterpret_cast<const float*>(jb.v + (size_t)y * jb.pitch_v) + x * 4 + jb.v_elem;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        mm[c] = (double)m[jb.m_nch == 1 ? 0 : c];
        vv[c] = (double)v[c];
        if (mm[c] < 0.001 || mm[c] > 0.5) ok = false;
        if (vv[c] < 0 || vv[c] > 0.001) ok = false;
    }
    return ok;
}

__device__ __forceinline__ clim_fx128 wave_sum(clim_fx128 a) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) a = clim_fx_add(a, fx_shfl_xor(a, m));
    return a;
}

// block-wide sum of K fixed-point values; result valid in thread 0's out[]
template <int K>
__device__ __forceinline__ void block_sum(clim_fx128* v, clim_fx128 (*lds)[K]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        clim_fx128 s = wave_sum(v[k]);
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

template <int NCH>
__device__ void fit_solve(const clim_fx128* tot, unsigned long long n, double* A, double* B) {
    const double N = (double)n;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        const double s_x = clim_fx_to_double(tot[4 * c + 0]);
        const double s_y = clim_fx_to_double(tot[4 * c + 1]);
        const double s_xx = clim_fx_to_double(tot[4 * c + 2]);
        const double s_xy = clim_fx_to_double(tot[4 * c + 3]);
        B[c] = fmax((N * s_xy - s_x * s_y) / (N * s_xx - s_x * s_x), 1e-8);
        A[c] = fmax((s_y - B[c] * s_x) / N, 1e-8);
    }
}

template <int NCH>
__global__ void __launch_bounds__(256) k_fit(FitJobs jobs, int pass) {
    constexpr int K = 4 * NCH + 1;  // 4 sums per channel + count
    __shared__ clim_fx128 lds[4][K];
    __shared__ unsigned int s_last;
    const FitJob& jb = jobs.j[blockIdx.y];
    FitState* st = jb.st;
    if (pass > 0 && st->status == 1) return;  // no sample in range: fallback already written

    clim_fx128 acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k].hi = acc[k].lo = 0;
    double A[NCH], B[NCH], e2[NCH];
    if (pass > 0) {
#pragma unroll
        for (int c = 0; c < NCH; c++) { A[c] = st->A[c]; B[c] = st->B[c]; e2[c] = st->err2[c]; }
    }
    const long long count = (long long)jb.w * jb.h;
    unsigned long long n = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < count; i += (long long)gridDim.x * 256) {
        double mm[NCH], vv[NCH];
        if (!fit_load<NCH>(jb, i, mm, vv)) continue;
        if (pass == 1) {
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const double diff = (A[c] + B[c] * mm[c]) - vv[c];
                acc[c] = clim_fx_add(acc[c], clim_fx_from_double(diff * diff));
            }
            continue;
        }
        if (pass == 2) {
            bool inlier = true;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                const double diff = (A[c] + B[c] * mm[c]) - vv[c];
                if (!(diff * diff <= e2[c])) inlier = false;
            }
            if (!inlier) continue;
        }
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            acc[4 * c + 0] = clim_fx_add(acc[4 * c + 0], clim_fx_from_double(mm[c]));
            acc[4 * c + 1] = clim_fx_add(acc[4 * c + 1], clim_fx_from_double(vv[c]));
            acc[4 * c + 2] = clim_fx_add(acc[4 * c + 2], clim_fx_from_double(mm[c] * mm[c]));
            acc[4 * c + 3] = clim_fx_add(acc[4 * c + 3], clim_fx_from_double(mm[c] * vv[c]));
        }
        n++;
    }
    acc[K - 1].lo = n;

    block_sum<K>(acc, lds);
    clim_fx128* part = jb.partials + (size_t)blockIdx.x * K;
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; k++) part[k] = acc[k];
        __threadfence();
        const unsigned int prev = atomicAdd(&st->counter, 1u);
        s_last = prev == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();

    // last block: combine the partials of every block
    clim_fx128 tot[K];
#pragma unroll
    for (int k = 0; k < K; k++) tot[k].hi = tot[k].lo = 0;
    for (unsigned int b = threadIdx.x; b < gridDim.x; b += 256) {
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
    if (threadIdx.x != 0) return;

    const unsigned long long cnt = tot[K - 1].lo;
    float* out = st->out;
    if (pass == 0) {
        st->n0 = cnt;
        if (cnt == 0) {
            st->status = 1;
            for (int k = 0; k < 2 * NCH; k++) out[k] = jb.nlf_in[k];
        } else {
            st->status = 0;
            fit_solve<NCH>(tot, cnt, st->A, st->B);
        }
    } else if (pass == 1) {
        for (int c = 0; c < NCH; c++) st->err2[c] = sqrt(clim_fx_to_double(tot[c]) / (double)st->n0);
    } else {
        if (cnt > 0) {
            fit_solve<NCH>(tot, cnt, st->A, st->B);
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

// nblocks: blocks per job (the grid row length); partials of each job must hold
// nblocks * (4*nch + 1) fixed-point values.
void launch_fit(const FitJobs& jobs, int njobs, int nch, int nblocks, int pass, hipStream_t st) {
    dim3 grd(nblocks, njobs), blk(256);
    if (nch == 4)
        hipLaunchKernelGGL(k_fit<4>, grd, blk, 0, st, jobs, pass);
    else
        hipLaunchKernelGGL(k_fit<3>, grd, blk, 0, st, jobs, pass);
}

}  // namespace clim
