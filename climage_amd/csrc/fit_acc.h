// fit_acc.h - per-thread and per-block accumulation of the noise-model
// regression sums (internal), shared by the streaming regression kernels and
// the statistics kernels that feed the first regression pass.
#pragma once

#include "device_common.h"

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

// One DPP step of a wave reduction on a 128-bit value: the value of the lane
// that DPP control CTRL selects (zero in rows outside ROW_MASK).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ clim_fx128 fx_dpp(clim_fx128 a) {
    const uint32_t w0 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)a.lo, CTRL, ROW_MASK, 0xF, false);
    const uint32_t w1 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(a.lo >> 32), CTRL, ROW_MASK, 0xF, false);
    const uint32_t w2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)a.hi, CTRL, ROW_MASK, 0xF, false);
    const uint32_t w3 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(a.hi >> 32), CTRL, ROW_MASK, 0xF, false);
    clim_fx128 r;
    r.lo = ((uint64_t)w1 << 32) | w0;
    r.hi = ((uint64_t)w3 << 32) | w2;
    return r;
}

// The sum of a 128-bit value over the 64 lanes of a wave, in lane 63.  DPP
// row operations (quad permutations, half-row and row mirrors, row
// broadcasts) move the value between lanes inside the VALU, without an LDS
// round trip.  Integer sums do not depend on the order lanes are combined in.
__device__ __forceinline__ clim_fx128 wave_sum_dpp(clim_fx128 a) {
    a = clim_fx_add(a, fx_dpp<0xB1, 0xF>(a));   // quad_perm 1,0,3,2
    a = clim_fx_add(a, fx_dpp<0x4E, 0xF>(a));   // quad_perm 2,3,0,1
    a = clim_fx_add(a, fx_dpp<0x141, 0xF>(a));  // row_half_mirror
    a = clim_fx_add(a, fx_dpp<0x140, 0xF>(a));  // row_mirror
    a = clim_fx_add(a, fx_dpp<0x142, 0xA>(a));  // row_bcast15, rows 1 and 3
    a = clim_fx_add(a, fx_dpp<0x143, 0xC>(a));  // row_bcast31, rows 2 and 3
    return a;
}

// Block-wide sums of K fixed-point values per thread of a 256-thread block,
// four values at a time: every thread writes them to buf (1024 entries of
// LDS), then wave w folds value w's 256 entries (four per lane) and reduces
// them with DPP.  The K totals land in tot (LDS), visible to every thread on
// return.  Only one value per thread is in flight in the reduction itself.
template <int K>
__device__ __forceinline__ void block_sum(const clim_fx128* v, clim_fx128* buf, clim_fx128* tot) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 4) {
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (k0 + c < K) buf[c * 256 + tid] = v[k0 + c];
        __syncthreads();
        if (k0 + wave < K) {
            const clim_fx128* col = buf + wave * 256;
            clim_fx128 a = col[lane];
            a = clim_fx_add(a, col[lane + 64]);
            a = clim_fx_add(a, col[lane + 128]);
            a = clim_fx_add(a, col[lane + 192]);
            a = wave_sum_dpp(a);
            if (lane == 63) tot[k0 + wave] = a;
        }
        __syncthreads();
    }
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
    __device__ __forceinline__ static bool in_range(const float* m, const float* v) {
        constexpr float kMeanMin = 0x1.0624dep-10f, kVarMax = 0x1.0624dcp-10f;
        bool ok = true;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const float mm = m[SHARED_M ? 0 : c], vv = v[c];
            ok = ok && mm >= kMeanMin && mm <= 0.5f && vv >= 0.0f && vv <= kVarMax;
        }
        return ok;
    }
    __device__ __forceinline__ void add(const float* m, const float* v, bool inside) {
        add_tested(m, v, inside && in_range(m, v));
    }
    // a sample whose range test (and `inside`) gave `ok`
    __device__ __forceinline__ void add_tested(const float* m, const float* v, bool ok) {
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

// Block-sums the per-thread accumulators of a 256-thread (1-D) block and
// writes them as partial `bi` of the job: partials are stored k-major,
// partials[k * stride + bi].  buf: 1024 entries of LDS scratch (the caller's
// tile storage may serve once it is no longer needed), tot: K entries of LDS.
template <int NCH, bool SHARED_M, bool REFIT>
__device__ __forceinline__ void store_block_partials(const Acc<NCH, SHARED_M, REFIT>& acc, clim_fx128* buf,
                                                     clim_fx128* tot, clim_fx128* partials, int bi, int stride) {
    constexpr int K = Layout<NCH, SHARED_M, REFIT>::K;
    clim_fx128 sums[K];
    acc.store(sums);
    block_sum<K>(sums, buf, tot);
    if ((int)threadIdx.x < K) partials[(size_t)threadIdx.x * stride + bi] = tot[threadIdx.x];
}

// The slot of this lane's sample in the block's run of a sample list: the
// wave's lanes with `ok` take consecutive slots after the block's LDS counter
// (every lane of a full wave calls this; the slots of different waves do not
// follow a fixed order, which the exact sums do not depend on).
__device__ __forceinline__ int list_slot(bool ok, int* count) {
    const unsigned long long mask = __ballot(ok);
    const int lane = __lane_id();
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    int base = 0;
    if (lane == 0 && mask) base = atomicAdd(count, __popcll(mask));
    return __builtin_amdgcn_readfirstlane(base) + rank;
}
// the same for two samples per lane, with one counter update per wave
__device__ __forceinline__ void list_slots2(bool ok0, bool ok1, int* count, int& s0, int& s1) {
    const unsigned long long m0 = __ballot(ok0), m1 = __ballot(ok1);
    const unsigned long long below = (1ull << __lane_id()) - 1ull;
    const int c0 = __popcll(m0);
    int base = 0;
    if (__lane_id() == 0 && (m0 | m1)) base = atomicAdd(count, c0 + __popcll(m1));
    base = __builtin_amdgcn_readfirstlane(base);
    s0 = base + __popcll(m0 & below);
    s1 = base + c0 + __popcll(m1 & below);
}

// the job of block b: jobs own consecutive runs of blocks starting at block_base
template <typename Jobs>
__device__ __forceinline__ int job_of_block(const Jobs& jobs, int b) {
    int ji = 0;
    while (ji + 1 < jobs.njobs && b >= jobs.j[ji + 1].block_base) ji++;
    return ji;
}

}  // namespace clim
