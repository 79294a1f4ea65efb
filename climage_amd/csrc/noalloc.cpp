// noalloc.cpp - clim_compute_raw_noise_statistics_noalloc: the raw noise
// statistics and regression of clim_compute_raw_noise_statistics with every
// work buffer in one static device array of the library, for callers that
// cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// The array `scratch` holds, from offset 0 (each part 256-byte aligned):
//   FitState            the fit's totals, model and reduction counter
//   partials            kFitBlocksPerJob x kFitK 128-bit block sums
//   counts              kFitBlocksPerJob samples listed per statistics block
//   mean list, var list kNoallocMaxSlots float4 slots each
// In host code `scratch` names a host-side stand-in, not device memory:
// clim_noalloc_init (called by clim_create with the context's device current)
// takes the array's device address with hipGetSymbolAddress once per device,
// and only that address is passed to kernels.  Calls on one device share the
// array and it keeps the last call's contents; the reduction counter is
// cleared by a fill kernel at the start of each call.  Two calls that overlap
// on one GPU (two streams or two contexts) race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "ctx.h"

namespace {

constexpr size_t align256(size_t v) { return (v + 255) / 256 * 256; }

constexpr size_t kOffFit = 0;
constexpr size_t kOffPartials = align256(kOffFit + sizeof(clim::FitState));
constexpr size_t kOffCounts = align256(kOffPartials + sizeof(clim_fx128) * clim::kFitBlocksPerJob * kFitK);
constexpr size_t kOffMean = align256(kOffCounts + sizeof(unsigned int) * clim::kFitBlocksPerJob);
constexpr size_t kListBytes = (size_t)CLIM_NOALLOC_MAX_QUADS * 16;
constexpr size_t kOffVar = kOffMean + kListBytes;
constexpr size_t kScratchBytes = kOffVar + kListBytes;
static_assert(kListBytes % 256 == 0, "lists keep 256-byte alignment");
static_assert(kScratchBytes <= CLIM_NOALLOC_SCRATCH_BYTES, "the documented size covers the layout");

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

static __device__ __align__(256) unsigned char scratch[CLIM_NOALLOC_SCRATCH_BYTES];

// zeroes a FitState (its reduction counter must start at 0) with vector stores
__global__ void k_noalloc_clear(unsigned int* p, int words) {
    for (int i = threadIdx.x; i < words; i += blockDim.x) p[i] = 0u;
}

// the fitted model (A[4], B[4]) and the status into the caller's device memory
__global__ void k_noalloc_out(const clim::FitState* st, float* nlf, int32_t* status) {
    const int t = threadIdx.x;
    if (t < 8) nlf[t] = st->out[t];
    if (t == 0 && status) *status = st->status;
}

// Called by clim_create with `device` current: the array's device address on
// that device, taken once per device and kept for the process.
void clim_noalloc_init(int device, void** ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_done[device]) {
        void* p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? p : nullptr;
        g_scratch_done[device] = true;
    }
    *ptr = g_scratch_ptr[device];
}

extern "C" clim_status clim_compute_raw_noise_statistics_noalloc(clim_ctx* ctx, const clim_image* scaled,
                                                                 int32_t bayer, float* nlf_dev,
                                                                 int32_t* status_dev) {
    if (!ctx || !nlf_dev) return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics (no-alloc): null argument");
    if (bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics (no-alloc): bad bayer");
    if (!scaled || !scaled->data || scaled->format != CLIM_FMT_R32F || scaled->width <= 0 || scaled->height <= 0 ||
        scaled->pitch_bytes < scaled->width * 4)
        return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics (no-alloc): scaled must be an R32F image");
    if ((scaled->width & 1) || (scaled->height & 1))
        return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics (no-alloc): odd mosaic size");
    int aw = 0, ah = 0;
    clim::fit_list_dims(true, scaled->width, scaled->height, &aw, &ah);
    if ((long long)aw * ah > CLIM_NOALLOC_MAX_QUADS)
        return ctx_fail(ctx, CLIM_EINVAL,
                        "computeRawNoiseStatistics (no-alloc): mosaic of " + std::to_string(scaled->width) + "x" +
                            std::to_string(scaled->height) + " needs " + std::to_string((long long)aw * ah) +
                            " sample slots, more than the static scratch's " +
                            std::to_string(CLIM_NOALLOC_MAX_QUADS));
    if (!ctx->noalloc_scratch)
        return ctx_fail(ctx, CLIM_EHIP, "computeRawNoiseStatistics (no-alloc): no device address for the static "
                                        "scratch array on this device");
    char* base = static_cast<char*>(ctx->noalloc_scratch);
    clim::FitState* st = reinterpret_cast<clim::FitState*>(base + kOffFit);

    clim::FitJobs jobs{};
    jobs.njobs = 1;
    clim::FitJob& jb = jobs.j[0];
    jb.src = clim::Img{(char*)scaled->data, scaled->width, scaled->height, scaled->pitch_bytes};
    jb.b = make_bayer(bayer);
    jb.m = clim::Img{base + kOffMean, aw, ah, aw * 16};
    jb.v = clim::Img{base + kOffVar, aw, ah, aw * 16};
    jb.cap = (long long)aw * ah;
    jb.st = st;
    jb.partials = reinterpret_cast<clim_fx128*>(base + kOffPartials);
    jb.counts = reinterpret_cast<unsigned int*>(base + kOffCounts);
    for (int k = 0; k < 8; k++) {
        jb.nlf_in[k] = 0.0f;
        jb.mult[k] = 1.0f;
    }

    static_assert(sizeof(clim::FitState) % 4 == 0, "whole words");
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(256), 0, ctx->stream, reinterpret_cast<unsigned int*>(st),
                       (int)(sizeof(clim::FitState) / 4));
    CLIM_TRY_HIP(ctx, hipGetLastError());
    if (!clim::launch_raw_stats_fit(jobs, nullptr, 0.0f, ctx->stream))
        return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics (no-alloc): sample list smaller than the image");
    CLIM_TRY_HIP(ctx, hipGetLastError());
    clim::launch_fit_rest(jobs, true, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    hipLaunchKernelGGL(k_noalloc_out, dim3(1), dim3(64), 0, ctx->stream, st, nlf_dev, status_dev);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}
