// ctx.h - the library context behind clim_ctx (internal).
#pragma once

#include <string>
#include <vector>

#include "climage.h"
#include "kernels.h"
#include "strength.h"

// streams a clim_raw_to_rgb_batch call spreads its frames over (on batches
// of 72 12 MP frames three measured 1.4 % faster than two, four 3 % slower)
#ifndef CLIM_BATCH_LANES
#define CLIM_BATCH_LANES 3
#endif

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// device image owned by the context
// A device image the context owns.  The allocation only grows: img is a view
// of the requested size on an allocation of cap_w x cap_h texels (the pitch
// of cap_w), so frames of different sizes reuse it.
struct OwnedImg {
    clim::Img img{nullptr, 0, 0, 0};
    int format = 0;
    int cap_w = 0, cap_h = 0;
};

// work buffers of one frame geometry (RawConverter::allocateTextures,
// cpp/raw_converter.cpp:24-40, plus the pyramid of cpp/pyramidal_denoise.cpp:69-93)
struct Workspace {
    int w = 0, h = 0, levels = 0, algorithm = -1;
    OwnedImg scaled, green, rgbA, rgbB, raw_mean, raw_var;
    OwnedImg pyr[CLIM_MAX_LEVELS];    // pyr[i], i >= 1: downsampled level i
    OwnedImg den[CLIM_MAX_LEVELS];    // den[i], i >= 1: denoised level i (den[0] is rgbA)
    OwnedImg stats[CLIM_MAX_LEVELS];  // noiseStatistics of level i
};

constexpr int kFitJobs = 1 + CLIM_MAX_LEVELS;  // raw + levels
constexpr int kFitK = 21;                       // largest accumulator count (raw first pass: 5 sums x 4 channels + count)

enum Stage {
    ST_SCALE = 0, ST_RAW_STATS, ST_RAW_FIT, ST_GREEN, ST_RED_BLUE, ST_DESPECKLE,
    ST_DOWNSAMPLE, ST_LEVEL_STATS, ST_LEVEL_FIT, ST_DENOISE, ST_REASSEMBLE, ST_TONE, ST_RAW_DENOISE, ST_STRENGTH
};

// learned denoise strength (strength.h, strength.cpp) on one context
struct SnTrainer;
struct SnState {
    int mode = 0;                       // clim_strength_mode
    clim::SnWeights wt{};               // device weights loaded on this context (null until set)
    const clim::SnWeights* use = nullptr;  // the weights inference reads (a batch lane: its owner's)
    DevBuf map, gate_sum, open;         // per-frame buffers (grid of level 3)
    int gw = 0, gh = 0, levels = 0;
    SnTrainer* trainer = nullptr;       // training / gradient buffers, created on first use
};

struct ProfRec {
    int stage;
    long long texels;
    hipEvent_t a, b;
};

struct clim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    Workspace ws;
    clim::FitState* fit = nullptr;            // device [kFitJobs]
    clim_fx128* partials = nullptr;           // device [kFitJobs][kFitBlocksPerJob][kFitK]
    unsigned int* fit_counts = nullptr;       // device [kFitJobs][kFitBlocksPerJob] samples listed per block
    clim::FitState* fit_host = nullptr;       // pinned mirror for read-back
    clim::SceneParams* scene = nullptr;       // device
    clim::SceneParams* scene_host = nullptr;  // pinned staging
    // scratch of the stand-alone guided filter
    clim::GuidedScratch op_guided{};
    OwnedImg op_guided_imgs[8];
    // sample lists of the stand-alone regression operators
    OwnedImg op_mean, op_var, op_stats;
    // clim_raw_to_rgb_batch: the further lanes (contexts of their own on the
    // same device: stream, work buffers, fit slots; frame i runs on lane
    // i % kBatchLanes, lane 0 being this context), the events that order the
    // lanes' streams (ev[0]: the batch's start on this stream, ev[k]: lane k's
    // end), and device copies of the frames' fit states
    static constexpr int kBatchLanes = CLIM_BATCH_LANES;
    clim_ctx* lanes[kBatchLanes - 1] = {};
    hipEvent_t lane_ev[kBatchLanes] = {};
    clim::FitState* fit_hist = nullptr;
    size_t fit_hist_n = 0;
    // stage profiling
    bool prof = false;
    uint32_t prof_mask = 0xffffffffu;  // stages that record events
    std::vector<hipEvent_t> ev_pool;
    std::vector<ProfRec> prof_recs;
    size_t ev_next = 0;
    SnState sn;
    int precision = 0;  // clim_precision
    // device address of the static scratch array of
    // clim_compute_raw_noise_statistics_noalloc on this context's device
    void* noalloc_scratch = nullptr;
};

// brackets the launches issued in its scope with two events when profiling
struct ProfScope {
    clim_ctx* ctx;
    int idx = -1;
    ProfScope(clim_ctx* c, int stage, long long texels);
    ~ProfScope();
};

// helpers shared by the translation units of the library
clim_status ctx_fail(clim_ctx* ctx, clim_status s, const std::string& msg);
clim_status ctx_hip(clim_ctx* ctx, hipError_t e, const char* what);
clim_status owned_alloc(clim_ctx* ctx, OwnedImg& o, int w, int h, int format);
void owned_free(OwnedImg& o);
clim::Bayer make_bayer(int bayer);
int scene_colors(int bayer);
// the static scratch array's device address on `device` (current), taken once
// per device (noalloc.cpp); null when hipGetSymbolAddress failed
void clim_noalloc_init(int device, void** ptr);
// learned strength (strength.cpp): the network on level 3 of a frame's
// pyramid after the per-level fits, filling sr for the level filters (off
// unless a mode and weights are set); the context's strength buffers freed
clim_status strength_infer(clim_ctx* ctx, const clim::Img* lvl, int levels, int algorithm, clim::StrengthRef* sr);
void sn_release(clim_ctx* ctx);

#define CLIM_TRY_HIP(ctx, call)                                  \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return ctx_hip((ctx), e_, #call);  \
    } while (0)

#define CLIM_TRY(call)                      \
    do {                                    \
        clim_status s_ = (call);            \
        if (s_ != CLIM_OK) return s_;       \
    } while (0)
