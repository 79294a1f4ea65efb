// burst.cpp - a burst of frames recorded once as one HIP graph and replayed
// (clim_burst_*).
//
// Each frame is RawConverter::demosaicImage (cpp/raw_converter.cpp:73-186),
// as clim_raw_to_rgb computes it; the results are those of the single-frame
// calls bit for bit.  The burst differs only in how the frames' work is laid
// out on the device:
//  - every frame keeps its own green plane and despeckled YCbCr image, so
//    the full-resolution stencil stages (interpolateGreen, and
//    interpolateRedBlue + highlights + YCbCr + despeckle) of all frames run
//    as two launches of persistent blocks over one cross-frame tile list;
//  - the raw noise fits before them and the pyramids after them run per
//    frame, spread over lanes (contexts of the burst's own with a stream and
//    work buffers each), so one frame's small launches overlap another's;
//  - the whole sequence is captured once into a HIP graph: a replay is one
//    hipGraphLaunch on the caller's context stream, with no allocation and
//    no host work per frame;
//  - the persistent launches' shape (register budget as waves per SIMD,
//    blocks per CU, tiles per take) is chosen at creation by timing every
//    candidate on the burst's own frames; then the burst is recorded twice,
//    with the cross-frame launches and with every frame's demosaic on its
//    lane (the per-frame kernels), both graphs are timed, and the faster is
//    kept; the choices are kept per process for bursts of the same frame
//    sizes.
// The burst owns every buffer its graph touches apart from the caller's
// images, which must outlive it; the learned-strength weights and mode of
// the context are copied in at creation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "pipeline.h"

namespace {

constexpr int kLanes = 3;  // frame i runs on lane i mod 3

struct Tuned {
    int waves[2], blocks_per_cu[2], grab[2];
    float us[2], us_first[2];
    int candidates[2];
    int form;          // 0: cross-frame stencil launches, 1: per-frame kernels
    float form_ms[2];  // a replay of each form at creation
    int form_known;    // the form was chosen by timing (a cached choice can be reused)
};  // a clim_burst_tuning without the per-burst fields
std::mutex g_tune_mu;
std::map<std::string, Tuned> g_tune;  // frame-size signature -> choice

}  // namespace

struct clim_burst {
    clim_ctx* ctx = nullptr;  // the caller's context (errors, the replay stream)
    int n = 0;
    clim_ctx* lane[kLanes] = {};
    std::vector<clim_demosaic_params> params;  // FrameCall::p points here
    std::vector<FrameCall> calls;
    std::vector<OwnedImg> green, ycc;  // per frame
    std::vector<int> cross;            // frames in the cross-frame launches
    clim::FitState* fits = nullptr;    // [n] the frames' raw fits
    clim::BurstFrame* table = nullptr; // [cross.size()]
    clim::BurstQueue* queues = nullptr;  // [2]
    int total[2] = {0, 0};
    int cus = 0;
    bool half = false;  // FP16 YCbCr images (the mixed-precision form)
    Tuned tune{};
    bool cached = false;
    hipEvent_t fork = nullptr, mid = nullptr, lane_end[kLanes] = {}, lane_mid[kLanes] = {};
    bool timed = false;
    hipEvent_t t_start = nullptr;
    std::vector<hipEvent_t> t_frame;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static clim_status burst_fail(clim_burst* b, clim_status s, const std::string& msg) {
    return ctx_fail(b->ctx, s, "burst: " + msg);
}

#define BURST_TRY_HIP(b, call)                                               \
    do {                                                                     \
        hipError_t e_ = (call);                                              \
        if (e_ != hipSuccess) return ctx_hip((b)->ctx, e_, "burst: " #call); \
    } while (0)

// lane of frame i, and its context
static int lane_of(int i) { return i % kLanes; }

// the graph's work: raw fits (lanes), the cross-frame stencil launches (lane
// 0), the pyramids (lanes); lane 0's stream is the origin.  perframe: every
// frame's demosaic runs on its lane after its fit (no cross-frame launches,
// no join between the stages)
static clim_status record(clim_burst* b, bool with_times, bool perframe) {
    hipStream_t s0 = b->lane[0]->stream;
    auto lane_error = [&](clim_ctx* c, clim_status st, int i) {
        return burst_fail(b, st, "frame " + std::to_string(i) + ": " + c->last_error);
    };
    if (with_times) BURST_TRY_HIP(b, hipEventRecordWithFlags(b->t_start, s0, hipEventRecordExternal));
    BURST_TRY_HIP(b, hipEventRecord(b->fork, s0));
    for (int k = 1; k < kLanes; k++) BURST_TRY_HIP(b, hipStreamWaitEvent(b->lane[k]->stream, b->fork, 0));
    for (int i = 0; i < b->n; i++) {
        clim_ctx* c = b->lane[lane_of(i)];
        FrameCall& f = b->calls[i];
        clim_status st = frame_workspace(c, f);
        if (st == CLIM_OK) st = frame_raw_fit(c, f, b->fits + i);
        if (st == CLIM_OK && (perframe || std::find(b->cross.begin(), b->cross.end(), i) == b->cross.end()))
            st = frame_demosaic(c, f, b->green[i].img, b->ycc[i].img, 0);
        if (st == CLIM_OK && perframe) st = frame_pyramid(c, f, b->ycc[i].img);
        if (st != CLIM_OK) return lane_error(c, st, i);
        if (perframe && with_times)
            BURST_TRY_HIP(b, hipEventRecordWithFlags(b->t_frame[i], c->stream, hipEventRecordExternal));
    }
    if (perframe) {
        for (int k = 1; k < kLanes; k++) {
            BURST_TRY_HIP(b, hipEventRecord(b->lane_end[k], b->lane[k]->stream));
            BURST_TRY_HIP(b, hipStreamWaitEvent(s0, b->lane_end[k], 0));
        }
        return CLIM_OK;
    }
    for (int k = 1; k < kLanes; k++) {
        BURST_TRY_HIP(b, hipEventRecord(b->lane_mid[k], b->lane[k]->stream));
        BURST_TRY_HIP(b, hipStreamWaitEvent(s0, b->lane_mid[k], 0));
    }
    if (!b->cross.empty()) {
        for (int kind = 0; kind < 2; kind++)
            clim::launch_demosaic_burst(kind, b->tune.waves[kind], b->half, b->table, b->total[kind],
                                        b->queues + kind, b->tune.blocks_per_cu[kind] * b->cus, b->tune.grab[kind],
                                        s0);
        BURST_TRY_HIP(b, hipGetLastError());
    }
    BURST_TRY_HIP(b, hipEventRecord(b->mid, s0));
    for (int k = 1; k < kLanes; k++) BURST_TRY_HIP(b, hipStreamWaitEvent(b->lane[k]->stream, b->mid, 0));
    for (int i = 0; i < b->n; i++) {
        clim_ctx* c = b->lane[lane_of(i)];
        const FrameCall& f = b->calls[i];
        clim_status st = frame_workspace(c, f);
        if (st == CLIM_OK) st = frame_pyramid(c, f, b->ycc[i].img);
        if (st != CLIM_OK) return lane_error(c, st, i);
        if (with_times) BURST_TRY_HIP(b, hipEventRecordWithFlags(b->t_frame[i], c->stream, hipEventRecordExternal));
    }
    for (int k = 1; k < kLanes; k++) {
        BURST_TRY_HIP(b, hipEventRecord(b->lane_end[k], b->lane[k]->stream));
        BURST_TRY_HIP(b, hipStreamWaitEvent(s0, b->lane_end[k], 0));
    }
    return CLIM_OK;
}

// time one cross-frame launch shape (us per launch, lane 0's stream)
static clim_status time_shape(clim_burst* b, int kind, int waves, int bpc, int grab, float* us) {
    hipStream_t s0 = b->lane[0]->stream;
    hipEvent_t e0, e1;
    BURST_TRY_HIP(b, hipEventCreate(&e0));
    BURST_TRY_HIP(b, hipEventCreate(&e1));
    constexpr int kReps = 2;
    hipError_t e = hipSuccess;
    clim::launch_demosaic_burst(kind, waves, b->half, b->table, b->total[kind], b->queues + kind, bpc * b->cus, grab,
                                s0);
    if (e == hipSuccess) e = hipEventRecord(e0, s0);
    for (int r = 0; r < kReps && e == hipSuccess; r++)
        clim::launch_demosaic_burst(kind, waves, b->half, b->table, b->total[kind], b->queues + kind, bpc * b->cus,
                                    grab, s0);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, s0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return ctx_hip(b->ctx, e, "burst: autotune timing");
    *us = ms * 1e3f / kReps;
    return CLIM_OK;
}

// Candidates of one kind: register budget (waves per SIMD the build targets),
// persistent blocks per CU (at most what the budget and the kernel's LDS let
// reside), tiles per take.  The first candidate is the plain shape.
static clim_status autotune(clim_burst* b, int kind) {
    struct Shape { int waves, bpc, grab; };
    std::vector<Shape> cand;
    const int waves_set[2][2] = {{8, 6}, {4, 4}};
    const int max_bpc[2][2] = {{8, 6}, {4, 4}};  // resident 256-thread blocks per CU (registers, LDS)
    for (int wi = 0; wi < (kind == 0 ? 2 : 1); wi++)
        for (int bpc : {1, 2, 4, 6, 8})
            for (int grab : {1, 2, 4})
                if (bpc <= max_bpc[kind][wi]) cand.push_back({waves_set[kind][wi], bpc, grab});
    float best = 1e30f, first = 0.0f;
    Shape pick = cand[0];
    for (size_t c = 0; c < cand.size(); c++) {
        float us = 0.0f;
        CLIM_TRY(time_shape(b, kind, cand[c].waves, cand[c].bpc, cand[c].grab, &us));
        if (c == 0) first = us;
        if (us < best) {
            best = us;
            pick = cand[c];
        }
    }
    b->tune.waves[kind] = pick.waves;
    b->tune.blocks_per_cu[kind] = pick.bpc;
    b->tune.grab[kind] = pick.grab;
    b->tune.us[kind] = best;
    b->tune.us_first[kind] = first;
    b->tune.candidates[kind] = (int)cand.size();
    return CLIM_OK;
}

extern "C" void clim_burst_destroy(clim_burst* b) {
    if (!b) return;
    if (b->lane[0]) (void)hipStreamSynchronize(b->lane[0]->stream);
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& o : b->green) owned_free(o);
    for (auto& o : b->ycc) owned_free(o);
    if (b->fits) (void)hipFree(b->fits);
    if (b->table) (void)hipFree(b->table);
    if (b->queues) (void)hipFree(b->queues);
    hipEvent_t evs[] = {b->fork, b->mid, b->t_start};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < kLanes; k++) {
        if (b->lane_end[k]) (void)hipEventDestroy(b->lane_end[k]);
        if (b->lane_mid[k]) (void)hipEventDestroy(b->lane_mid[k]);
    }
    for (hipEvent_t e : b->t_frame)
        if (e) (void)hipEventDestroy(e);
    for (int k = 0; k < kLanes; k++)
        if (b->lane[k]) clim_destroy(b->lane[k]);
    delete b;
}

static clim_status burst_build(clim_burst* b, int32_t n, const clim_image* const* raws,
                               clim_image* const* out_rgba8, clim_image* const* out_f32,
                               const clim_demosaic_params* params, uint32_t flags) {
    clim_ctx* ctx = b->ctx;
    b->n = n;
    b->params.assign(params, params + n);
    b->calls.resize(n);
    b->green.resize(n);
    b->ycc.resize(n);
    b->timed = (flags & CLIM_BURST_FRAME_TIMES) != 0;
    b->half = ctx->precision == CLIM_PRECISION_MIXED;
    BURST_TRY_HIP(b, hipSetDevice(ctx->device));
    BURST_TRY_HIP(b, hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    // lanes: contexts of the burst's own, with the context's learned strength
    std::vector<float> weights;
    if (ctx->sn.wt.master) {
        weights.resize(CLIM_STRENGTH_PARAMS);
        CLIM_TRY(clim_strength_get_weights(ctx, weights.data(), (int64_t)weights.size()));
    }
    for (int k = 0; k < kLanes; k++) {
        const clim_status s = clim_create(&b->lane[k], ctx->device);
        if (s != CLIM_OK) return burst_fail(b, s, "lane context");
        b->lane[k]->precision = ctx->precision;
        if (!weights.empty()) {
            if (clim_strength_set_weights(b->lane[k], weights.data(), (int64_t)weights.size()) != CLIM_OK ||
                clim_strength_set_mode(b->lane[k], ctx->sn.mode) != CLIM_OK)
                return burst_fail(b, CLIM_EINVAL, "lane strength: " + b->lane[k]->last_error);
        }
    }
    for (int i = 0; i < n; i++) {
        clim_ctx* c = b->lane[lane_of(i)];
        const clim_status s = frame_prepare(c, raws[i], out_rgba8[i], out_f32 ? out_f32[i] : nullptr,
                                            &b->params[i], &b->calls[i]);
        if (s != CLIM_OK) return burst_fail(b, s, "frame " + std::to_string(i) + ": " + c->last_error);
        const FrameCall& f = b->calls[i];
        CLIM_TRY(owned_alloc(ctx, b->green[i], f.w, f.h, CLIM_FMT_R32F));
        CLIM_TRY(owned_alloc(ctx, b->ycc[i], f.w, f.h,
                             ctx->precision == CLIM_PRECISION_MIXED ? CLIM_FMT_RGBA16F : CLIM_FMT_RGBA32F));
        // the high-noise frames read their pre-denoised plane from the lane's
        // work buffers: their demosaic stays with the lane
        if (!f.high_noise) b->cross.push_back(i);
    }
    BURST_TRY_HIP(b, hipMalloc(&b->fits, sizeof(clim::FitState) * n));
    BURST_TRY_HIP(b, hipMemset(b->fits, 0, sizeof(clim::FitState) * n));
    BURST_TRY_HIP(b, hipMalloc(&b->queues, sizeof(clim::BurstQueue) * 2));
    BURST_TRY_HIP(b, hipMemset(b->queues, 0, sizeof(clim::BurstQueue) * 2));
    // the cross-frame table (the frames' raw fits are the burst's own states)
    std::vector<clim::BurstFrame> table(b->cross.size());
    std::string sig;
    for (size_t j = 0; j < b->cross.size(); j++) {
        const int i = b->cross[j];
        const FrameCall& f = b->calls[i];
        clim::BurstFrame& t = table[j];
        t.raw = f.raw.src<true>();
        t.green = b->green[i].img;
        t.ycc = b->ycc[i].img;
        const int bayer = f.p->bayerPattern;
        t.gxo = 0;
        t.gyo = (bayer == 2 || bayer == 3) ? 1 : 0;  // raw_green site, as launch_green
        t.lv = clim::ParamRef{b->fits[i].out + 5, {0}};
        t.cv = clim::ParamRef{b->fits[i].out, {0}};
        t.b = f.mosaic.b;
        t.red_is_c1 = (bayer == 3 || bayer == 0) ? 1 : 3;  // as launch_rb_despeckle
        t.clip = 1.0f;
        std::memcpy(t.c2y, f.c2y, sizeof(t.c2y));
        t.var_a = f.p->noiseModel.pyramidNlf[0][0];
        t.var_b = f.p->noiseModel.pyramidNlf[0][3];
        for (int kind = 0; kind < 2; kind++) {
            int tx = 0, tiles = 0;
            clim::burst_tiles(f.w, f.h, kind, &tx, &tiles);
            t.tiles_x[kind] = tx;
            b->total[kind] += tiles;
            t.tile_end[kind] = b->total[kind];
        }
        sig += std::to_string(f.w) + "x" + std::to_string(f.h) + ";";
    }
    if (!table.empty()) {
        BURST_TRY_HIP(b, hipMalloc(&b->table, sizeof(clim::BurstFrame) * table.size()));
        BURST_TRY_HIP(b, hipMemcpy(b->table, table.data(), sizeof(clim::BurstFrame) * table.size(),
                                   hipMemcpyHostToDevice));
    }
    BURST_TRY_HIP(b, hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    BURST_TRY_HIP(b, hipEventCreateWithFlags(&b->mid, hipEventDisableTiming));
    for (int k = 1; k < kLanes; k++) {
        BURST_TRY_HIP(b, hipEventCreateWithFlags(&b->lane_end[k], hipEventDisableTiming));
        BURST_TRY_HIP(b, hipEventCreateWithFlags(&b->lane_mid[k], hipEventDisableTiming));
    }
    if (b->timed) {
        BURST_TRY_HIP(b, hipEventCreate(&b->t_start));
        b->t_frame.assign(n, nullptr);
        for (int i = 0; i < n; i++) BURST_TRY_HIP(b, hipEventCreate(&b->t_frame[i]));
    }
    // the plain shape for the first (eager) run, which also sizes every lane
    // buffer the graph will use
    for (int kind = 0; kind < 2; kind++) {
        b->tune.waves[kind] = kind == 0 ? 8 : 4;
        b->tune.blocks_per_cu[kind] = 1;
        b->tune.grab[kind] = 1;
    }
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(sig);
        if (it != g_tune.end() && !sig.empty()) {
            b->tune = it->second;
            b->cached = true;
        }
    }
    BURST_TRY_HIP(b, hipStreamSynchronize(ctx->stream));  // the caller's queued work (e.g. the inputs) first
    CLIM_TRY(record(b, false, false));
    BURST_TRY_HIP(b, hipStreamSynchronize(b->lane[0]->stream));
    if (!b->cached && !b->cross.empty()) {
        // the frames' fits are in place: time the shapes of each stage (the
        // green planes the second stage reads are the chosen shape's)
        CLIM_TRY(autotune(b, 0));
        clim::launch_demosaic_burst(0, b->tune.waves[0], b->half, b->table, b->total[0], b->queues,
                                    b->tune.blocks_per_cu[0] * b->cus, b->tune.grab[0], b->lane[0]->stream);
        CLIM_TRY(autotune(b, 1));
        BURST_TRY_HIP(b, hipStreamSynchronize(b->lane[0]->stream));
        b->tune.form_known = 0;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[sig] = b->tune;
    }
    // the per-frame form once eagerly (its lane buffers sized), then both
    // forms captured; unless the choice is known, both are timed and the
    // faster is kept
    CLIM_TRY(record(b, false, true));
    BURST_TRY_HIP(b, hipStreamSynchronize(b->lane[0]->stream));
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    auto release = [&]() {
        for (int k = 0; k < 2; k++) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            exec[k] = nullptr;
            graph[k] = nullptr;
        }
    };
    const int forced = (flags & CLIM_BURST_CROSS_FRAME) ? 0 : (flags & CLIM_BURST_PER_FRAME) ? 1 : -1;
    const bool choose = forced < 0 && !(b->cached && b->tune.form_known);
    if (forced >= 0) {
        b->tune.form = b->cross.empty() ? 1 : forced;
        b->tune.form_ms[0] = b->tune.form_ms[1] = 0.0f;
    }
    for (int form = 0; form < 2; form++) {
        if (!choose && form != b->tune.form) continue;
        if (form == 0 && b->cross.empty()) continue;
        BURST_TRY_HIP(b, hipStreamBeginCapture(b->lane[0]->stream, hipStreamCaptureModeThreadLocal));
        const clim_status rs = record(b, b->timed, form == 1);
        const hipError_t ce = hipStreamEndCapture(b->lane[0]->stream, &graph[form]);
        if (rs != CLIM_OK || ce != hipSuccess) {
            release();
            if (rs != CLIM_OK) return rs;
            return ctx_hip(ctx, ce, "burst: hipStreamEndCapture");
        }
        const hipError_t ie = hipGraphInstantiate(&exec[form], graph[form], nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            release();
            return ctx_hip(ctx, ie, "burst: hipGraphInstantiate");
        }
    }
    if (choose) {
        // one replay of each to settle, then three timed
        b->tune.form_ms[0] = b->tune.form_ms[1] = 0.0f;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hipError_t e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        for (int form = 0; form < 2 && e == hipSuccess; form++) {
            if (!exec[form]) continue;
            e = hipGraphLaunch(exec[form], ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
            for (int r = 0; r < 3 && e == hipSuccess; r++) e = hipGraphLaunch(exec[form], ctx->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            b->tune.form_ms[form] = ms / 3.0f;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (e != hipSuccess) {
            release();
            return ctx_hip(ctx, e, "burst: timing the recorded forms");
        }
        b->tune.form = !exec[0] || (exec[1] && b->tune.form_ms[1] < b->tune.form_ms[0]) ? 1 : 0;
        b->tune.form_known = 1;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[sig] = b->tune;
    }
    b->graph = graph[b->tune.form];
    b->exec = exec[b->tune.form];
    graph[b->tune.form] = nullptr;
    exec[b->tune.form] = nullptr;
    release();
    return CLIM_OK;
}

extern "C" clim_status clim_burst_create(clim_ctx* ctx, int32_t n, const clim_image* const* raws,
                                         clim_image* const* out_rgba8, clim_image* const* out_f32,
                                         const clim_demosaic_params* params, uint32_t flags, clim_burst** out) {
    if (!ctx || !out || n <= 0 || !raws || !out_rgba8 || !params)
        return ctx_fail(ctx, CLIM_EINVAL, "burst: bad argument");
    if (flags & ~(uint32_t)(CLIM_BURST_FRAME_TIMES | CLIM_BURST_CROSS_FRAME | CLIM_BURST_PER_FRAME))
        return ctx_fail(ctx, CLIM_EINVAL, "burst: unknown flag");
    if ((flags & CLIM_BURST_CROSS_FRAME) && (flags & CLIM_BURST_PER_FRAME))
        return ctx_fail(ctx, CLIM_EINVAL, "burst: CLIM_BURST_CROSS_FRAME and CLIM_BURST_PER_FRAME together");
    *out = nullptr;
    clim_burst* b = new clim_burst();
    b->ctx = ctx;
    const clim_status s = burst_build(b, n, raws, out_rgba8, out_f32, params, flags);
    if (s != CLIM_OK) {
        const std::string msg = ctx->last_error;
        b->ctx = nullptr;  // the context's error stays as set
        clim_burst_destroy(b);
        ctx->last_error = msg;
        return s;
    }
    *out = b;
    return CLIM_OK;
}

extern "C" clim_status clim_burst_launch(clim_burst* b) {
    if (!b || !b->exec) return CLIM_EINVAL;
    BURST_TRY_HIP(b, hipGraphLaunch(b->exec, b->ctx->stream));
    return CLIM_OK;
}

extern "C" clim_status clim_burst_frame_times(clim_burst* b, float* ms) {
    if (!b || !ms) return CLIM_EINVAL;
    if (!b->timed) return burst_fail(b, CLIM_EINVAL, "created without CLIM_BURST_FRAME_TIMES");
    BURST_TRY_HIP(b, hipStreamSynchronize(b->ctx->stream));
    for (int i = 0; i < b->n; i++) BURST_TRY_HIP(b, hipEventElapsedTime(ms + i, b->t_start, b->t_frame[i]));
    return CLIM_OK;
}

extern "C" clim_status clim_burst_tuning_read(const clim_burst* b, clim_burst_tuning* t) {
    if (!b || !t) return CLIM_EINVAL;
    for (int k = 0; k < 2; k++) {
        t->waves[k] = b->tune.waves[k];
        t->blocks_per_cu[k] = b->tune.blocks_per_cu[k];
        t->grab[k] = b->tune.grab[k];
        t->us[k] = b->tune.us[k];
        t->us_first[k] = b->tune.us_first[k];
        t->candidates[k] = b->tune.candidates[k];
        t->tiles[k] = b->total[k];
    }
    t->cross_frames = (int32_t)b->cross.size();
    t->cached = b->cached ? 1 : 0;
    t->form = b->tune.form;
    t->form_ms[0] = b->tune.form_ms[0];
    t->form_ms[1] = b->tune.form_ms[1];
    return CLIM_OK;
}
