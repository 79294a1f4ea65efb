// strength.cpp - the learned denoise strength: weights, inference inside the
// pyramid, training on frames the pipeline generates, and the C ABI entry
// points clim_strength_* (include/climage.h).
//
// Plug point: the per-level filter of PyramidalDenoise
// (ImageDenoiser::denoise, cpp/pyramidal_denoise.hpp:24-34).  The network
// runs once per frame after the per-level noise fits, on the grid of pyramid
// level 3, and writes a map of per-level multipliers of the bilateral range
// sigma (cl/demosaic.cl:735-765) or of the guided filter's epsilon
// (cl/demosaic.cl:767-851).  A level takes the map only when the gate is
// open: the network's predicted error of the level against the plain
// strength is below it (mode GATED), or the caller forces it open or shut.
// Shut, the level runs the plain filter (the same instructions and results
// as without the network).
//
// Training data: seeded synthetic frames with their clean scenes
// (clim_generate_bayer) go through the plain pipeline; each level is then
// denoised at seven constant strengths, and the error of each against the
// clean scene's pyramid (camera RGB) on the grid of level 3 forms the error
// curves the loss interpolates.  The plain pyramid is the judge: the loss of
// each crop is normalised by its plain-strength error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "ctx.h"

int inverse3x3_scaled(const float mf[9], double scale, float out[9]);

using clim::Img;

static_assert(clim::kSnParams == CLIM_STRENGTH_PARAMS, "the ABI's parameter count is the network's");

namespace {

template <typename T>
hipError_t dev_alloc(T** p, size_t n) {
    return hipMalloc(reinterpret_cast<void**>(p), sizeof(T) * std::max<size_t>(n, 1));
}

clim_status devbuf(clim_ctx* ctx, DevBuf& b, size_t bytes) {
    if (b.p && b.bytes >= bytes) return CLIM_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    CLIM_TRY_HIP(ctx, hipMalloc(&b.p, std::max<size_t>(bytes, 256)));
    b.bytes = std::max<size_t>(bytes, 256);
    return CLIM_OK;
}

void free_buf(DevBuf& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

clim_status alloc_weights(clim_ctx* ctx, clim::SnWeights& w) {
    if (w.master) return CLIM_OK;
    CLIM_TRY_HIP(ctx, dev_alloc(&w.master, clim::kSnParams));
    for (int l = 0; l < clim::kSnLayers; l++) {
        CLIM_TRY_HIP(ctx, dev_alloc(&w.fwd[l], 32 * (size_t)clim::sn_kpad(l)));
        if (l > 0) CLIM_TRY_HIP(ctx, dev_alloc(&w.dgrad[l], 32 * 288));
        CLIM_TRY_HIP(ctx, dev_alloc(&w.bias[l], 32));
    }
    return CLIM_OK;
}

void free_weights(clim::SnWeights& w) {
    if (w.master) (void)hipFree(w.master);
    for (int l = 0; l < clim::kSnLayers; l++) {
        if (w.fwd[l]) (void)hipFree(w.fwd[l]);
        if (w.dgrad[l]) (void)hipFree(w.dgrad[l]);
        if (w.bias[l]) (void)hipFree(w.bias[l]);
    }
    w = clim::SnWeights{};
}

// bfloat16 of a float, round to nearest even (the device's conversion)
uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float host_from_bf16(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// master weights as the operands see them: weights rounded to bfloat16,
// biases kept in FP32
void round_weights(std::vector<float>& p) {
    for (int l = 0; l < clim::kSnLayers; l++) {
        const int off = clim::sn_layer_offset(l), nw = 9 * clim::sn_cin(l) * clim::sn_cout(l);
        for (int i = 0; i < nw; i++) p[off + i] = host_from_bf16(host_bf16(p[off + i]));
    }
}

}  // namespace

// training and gradient buffers of one context
struct SnTrainer {
    clim::SnWeights wt{};
    float *grad = nullptr, *m = nullptr, *v = nullptr;
    size_t cap_px = 0;  // pixels of the batch buffers
    uint16_t *feat = nullptr, *act[3] = {}, *dzA = nullptr, *dzB = nullptr;
    float *out32 = nullptr, *curves = nullptr, *norm = nullptr, *wpart = nullptr, *lpart = nullptr;
    int cap_img = 0, lpart_cap = 0;

    void free_batch() {
        uint16_t** u[] = {&feat, &act[0], &act[1], &act[2], &dzA, &dzB};
        for (auto p : u)
            if (*p) (void)hipFree(*p), *p = nullptr;
        float** f[] = {&out32, &curves, &norm, &lpart};
        for (auto p : f)
            if (*p) (void)hipFree(*p), *p = nullptr;
        cap_px = 0;
        cap_img = 0;
        lpart_cap = 0;
    }
    ~SnTrainer() {
        free_batch();
        free_weights(wt);
        float** f[] = {&grad, &m, &v, &wpart};
        for (auto p : f)
            if (*p) (void)hipFree(*p), *p = nullptr;
    }
};

void sn_release(clim_ctx* ctx) {
    SnState& s = ctx->sn;
    delete s.trainer;
    s.trainer = nullptr;
    free_weights(s.wt);
    DevBuf* b[] = {&s.map, &s.gate_sum, &s.open};
    for (DevBuf* p : b) free_buf(*p);
    s.use = nullptr;
}

static clim_status trainer_get(clim_ctx* ctx, SnTrainer** out) {
    SnState& s = ctx->sn;
    if (!s.trainer) s.trainer = new SnTrainer();
    SnTrainer* t = s.trainer;
    CLIM_TRY(alloc_weights(ctx, t->wt));
    if (!t->grad) {
        CLIM_TRY_HIP(ctx, dev_alloc(&t->grad, clim::kSnParams));
        CLIM_TRY_HIP(ctx, dev_alloc(&t->m, clim::kSnParams));
        CLIM_TRY_HIP(ctx, dev_alloc(&t->v, clim::kSnParams));
        CLIM_TRY_HIP(ctx, dev_alloc(&t->wpart, (size_t)clim::kSnWgradBlocks * (9 * 32 * 32 + 32)));
    }
    *out = t;
    return CLIM_OK;
}

static clim_status trainer_batch(clim_ctx* ctx, SnTrainer* t, int n, size_t px) {
    if (px <= t->cap_px && n <= t->cap_img) return CLIM_OK;
    t->free_batch();
    CLIM_TRY_HIP(ctx, dev_alloc(&t->feat, px * 8));
    for (int i = 0; i < 3; i++) CLIM_TRY_HIP(ctx, dev_alloc(&t->act[i], px * 32));
    CLIM_TRY_HIP(ctx, dev_alloc(&t->dzA, px * 32));
    CLIM_TRY_HIP(ctx, dev_alloc(&t->dzB, px * 32));
    CLIM_TRY_HIP(ctx, dev_alloc(&t->out32, px * 32));
    CLIM_TRY_HIP(ctx, dev_alloc(&t->curves, px * clim::kSnCurves));
    CLIM_TRY_HIP(ctx, dev_alloc(&t->norm, (size_t)n * 10));
    t->lpart_cap = (int)((px + 255) / 256) * 2;
    CLIM_TRY_HIP(ctx, dev_alloc(&t->lpart, (size_t)t->lpart_cap));
    t->cap_px = px;
    t->cap_img = n;
    return CLIM_OK;
}

// forward of a batch: feat -> act[0..2] -> out32 (raw outputs)
static void sn_forward_train(const clim::SnWeights& wt, SnTrainer* t, int n, int h, int w, hipStream_t st) {
    const uint16_t* in = t->feat;
    for (int l = 0; l < clim::kSnLayers; l++) {
        clim::SnConv a{};
        a.x = in;
        a.wa = wt.fwd[l];
        a.bias = wt.bias[l];
        a.n = n;
        a.h = h;
        a.w = w;
        if (l < clim::kSnLayers - 1) {
            a.y = t->act[l];
            clim::launch_sn_conv(a, clim::sn_cin(l), clim::SN_RELU, st);
            in = t->act[l];
        } else {
            a.y32 = t->out32;
            clim::launch_sn_conv(a, clim::sn_cin(l), clim::SN_RAW, st);
        }
    }
}

// loss and the full gradient of a batch already in the trainer's buffers;
// returns the loss block count (partials in t->lpart)
static int sn_backward_train(const clim::SnWeights& wt, SnTrainer* t, int n, int h, int w, hipStream_t st) {
    const int hw = h * w;
    const float scale = 1.0f / ((float)n * (float)hw * 10.0f);
    const int nb = clim::launch_sn_loss(t->out32, t->curves, t->norm, n, hw, scale, t->dzA, t->lpart, st);
    uint16_t* dz = t->dzA;
    uint16_t* other = t->dzB;
    for (int l = clim::kSnLayers - 1; l >= 0; l--) {
        const uint16_t* x = l > 0 ? t->act[l - 1] : t->feat;
        clim::launch_sn_wgrad(x, dz, n, h, w, clim::sn_cin(l), clim::sn_cout(l), t->wpart,
                              t->grad + clim::sn_layer_offset(l), st);
        if (l > 0) {
            clim::SnConv a{};
            a.x = dz;
            a.wa = wt.dgrad[l];
            a.bias = nullptr;
            a.n = n;
            a.h = h;
            a.w = w;
            a.y = other;
            a.mask = t->act[l - 1];
            clim::launch_sn_conv(a, 32, clim::SN_MASK, st);
            std::swap(dz, other);
        }
    }
    return nb;
}

// --------------------------------------------------------------- inference

// Runs the network on level 3 of the frame's pyramid (lvl[0..levels)) with
// the calibrated per-level noise models in the context's fit slots, and
// fills `sr` for the level filters.  algorithm: which filter each level runs
// (GuidedFast runs the bilateral on the last level, as the reference's
// pyramid does).
clim_status strength_infer(clim_ctx* ctx, const Img* lvl, int levels, int algorithm, clim::StrengthRef* sr) {
    SnState& s = ctx->sn;
    *sr = clim::StrengthRef{nullptr, 0, 0, nullptr, 0.0f};
    if (s.mode == CLIM_STRENGTH_OFF || !s.use || levels <= clim::kSnGrid ||
        (algorithm != CLIM_DENOISE_BILATERAL && algorithm != CLIM_DENOISE_GUIDED_FAST))
        return CLIM_OK;
    const int gw = lvl[clim::kSnGrid].w, gh = lvl[clim::kSnGrid].h;
    const size_t px = (size_t)gw * gh;
    CLIM_TRY(devbuf(ctx, s.map, px * clim::kSnLevels * sizeof(float)));
    if (!s.gate_sum.p) {
        CLIM_TRY(devbuf(ctx, s.gate_sum, clim::kSnLevels * sizeof(long long)));
        CLIM_TRY_HIP(ctx, hipMemset(s.gate_sum.p, 0, s.gate_sum.bytes));  // later frames: the gate kernel clears
    }
    CLIM_TRY(devbuf(ctx, s.open, 3 * clim::kSnLevels * sizeof(int)));
    s.gw = gw;
    s.gh = gh;
    s.levels = levels;
    hipStream_t st = ctx->stream;
    ProfScope ps(ctx, ST_STRENGTH, (long long)px);
    // GuidedFast runs the bilateral on its last level (the reference's pyramid does)
    const int guided_mask = algorithm == CLIM_DENOISE_GUIDED_FAST ? ((1 << (levels - 1)) - 1) : 0;
    clim::launch_sn_infer(lvl[clim::kSnGrid], ctx->fit[1].cal, ctx->fit[1 + clim::kSnGrid].cal, *s.use,
                          static_cast<float*>(s.map.p),
                          guided_mask, static_cast<long long*>(s.gate_sum.p), st);
    clim::launch_sn_gate(static_cast<long long*>(s.gate_sum.p), levels, s.mode, static_cast<int*>(s.open.p), st);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    *sr = clim::StrengthRef{static_cast<const float*>(s.map.p), gw, gh, static_cast<const int*>(s.open.p), 0.0f};
    return CLIM_OK;
}

// ------------------------------------------------------------------- C ABI

extern "C" clim_status clim_strength_set_weights(clim_ctx* ctx, const float* params, int64_t n) {
    if (!ctx || !params) return ctx_fail(ctx, CLIM_EINVAL, "strength_set_weights: null argument");
    if (n != clim::kSnParams)
        return ctx_fail(ctx, CLIM_EINVAL, "strength_set_weights: expected " + std::to_string(clim::kSnParams) +
                                              " parameters, got " + std::to_string(n));
    for (int64_t i = 0; i < n; i++)
        if (!std::isfinite(params[i])) return ctx_fail(ctx, CLIM_EINVAL, "strength_set_weights: non-finite parameter");
    std::vector<float> p(params, params + n);
    round_weights(p);
    SnState& s = ctx->sn;
    CLIM_TRY(alloc_weights(ctx, s.wt));
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(s.wt.master, p.data(), sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    clim::launch_sn_prep(s.wt, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s.use = &s.wt;
    return CLIM_OK;
}

extern "C" clim_status clim_strength_get_weights(clim_ctx* ctx, float* params, int64_t n) {
    if (!ctx || !params) return ctx_fail(ctx, CLIM_EINVAL, "strength_get_weights: null argument");
    if (n != clim::kSnParams) return ctx_fail(ctx, CLIM_EINVAL, "strength_get_weights: wrong parameter count");
    if (!ctx->sn.wt.master) return ctx_fail(ctx, CLIM_EINVAL, "strength_get_weights: no weights set");
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(params, ctx->sn.wt.master, sizeof(float) * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CLIM_OK;
}

extern "C" clim_status clim_strength_set_mode(clim_ctx* ctx, int32_t mode) {
    if (!ctx || mode < CLIM_STRENGTH_OFF || mode > CLIM_STRENGTH_FORCE_LEARNED)
        return ctx_fail(ctx, CLIM_EINVAL, "strength_set_mode: bad mode");
    if (mode != CLIM_STRENGTH_OFF && !ctx->sn.use)
        return ctx_fail(ctx, CLIM_EINVAL, "strength_set_mode: no weights set");
    ctx->sn.mode = mode;
    return CLIM_OK;
}

extern "C" clim_status clim_strength_gate_read(clim_ctx* ctx, int32_t open[5], float mean_pred[5]) {
    if (!ctx || !open) return ctx_fail(ctx, CLIM_EINVAL, "strength_gate_read: null argument");
    SnState& s = ctx->sn;
    if (!s.open.p) return ctx_fail(ctx, CLIM_EINVAL, "strength_gate_read: no frame ran the network");
    int v[3 * clim::kSnLevels];
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(v, s.open.p, sizeof(v), hipMemcpyDeviceToHost, ctx->stream));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int l = 0; l < clim::kSnLevels; l++) {
        open[l] = v[l];
        if (mean_pred) {
            const long long sum = ((long long)v[clim::kSnLevels + l] << 32) |
                                  (long long)(unsigned int)v[2 * clim::kSnLevels + l];
            mean_pred[l] = (float)((double)sum / 65536.0 / ((double)s.gw * s.gh));
        }
    }
    return CLIM_OK;
}

extern "C" clim_status clim_strength_map_read(clim_ctx* ctx, float* map, int64_t capacity, int32_t* gw,
                                              int32_t* gh) {
    if (!ctx || !gw || !gh) return ctx_fail(ctx, CLIM_EINVAL, "strength_map_read: null argument");
    SnState& s = ctx->sn;
    if (!s.map.p) return ctx_fail(ctx, CLIM_EINVAL, "strength_map_read: no frame ran the network");
    *gw = s.gw;
    *gh = s.gh;
    const int64_t n = (int64_t)s.gw * s.gh * clim::kSnLevels;
    if (!map) return CLIM_OK;
    if (capacity < n) return ctx_fail(ctx, CLIM_EINVAL, "strength_map_read: buffer too small");
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(map, s.map.p, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CLIM_OK;
}

extern "C" clim_status clim_strength_loss_grad(clim_ctx* ctx, const float* params, const float* feat,
                                               const float* curves, const float* norm, int32_t n, int32_t h,
                                               int32_t w, float* out_raw, float* grad, double loss[2]) {
    if (!ctx || !params || !feat || !curves || !norm || n <= 0 || h <= 0 || w <= 0)
        return ctx_fail(ctx, CLIM_EINVAL, "strength_loss_grad: bad argument");
    SnTrainer* t = nullptr;
    CLIM_TRY(trainer_get(ctx, &t));
    const size_t px = (size_t)n * h * w;
    CLIM_TRY(trainer_batch(ctx, t, n, px));
    hipStream_t st = ctx->stream;
    std::vector<float> p(params, params + clim::kSnParams);
    round_weights(p);
    std::vector<uint16_t> fb(px * 8);
    for (size_t i = 0; i < px * 8; i++) fb[i] = host_bf16(feat[i]);
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(t->wt.master, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice, st));
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(t->feat, fb.data(), sizeof(uint16_t) * fb.size(), hipMemcpyHostToDevice, st));
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(t->curves, curves, sizeof(float) * px * clim::kSnCurves, hipMemcpyHostToDevice,
                                     st));
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(t->norm, norm, sizeof(float) * n * 10, hipMemcpyHostToDevice, st));
    clim::launch_sn_prep(t->wt, st);
    sn_forward_train(t->wt, t, n, h, w, st);
    const int nb = sn_backward_train(t->wt, t, n, h, w, st);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    std::vector<float> lp((size_t)nb * 2);
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(lp.data(), t->lpart, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, st));
    if (out_raw)
        CLIM_TRY_HIP(ctx, hipMemcpyAsync(out_raw, t->out32, sizeof(float) * px * 32, hipMemcpyDeviceToHost, st));
    if (grad)
        CLIM_TRY_HIP(ctx, hipMemcpyAsync(grad, t->grad, sizeof(float) * clim::kSnParams, hipMemcpyDeviceToHost, st));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(st));
    if (loss) {
        double a = 0, b = 0;
        for (int i = 0; i < nb; i++) {
            a += lp[2 * i];
            b += lp[2 * i + 1];
        }
        loss[0] = a / ((double)px * 10.0);
        loss[1] = b / ((double)px * 10.0);
    }
    return CLIM_OK;
}

// ----------------------------------------------------------------- training

// He-normal initial weights (seeded), zero biases; the last layer starts
// small so that the first strengths are close to the plain ones.
static void init_weights(uint32_t seed, std::vector<float>& p) {
    std::mt19937 rng(seed);
    p.assign(clim::kSnParams, 0.0f);
    for (int l = 0; l < clim::kSnLayers; l++) {
        const int cin = clim::sn_cin(l), cout = clim::sn_cout(l), off = clim::sn_layer_offset(l);
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * cin)) * (l == clim::kSnLayers - 1 ? 0.1f : 1.0f));
        for (int i = 0; i < 9 * cin * cout; i++) p[off + i] = nd(rng);
    }
    round_weights(p);
}

extern "C" clim_status clim_strength_train(clim_ctx* ctx, const clim_strength_train_config* cfg, float* log,
                                           int64_t log_n) {
    if (!ctx || !cfg) return ctx_fail(ctx, CLIM_EINVAL, "strength_train: null argument");
    const clim_strength_train_config c = *cfg;
    if (c.frames <= 0 || c.width < 256 || c.height < 256 || (c.width & 7) || (c.height & 7) || c.crop <= 0 ||
        c.epochs <= 0 || c.batch <= 0 || !(c.lr > 0.0f) || c.bayer < 0 || c.bayer > 3 || c.noise_levels <= 0 ||
        c.noise_levels > 8)
        return ctx_fail(ctx, CLIM_EINVAL, "strength_train: bad configuration");
    const int levels = 5;
    const int gw = c.width >> clim::kSnGrid, gh = c.height >> clim::kSnGrid, cs = c.crop;
    if (gw < cs || gh < cs) return ctx_fail(ctx, CLIM_EINVAL, "strength_train: crop larger than the grid");
    const int cx = gw / cs, cy = gh / cs, per_frame = cx * cy;
    const long long ncrops = (long long)per_frame * c.frames;
    if (log && log_n < (int64_t)c.epochs * 2) return ctx_fail(ctx, CLIM_EINVAL, "strength_train: log too small");
    hipStream_t st = ctx->stream;
    SnTrainer* t = nullptr;
    CLIM_TRY(trainer_get(ctx, &t));
    const int nb = std::min<int>(c.batch, (int)ncrops);
    CLIM_TRY(trainer_batch(ctx, t, nb, (size_t)nb * cs * cs));

    // frame buffers and the training set
    clim_image raw{}, clean{}, out8{};
    CLIM_TRY(clim_image_alloc(ctx, c.width, c.height, CLIM_FMT_R16U, &raw));
    CLIM_TRY(clim_image_alloc(ctx, c.width, c.height, CLIM_FMT_RGBA32F, &clean));
    CLIM_TRY(clim_image_alloc(ctx, c.width, c.height, CLIM_FMT_RGBA8, &out8));
    OwnedImg cl[5], tmp;
    DevBuf ffeat, fcurves, sfeat, scurves, snorm, origins, perm;
    auto cleanup = [&]() {
        clim_image_free(ctx, &raw);
        clim_image_free(ctx, &clean);
        clim_image_free(ctx, &out8);
        for (auto& o : cl) owned_free(o);
        owned_free(tmp);
        DevBuf* b[] = {&ffeat, &fcurves, &sfeat, &scurves, &snorm, &origins, &perm};
        for (DevBuf* p : b) free_buf(*p);
    };
    auto fail = [&](clim_status s) {
        (void)hipStreamSynchronize(st);
        cleanup();
        return s;
    };
#define SN_TRY(call)                          \
    do {                                      \
        clim_status s_ = (call);              \
        if (s_ != CLIM_OK) return fail(s_);   \
    } while (0)
#define SN_TRY_HIP(call)                                               \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return fail(ctx_hip(ctx, e_, #call));    \
    } while (0)
    for (int l = 0; l < levels; l++) SN_TRY(owned_alloc(ctx, cl[l], c.width >> l, c.height >> l, CLIM_FMT_RGBA32F));
    SN_TRY(owned_alloc(ctx, tmp, c.width, c.height, CLIM_FMT_RGBA32F));
    SN_TRY(devbuf(ctx, ffeat, (size_t)gw * gh * 8 * sizeof(uint16_t)));
    SN_TRY(devbuf(ctx, fcurves, (size_t)gw * gh * clim::kSnCurves * sizeof(float)));
    SN_TRY(devbuf(ctx, sfeat, (size_t)ncrops * cs * cs * 8 * sizeof(uint16_t)));
    SN_TRY(devbuf(ctx, scurves, (size_t)ncrops * cs * cs * clim::kSnCurves * sizeof(float)));
    SN_TRY(devbuf(ctx, snorm, (size_t)ncrops * 10 * sizeof(float)));
    SN_TRY(devbuf(ctx, origins, (size_t)per_frame * sizeof(int2)));
    const int steps = (int)((ncrops + nb - 1) / nb);
    SN_TRY(devbuf(ctx, perm, (size_t)steps * nb * sizeof(int)));
    {
        std::vector<int2> o(per_frame);
        for (int j = 0; j < cy; j++)
            for (int i = 0; i < cx; i++) o[j * cx + i] = make_int2(i * cs, j * cs);
        SN_TRY_HIP(hipMemcpy(origins.p, o.data(), sizeof(int2) * o.size(), hipMemcpyHostToDevice));
    }

    // 1. the training set: frames through the plain pipeline, error curves
    const int saved_mode = ctx->sn.mode;
    ctx->sn.mode = CLIM_STRENGTH_OFF;
    static const float kNoiseA[8] = {0.0005f, 0.001f, 0.002f, 0.004f, 0.008f, 0.016f, 0.003f, 0.006f};
    float c2y[9], y2c[9], ident[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    clim_cam_ycbcr(ident, c2y);
    inverse3x3_scaled(c2y, 1.0, y2c);
    const Img cleanv{(char*)clean.data, clean.width, clean.height, clean.pitch_bytes};
    for (int f = 0; f < c.frames; f++) {
        const float a = kNoiseA[f % c.noise_levels];
        clim_demosaic_params p;
        clim_default_params(&p, c.bayer, 64.0f, 4095.0f);
        p.levels = levels;
        SN_TRY(clim_generate_bayer(ctx, &raw, &clean, c.bayer, c.seed + (uint32_t)f, c.seed + 100000u + (uint32_t)f,
                                   a, a / 200.0f, 64.0f, 4095.0f));
        SN_TRY(clim_raw_to_rgb(ctx, &raw, &out8, nullptr, &p, nullptr));
        // the workspace now holds the noisy pyramid (level 0: rgbB) and the
        // calibrated noise models (fit slots 1..levels)
        Img lv[5];
        lv[0] = ctx->ws.rgbB.img;
        for (int l = 1; l < levels; l++) lv[l] = ctx->ws.pyr[l].img;
        clim::launch_transform(cleanv, cl[0].img, c2y, st);
        for (int l = 0; l + 1 < levels; l++) clim::launch_downsample(cl[l].img, cl[l + 1].img, st);
        for (int l = 0; l < levels; l++) {
            const Img den{tmp.img.p, lv[l].w, lv[l].h, tmp.img.pitch};
            clim::ParamRef cal{};
            cal.dev = ctx->fit[1 + l].cal;
            for (int fl = 0; fl < 2; fl++) {
                for (int k = 0; k < clim::kSnCurveK; k++) {
                    const float sk = 0.5f * (float)(k - 3);
                    const clim::StrengthRef sr{nullptr, 0, 0, nullptr, std::exp2(fl ? 2.0f * sk : sk)};
                    if (fl == 0)
                        clim::launch_bilateral(lv[l], den, cal, 1.0f, 1.0f, st, &sr, l);
                    else
                        clim::launch_guided_fast(lv[l], den, cal, st, &sr, l);
                    clim::launch_sn_curve(den, cl[l].img, y2c, l, gw, gh, static_cast<float*>(fcurves.p),
                                          (l * 2 + fl) * clim::kSnCurveK + k, st);
                }
            }
        }
        clim::launch_sn_features(lv[clim::kSnGrid], ctx->fit[1].cal, ctx->fit[1 + clim::kSnGrid].cal,
                                 static_cast<uint16_t*>(ffeat.p), st);
        const size_t base = (size_t)f * per_frame;
        clim::launch_sn_gather(static_cast<uint16_t*>(ffeat.p), static_cast<float*>(fcurves.p), gw, cs,
                               static_cast<int2*>(origins.p), per_frame,
                               static_cast<uint16_t*>(sfeat.p) + base * cs * cs * 8,
                               static_cast<float*>(scurves.p) + base * cs * cs * clim::kSnCurves,
                               static_cast<float*>(snorm.p) + base * 10, st);
        SN_TRY_HIP(hipGetLastError());
        if (c.verbose && (f % 16 == 15 || f == c.frames - 1)) {
            SN_TRY_HIP(hipStreamSynchronize(st));
            std::fprintf(stderr, "strength_train: %d of %d frames generated\n", f + 1, c.frames);
        }
    }
    ctx->sn.mode = saved_mode;

    // 2. training: Adam over shuffled batches of crops
    std::vector<float> p0;
    init_weights(c.seed, p0);
    SN_TRY_HIP(hipMemcpyAsync(t->wt.master, p0.data(), sizeof(float) * p0.size(), hipMemcpyHostToDevice, st));
    SN_TRY_HIP(hipMemsetAsync(t->m, 0, sizeof(float) * clim::kSnParams, st));
    SN_TRY_HIP(hipMemsetAsync(t->v, 0, sizeof(float) * clim::kSnParams, st));
    std::mt19937 rng(c.seed ^ 0x5eedu);
    std::vector<int> order(ncrops), pidx((size_t)steps * nb);
    std::vector<float> lp((size_t)t->lpart_cap);
    int step_no = 0;
    for (int e = 0; e < c.epochs; e++) {
        std::iota(order.begin(), order.end(), 0);
        std::shuffle(order.begin(), order.end(), rng);
        for (size_t i = 0; i < pidx.size(); i++) pidx[i] = order[i % ncrops];
        SN_TRY_HIP(hipStreamSynchronize(st));  // the previous epoch's batches are assembled
        SN_TRY_HIP(hipMemcpyAsync(perm.p, pidx.data(), sizeof(int) * pidx.size(), hipMemcpyHostToDevice, st));
        // learning rate: constant, a tenth over the last fifth of the epochs
        const float lr = e >= c.epochs - std::max(1, c.epochs / 5) ? c.lr * 0.1f : c.lr;
        double ls = 0.0, lr_sum = 0.0;
        for (int s = 0; s < steps; s++) {
            clim::launch_sn_batch(static_cast<uint16_t*>(sfeat.p), static_cast<float*>(scurves.p),
                                  static_cast<float*>(snorm.p), static_cast<int*>(perm.p) + (size_t)s * nb, nb, cs,
                                  t->feat, t->curves, t->norm, st);
            clim::launch_sn_prep(t->wt, st);
            sn_forward_train(t->wt, t, nb, cs, cs, st);
            const int lb = sn_backward_train(t->wt, t, nb, cs, cs, st);
            clim::launch_sn_adam(t->wt.master, t->grad, t->m, t->v, clim::kSnParams, lr, 0.9f, 0.999f, 1e-8f,
                                 ++step_no, st);
            if (s == steps - 1 || (s % 64) == 63) {
                // the loss of this step (a sample of the epoch's)
                SN_TRY_HIP(hipMemcpyAsync(lp.data(), t->lpart, sizeof(float) * lb * 2, hipMemcpyDeviceToHost, st));
                SN_TRY_HIP(hipStreamSynchronize(st));
                double a = 0, b = 0;
                for (int i = 0; i < lb; i++) {
                    a += lp[2 * i];
                    b += lp[2 * i + 1];
                }
                const double denom = (double)nb * cs * cs * 10.0;
                ls = a / denom;
                lr_sum = b / denom;
            }
        }
        SN_TRY_HIP(hipGetLastError());
        if (log) {
            log[2 * e] = (float)ls;
            log[2 * e + 1] = (float)lr_sum;
        }
        if (c.verbose) std::fprintf(stderr, "strength_train: epoch %d: losses %.5f %.5f\n", e, ls, lr_sum);
    }
    // the trained weights become the context's inference weights
    std::vector<float> pt(clim::kSnParams);
    SN_TRY_HIP(hipMemcpyAsync(pt.data(), t->wt.master, sizeof(float) * pt.size(), hipMemcpyDeviceToHost, st));
    SN_TRY_HIP(hipStreamSynchronize(st));
    cleanup();
#undef SN_TRY
#undef SN_TRY_HIP
    return clim_strength_set_weights(ctx, pt.data(), (int64_t)pt.size());
}
