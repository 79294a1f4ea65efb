// capi.cpp - context, images and stand-alone operators of the C ABI
// (include/climage.h).  The pipeline entry points are in pipeline.cpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "ctx.h"

using clim::Img;
using clim::ParamRef;

// ------------------------------------------------------------ helpers

clim_status ctx_fail(clim_ctx* ctx, clim_status s, const std::string& msg) {
    if (ctx) ctx->last_error = msg;
    return s;
}

clim_status ctx_hip(clim_ctx* ctx, hipError_t e, const char* what) {
    return ctx_fail(ctx, e == hipErrorOutOfMemory ? CLIM_ENOMEM : CLIM_EHIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" clim_status clim_set_precision(clim_ctx* ctx, int32_t precision) {
    if (!ctx || (precision != CLIM_PRECISION_FP32 && precision != CLIM_PRECISION_MIXED))
        return ctx_fail(ctx, CLIM_EINVAL, "set_precision: bad precision");
    ctx->precision = precision;
    return CLIM_OK;
}

int32_t clim_format_texel_bytes(int32_t format) {
    switch (format) {
        case CLIM_FMT_R16U: return 2;
        case CLIM_FMT_R32F: return 4;
        case CLIM_FMT_RGBA32F: return 16;
        case CLIM_FMT_RGBA8: return 4;
        case CLIM_FMT_RG32F: return 8;
        case CLIM_FMT_RGBA16F: return 8;
        default: return 0;
    }
}

static size_t pitch_for(int w, int format) {
    const size_t row = (size_t)w * clim_format_texel_bytes(format);
    return (row + 255) / 256 * 256;
}

clim_status owned_alloc(clim_ctx* ctx, OwnedImg& o, int w, int h, int format) {
    const bool same = o.img.p && o.format == format;
    if (same && w <= o.cap_w && h <= o.cap_h) {
        o.img.w = w;
        o.img.h = h;
        o.img.half = format == CLIM_FMT_RGBA16F ? 1 : 0;
        return CLIM_OK;
    }
    const int cw = same ? std::max(w, o.cap_w) : w, ch = same ? std::max(h, o.cap_h) : h;
    owned_free(o);
    const size_t pitch = pitch_for(cw, format);
    void* p = nullptr;
    CLIM_TRY_HIP(ctx, hipMalloc(&p, pitch * (size_t)ch));
    o.img = Img{(char*)p, w, h, (int)pitch, format == CLIM_FMT_RGBA16F ? 1 : 0};
    o.format = format;
    o.cap_w = cw;
    o.cap_h = ch;
    return CLIM_OK;
}

void owned_free(OwnedImg& o) {
    if (o.img.p) (void)hipFree(o.img.p);
    o.img = Img{nullptr, 0, 0, 0};
    o.format = 0;
    o.cap_w = o.cap_h = 0;
}

// OpenCL Bayer table, cl/demosaic.cl:29-34
static const int kBayerOffsets[4][4][2] = {
    {{1, 0}, {0, 0}, {0, 1}, {1, 1}},
    {{0, 1}, {0, 0}, {1, 0}, {1, 1}},
    {{0, 0}, {0, 1}, {1, 1}, {1, 0}},
    {{1, 1}, {0, 1}, {0, 0}, {1, 0}},
};

clim::Bayer make_bayer(int bayer) {
    clim::Bayer b{};
    b.site_packed = 0;
    for (int c = 0; c < 4; c++) {
        b.ox[c] = kBayerOffsets[bayer][c][0];
        b.oy[c] = kBayerOffsets[bayer][c][1];
        const int parity = ((b.oy[c] & 1) << 1) | (b.ox[c] & 1);
        b.site_packed |= c << (parity * 2);
    }
    return b;
}

int scene_colors(int bayer) {
    static const int color_of[4] = {0, 1, 2, 1};
    int packed = 0;
    for (int c = 0; c < 4; c++) {
        const int parity = ((kBayerOffsets[bayer][c][1] & 1) << 1) | (kBayerOffsets[bayer][c][0] & 1);
        packed |= color_of[c] << (parity * 2);
    }
    return packed;
}

static clim_status check_img(clim_ctx* ctx, const clim_image* im, int format, const char* name) {
    if (!im || !im->data) return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": null image");
    if (im->format != format)
        return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": wrong format " + std::to_string(im->format));
    if (im->width <= 0 || im->height <= 0)
        return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": empty image");
    if (im->pitch_bytes < im->width * clim_format_texel_bytes(format))
        return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": pitch smaller than a row");
    if ((format == CLIM_FMT_RGBA32F || format == CLIM_FMT_RG32F) && (im->pitch_bytes % 16) != 0)
        return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": pitch must be a multiple of 16 bytes");
    return CLIM_OK;
}

static clim_status check_same(clim_ctx* ctx, const clim_image* a, const clim_image* b, const char* what) {
    if (a->width != b->width || a->height != b->height)
        return ctx_fail(ctx, CLIM_EINVAL, std::string(what) + ": image sizes differ");
    return CLIM_OK;
}

static Img view(const clim_image* im) {
    return Img{(char*)im->data, im->width, im->height, im->pitch_bytes, im->format == CLIM_FMT_RGBA16F ? 1 : 0};
}

static ParamRef host_params(const float* v, int n) {
    ParamRef r{};
    r.dev = nullptr;
    for (int i = 0; i < n && i < 8; i++) r.host[i] = v[i];
    return r;
}

// -------------------------------------------------------------- context

extern "C" const char* clim_status_string(clim_status s) {
    switch (s) {
        case CLIM_OK: return "ok";
        case CLIM_EINVAL: return "invalid argument";
        case CLIM_ENOMEM: return "out of memory";
        case CLIM_EHIP: return "HIP runtime error";
        case CLIM_ENUMERIC: return "numerical failure";
        case CLIM_ENODEV: return "no gfx950 device";
        default: return "unknown status";
    }
}

extern "C" const char* clim_last_error(const clim_ctx* ctx) { return ctx ? ctx->last_error.c_str() : ""; }

extern "C" clim_status clim_create(clim_ctx** out, int device) {
    if (!out) return CLIM_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return CLIM_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CLIM_ENODEV;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return CLIM_ENODEV;
    if (hipSetDevice(device) != hipSuccess) return CLIM_ENODEV;
    clim_ctx* ctx = new clim_ctx();
    ctx->device = device;
    clim_noalloc_init(device, &ctx->noalloc_scratch);
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&ctx->fit, sizeof(clim::FitState) * kFitJobs);
    if (e == hipSuccess) e = hipMemset(ctx->fit, 0, sizeof(clim::FitState) * kFitJobs);
    if (e == hipSuccess) e = hipMalloc(&ctx->partials, sizeof(clim_fx128) * kFitJobs * clim::kFitBlocksPerJob * kFitK);
    if (e == hipSuccess) e = hipMalloc(&ctx->fit_counts, sizeof(unsigned int) * kFitJobs * clim::kFitBlocksPerJob);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->fit_host, sizeof(clim::FitState) * kFitJobs, 0);
    if (e == hipSuccess) e = hipMalloc(&ctx->scene, sizeof(clim::SceneParams));
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->scene_host, sizeof(clim::SceneParams), 0);
    if (e != hipSuccess) {
        clim_destroy(ctx);
        return e == hipErrorOutOfMemory ? CLIM_ENOMEM : CLIM_EHIP;
    }
    *out = ctx;
    return CLIM_OK;
}

static void free_workspace(Workspace& ws) {
    OwnedImg* all[] = {&ws.scaled, &ws.green, &ws.rgbA, &ws.rgbB, &ws.raw_mean, &ws.raw_var};
    for (OwnedImg* o : all) owned_free(*o);
    for (int i = 0; i < CLIM_MAX_LEVELS; i++) {
        owned_free(ws.pyr[i]);
        owned_free(ws.den[i]);
        owned_free(ws.stats[i]);
    }
    ws.w = ws.h = ws.levels = 0;
    ws.algorithm = -1;
}

extern "C" void clim_destroy(clim_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_workspace(ctx->ws);
    sn_release(ctx);
    for (auto& o : ctx->op_guided_imgs) owned_free(o);
    owned_free(ctx->op_mean);
    owned_free(ctx->op_var);
    owned_free(ctx->op_stats);
    for (clim_ctx* l : ctx->lanes)
        if (l) clim_destroy(l);
    for (hipEvent_t e : ctx->lane_ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->fit_hist) (void)hipFree(ctx->fit_hist);
    if (ctx->fit) (void)hipFree(ctx->fit);
    if (ctx->fit_counts) (void)hipFree(ctx->fit_counts);
    if (ctx->partials) (void)hipFree(ctx->partials);
    if (ctx->fit_host) (void)hipHostFree(ctx->fit_host);
    if (ctx->scene) (void)hipFree(ctx->scene);
    if (ctx->scene_host) (void)hipHostFree(ctx->scene_host);
    for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

extern "C" clim_status clim_sync(clim_ctx* ctx) {
    if (!ctx) return CLIM_EINVAL;
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CLIM_OK;
}

extern "C" void* clim_stream(clim_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// --------------------------------------------------------------- images

extern "C" clim_status clim_image_alloc(clim_ctx* ctx, int32_t width, int32_t height, int32_t format,
                                        clim_image* out) {
    if (!ctx || !out) return CLIM_EINVAL;
    if (width <= 0 || height <= 0 || clim_format_texel_bytes(format) == 0)
        return ctx_fail(ctx, CLIM_EINVAL, "clim_image_alloc: bad size or format");
    const size_t pitch = pitch_for(width, format);
    void* p = nullptr;
    CLIM_TRY_HIP(ctx, hipMalloc(&p, pitch * (size_t)height));
    out->data = p;
    out->width = width;
    out->height = height;
    out->pitch_bytes = (int32_t)pitch;
    out->format = format;
    return CLIM_OK;
}

extern "C" clim_status clim_image_free(clim_ctx* ctx, clim_image* img) {
    if (!ctx || !img) return CLIM_EINVAL;
    if (img->data) {
        CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        CLIM_TRY_HIP(ctx, hipFree(img->data));
    }
    img->data = nullptr;
    return CLIM_OK;
}

extern "C" clim_status clim_copy_from_host(clim_ctx* ctx, const void* host, int64_t host_pitch, clim_image* dst) {
    if (!ctx || !host || !dst || !dst->data) return CLIM_EINVAL;
    const size_t row = (size_t)dst->width * clim_format_texel_bytes(dst->format);
    if (host_pitch < (int64_t)row) return ctx_fail(ctx, CLIM_EINVAL, "clim_copy_from_host: host pitch too small");
    CLIM_TRY_HIP(ctx, hipMemcpy2DAsync(dst->data, dst->pitch_bytes, host, host_pitch, row, dst->height,
                                       hipMemcpyHostToDevice, ctx->stream));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CLIM_OK;
}

extern "C" clim_status clim_copy_to_host(clim_ctx* ctx, const clim_image* src, void* host, int64_t host_pitch) {
    if (!ctx || !host || !src || !src->data) return CLIM_EINVAL;
    const size_t row = (size_t)src->width * clim_format_texel_bytes(src->format);
    if (host_pitch < (int64_t)row) return ctx_fail(ctx, CLIM_EINVAL, "clim_copy_to_host: host pitch too small");
    CLIM_TRY_HIP(ctx, hipMemcpy2DAsync(host, host_pitch, src->data, src->pitch_bytes, row, src->height,
                                       hipMemcpyDeviceToHost, ctx->stream));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CLIM_OK;
}

// ------------------------------------------------------------ operators

extern "C" clim_status clim_scale_raw_data(clim_ctx* ctx, const clim_image* raw, clim_image* scaled, int32_t bayer,
                                           const float scale_mul[4], float black_level) {
    if (!ctx || !scale_mul || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "scaleRawData: bad argument");
    CLIM_TRY(check_img(ctx, raw, CLIM_FMT_R16U, "raw"));
    CLIM_TRY(check_img(ctx, scaled, CLIM_FMT_R32F, "scaled"));
    CLIM_TRY(check_same(ctx, raw, scaled, "scaleRawData"));
    if ((raw->width & 1) || (raw->height & 1)) return ctx_fail(ctx, CLIM_EINVAL, "scaleRawData: odd mosaic size");
    clim::launch_scale_raw(view(raw), view(scaled), make_bayer(bayer), scale_mul, black_level, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_raw_noise_statistics(clim_ctx* ctx, const clim_image* scaled, int32_t bayer,
                                                 clim_image* mean, clim_image* var) {
    if (!ctx || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "rawNoiseStatistics: bad argument");
    CLIM_TRY(check_img(ctx, scaled, CLIM_FMT_R32F, "scaled"));
    CLIM_TRY(check_img(ctx, mean, CLIM_FMT_RGBA32F, "mean"));
    CLIM_TRY(check_img(ctx, var, CLIM_FMT_RGBA32F, "var"));
    if (scaled->width != 2 * mean->width || scaled->height != 2 * mean->height)
        return ctx_fail(ctx, CLIM_EINVAL, "rawNoiseStatistics: mean must be half the mosaic size");
    CLIM_TRY(check_same(ctx, mean, var, "rawNoiseStatistics"));
    clim::launch_raw_stats(view(scaled), make_bayer(bayer), view(mean), view(var), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

// a scaled FP32 mosaic as the demosaic kernels' input
static clim::MosaicArg scaled_mosaic(const clim_image* scaled, int bayer) {
    clim::MosaicArg m{};
    m.im = view(scaled);
    m.b = make_bayer(bayer);
    m.u16 = false;
    return m;
}

// the statistics kernel of the jobs with the first regression pass
// (raw fits: smul set reads the jobs' src as the uint16 mosaic scaled with
// (smul, black), null as the scaled FP32 mosaic)
clim_status run_stats_fit(clim_ctx* ctx, clim::FitJobs& jobs, bool raw, const float* smul, float black) {
    const bool ok = raw ? clim::launch_raw_stats_fit(jobs, smul, black, ctx->stream)
                        : clim::launch_noise_stats_fit(jobs, 0, jobs.njobs, nullptr, ctx->stream);
    if (!ok) return ctx_fail(ctx, CLIM_EINVAL, "noise statistics: sample list smaller than the image");
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

// a sample-list buffer (an RGBA32F allocation) for the statistics of a w x h
// level (raw: of a w x h mosaic)
clim_status fit_list_alloc(clim_ctx* ctx, OwnedImg& o, bool raw, int w, int h) {
    int aw = 0, ah = 0;
    clim::fit_list_dims(raw, w, h, &aw, &ah);
    return owned_alloc(ctx, o, aw, ah, CLIM_FMT_RGBA32F);
}

// the rest of the regression: finalisation, refit, finalisation
clim_status run_fit_rest(clim_ctx* ctx, clim::FitJobs& jobs, bool raw) {
    clim::launch_fit_rest(jobs, raw, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

clim::FitJob make_fit_job(clim_ctx* ctx, int slot, const Img& src, const Img& m, const Img& v, const float* nlf_in,
                          const float* mult) {
    clim::FitJob jb{};
    jb.src = src;
    jb.m = m;
    jb.v = v;
    jb.st = ctx->fit + slot;
    jb.partials = ctx->partials + (size_t)slot * clim::kFitBlocksPerJob * kFitK;
    jb.counts = ctx->fit_counts + (size_t)slot * clim::kFitBlocksPerJob;
    jb.cap = std::min((long long)m.pitch * m.h, (long long)v.pitch * v.h) / 16;
    for (int k = 0; k < 8; k++) {
        jb.nlf_in[k] = nlf_in ? nlf_in[k] : 0.0f;
        jb.mult[k] = mult ? mult[k] : 1.0f;
    }
    return jb;
}

static clim_status read_fit(clim_ctx* ctx, int slot, int n) {
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(ctx->fit_host + slot, ctx->fit + slot, sizeof(clim::FitState) * n,
                                     hipMemcpyDeviceToHost, ctx->stream));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CLIM_OK;
}

extern "C" clim_status clim_compute_raw_noise_statistics(clim_ctx* ctx, const clim_image* scaled, int32_t bayer,
                                                         float nlf_out[8]) {
    if (!ctx || !nlf_out) return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics: null argument");
    if (bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics: bad bayer");
    CLIM_TRY(check_img(ctx, scaled, CLIM_FMT_R32F, "scaled"));
    if ((scaled->width & 1) || (scaled->height & 1))
        return ctx_fail(ctx, CLIM_EINVAL, "computeRawNoiseStatistics: odd mosaic size");
    // rawNoiseStatistics into the context's mean/var sample lists, then the fit
    CLIM_TRY(fit_list_alloc(ctx, ctx->op_mean, true, scaled->width, scaled->height));
    CLIM_TRY(fit_list_alloc(ctx, ctx->op_var, true, scaled->width, scaled->height));
    clim::FitJobs jobs{};
    jobs.njobs = 1;
    jobs.j[0] = make_fit_job(ctx, 0, view(scaled), ctx->op_mean.img, ctx->op_var.img, nullptr, nullptr);
    jobs.j[0].b = make_bayer(bayer);
    CLIM_TRY(run_stats_fit(ctx, jobs, true, nullptr, 0.0f));
    CLIM_TRY(run_fit_rest(ctx, jobs, true));
    CLIM_TRY(read_fit(ctx, 0, 1));
    if (ctx->fit_host[0].status == 1)
        return ctx_fail(ctx, CLIM_ENUMERIC, "computeRawNoiseStatistics: no sample in range");
    std::memcpy(nlf_out, ctx->fit_host[0].out, 8 * sizeof(float));
    return CLIM_OK;
}

extern "C" clim_status clim_interpolate_green(clim_ctx* ctx, const clim_image* scaled, clim_image* green,
                                              int32_t bayer, float luma_variance) {
    if (!ctx || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "interpolateGreen: bad argument");
    CLIM_TRY(check_img(ctx, scaled, CLIM_FMT_R32F, "scaled"));
    CLIM_TRY(check_img(ctx, green, CLIM_FMT_R32F, "green"));
    CLIM_TRY(check_same(ctx, scaled, green, "interpolateGreen"));
    clim::launch_green(scaled_mosaic(scaled, bayer), view(green), bayer, host_params(&luma_variance, 1),
                       ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_interpolate_red_blue(clim_ctx* ctx, const clim_image* scaled, const clim_image* green,
                                                 clim_image* rgb, int32_t bayer, float chroma_variance,
                                                 int32_t rotate_180) {
    if (!ctx || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "interpolateRedBlue: bad argument");
    CLIM_TRY(check_img(ctx, scaled, CLIM_FMT_R32F, "scaled"));
    CLIM_TRY(check_img(ctx, green, CLIM_FMT_R32F, "green"));
    CLIM_TRY(check_img(ctx, rgb, CLIM_FMT_RGBA32F, "rgb"));
    CLIM_TRY(check_same(ctx, scaled, green, "interpolateRedBlue"));
    CLIM_TRY(check_same(ctx, scaled, rgb, "interpolateRedBlue"));
    clim::launch_red_blue(scaled_mosaic(scaled, bayer), view(green), view(rgb), bayer, make_bayer(bayer),
                          host_params(&chroma_variance, 1), rotate_180 ? 1 : 0, false, 1.0f, nullptr, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_blend_highlights(clim_ctx* ctx, const clim_image* in, float clip, clim_image* out) {
    if (!ctx) return CLIM_EINVAL;
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "blendHighlightsImage"));
    clim::launch_blend_highlights(view(in), view(out), clip, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_transform_image(clim_ctx* ctx, const clim_image* in, clim_image* out,
                                            const float transform[9]) {
    if (!ctx || !transform) return ctx_fail(ctx, CLIM_EINVAL, "transformImage: null transform");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "transformImage"));
    clim::launch_transform(view(in), view(out), transform, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_despeckle_image(clim_ctx* ctx, const clim_image* in, const float var_a[3],
                                            const float var_b[3], clim_image* out) {
    if (!ctx || !var_a || !var_b) return ctx_fail(ctx, CLIM_EINVAL, "despeckleImage: null noise model");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "despeckleImage"));
    if (in->data == out->data) return ctx_fail(ctx, CLIM_EINVAL, "despeckleImage: in and out must differ");
    clim::launch_despeckle(view(in), view(out), var_a[0], var_b[0], ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

// ------------------------------------------------ half-resolution preview

extern "C" clim_status clim_fast_debayer(clim_ctx* ctx, const clim_image* raw, clim_image* rgba, int32_t bayer) {
    if (!ctx || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "fasteDebayer: bad argument");
    CLIM_TRY(check_img(ctx, raw, CLIM_FMT_R32F, "raw"));
    CLIM_TRY(check_img(ctx, rgba, CLIM_FMT_RGBA32F, "rgba"));
    if (raw->width != 2 * rgba->width || raw->height != 2 * rgba->height)
        return ctx_fail(ctx, CLIM_EINVAL, "fasteDebayer: the raw image must be twice the RGBA image's size");
    clim::launch_fast_debayer(view(raw), view(rgba), make_bayer(bayer), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

// ------------------------------------------------ high-noise raw branch

static clim_status check_quads(clim_ctx* ctx, const clim_image* raw, const clim_image* rgba, const char* what) {
    if (raw->width != 2 * rgba->width || raw->height != 2 * rgba->height)
        return ctx_fail(ctx, CLIM_EINVAL, std::string(what) + ": the raw image must be twice the RGBA image's size");
    return CLIM_OK;
}

extern "C" clim_status clim_bayer_to_raw_rgba(clim_ctx* ctx, const clim_image* raw, clim_image* rgba,
                                              int32_t bayer) {
    if (!ctx || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "bayerToRawRGBA: bad argument");
    CLIM_TRY(check_img(ctx, raw, CLIM_FMT_R32F, "raw"));
    CLIM_TRY(check_img(ctx, rgba, CLIM_FMT_RGBA32F, "rgba"));
    CLIM_TRY(check_quads(ctx, raw, rgba, "bayerToRawRGBA"));
    clim::launch_bayer_to_raw_rgba(view(raw), view(rgba), make_bayer(bayer), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_raw_rgba_to_bayer(clim_ctx* ctx, const clim_image* rgba, clim_image* raw,
                                              int32_t bayer) {
    if (!ctx || bayer < 0 || bayer > 3) return ctx_fail(ctx, CLIM_EINVAL, "rawRGBAToBayer: bad argument");
    CLIM_TRY(check_img(ctx, rgba, CLIM_FMT_RGBA32F, "rgba"));
    CLIM_TRY(check_img(ctx, raw, CLIM_FMT_R32F, "raw"));
    CLIM_TRY(check_quads(ctx, raw, rgba, "rawRGBAToBayer"));
    clim::launch_raw_rgba_to_bayer(view(rgba), view(raw), make_bayer(bayer), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_despeckle_raw_rgba(clim_ctx* ctx, const clim_image* in, clim_image* out) {
    if (!ctx) return CLIM_EINVAL;
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "despeckleRawRGBAImage"));
    if (in->data == out->data) return ctx_fail(ctx, CLIM_EINVAL, "despeckleRawRGBAImage: in and out must differ");
    clim::launch_despeckle_raw_rgba(view(in), view(out), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_denoise_raw_rgba(clim_ctx* ctx, const clim_image* in, const float raw_variance[4],
                                             clim_image* out) {
    if (!ctx || !raw_variance) return ctx_fail(ctx, CLIM_EINVAL, "denoiseRawRGBAImage: null argument");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "denoiseRawRGBAImage"));
    if (in->data == out->data) return ctx_fail(ctx, CLIM_EINVAL, "denoiseRawRGBAImage: in and out must differ");
    clim::launch_denoise_raw_rgba(view(in), host_params(raw_variance, 4), view(out), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_downsample_image(clim_ctx* ctx, const clim_image* in, clim_image* out) {
    if (!ctx) return CLIM_EINVAL;
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    clim::launch_downsample(view(in), view(out), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_noise_statistics(clim_ctx* ctx, const clim_image* in, clim_image* out) {
    if (!ctx) return CLIM_EINVAL;
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "noiseStatistics"));
    clim::launch_noise_stats(view(in), view(out), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_compute_noise_statistics(clim_ctx* ctx, const clim_image* in, float nlf_out[6]) {
    if (!ctx || !nlf_out) return ctx_fail(ctx, CLIM_EINVAL, "computeNoiseStatistics: null argument");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(fit_list_alloc(ctx, ctx->op_stats, false, in->width, in->height));
    clim::FitJobs jobs{};
    jobs.njobs = 1;
    jobs.j[0] = make_fit_job(ctx, 1, view(in), ctx->op_stats.img, ctx->op_stats.img, nullptr, nullptr);
    CLIM_TRY(run_stats_fit(ctx, jobs, false, nullptr, 0.0f));
    CLIM_TRY(run_fit_rest(ctx, jobs, false));
    CLIM_TRY(read_fit(ctx, 1, 1));
    if (ctx->fit_host[1].status == 1)
        return ctx_fail(ctx, CLIM_ENUMERIC, "computeNoiseStatistics: no sample in range");
    std::memcpy(nlf_out, ctx->fit_host[1].out, 6 * sizeof(float));
    return CLIM_OK;
}

extern "C" clim_status clim_denoise_image(clim_ctx* ctx, const clim_image* in, const float var_a[3],
                                          const float var_b[3], float chroma_boost, float gradient_boost,
                                          clim_image* out) {
    if (!ctx || !var_a || !var_b) return ctx_fail(ctx, CLIM_EINVAL, "denoiseImage: null noise model");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "denoiseImage"));
    if (in->data == out->data) return ctx_fail(ctx, CLIM_EINVAL, "denoiseImage: in and out must differ");
    const float nlf[6] = {var_a[0], var_a[1], var_a[2], var_b[0], var_b[1], var_b[2]};
    clim::launch_bilateral(view(in), view(out), host_params(nlf, 6), chroma_boost, gradient_boost, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_denoise_image_guided(clim_ctx* ctx, const clim_image* in, const float var_a[3],
                                                 const float var_b[3], clim_image* out) {
    if (!ctx || !var_a || !var_b) return ctx_fail(ctx, CLIM_EINVAL, "denoiseImageGuided: null noise model");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "denoiseImageGuided"));
    if (in->data == out->data) return ctx_fail(ctx, CLIM_EINVAL, "denoiseImageGuided: in and out must differ");
    const float nlf[6] = {var_a[0], var_a[1], var_a[2], var_b[0], var_b[1], var_b[2]};
    clim::launch_guided_fast(view(in), view(out), host_params(nlf, 6), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

clim_status guided_scratch(clim_ctx* ctx, OwnedImg* imgs, int w, int h, clim::GuidedScratch* g) {
    for (int i = 0; i < 7; i++) CLIM_TRY(owned_alloc(ctx, imgs[i], w, h, CLIM_FMT_RGBA32F));
    CLIM_TRY(owned_alloc(ctx, imgs[7], w, h, CLIM_FMT_RG32F));
    g->mean_I = imgs[0].img;
    g->s1 = imgs[1].img;
    g->ipr = imgs[2].img;
    g->ipg = imgs[3].img;
    g->ipb = imgs[4].img;
    g->tmp4 = imgs[5].img;
    g->tmp2 = imgs[6].img;  // used as a float2 image (pitch is large enough)
    g->s2 = imgs[7].img;
    return CLIM_OK;
}

extern "C" clim_status clim_guided_filter(clim_ctx* ctx, const clim_image* in, int32_t filter_size,
                                          const float eps[3], clim_image* out) {
    if (!ctx || !eps || filter_size < 1 || !(filter_size & 1))
        return ctx_fail(ctx, CLIM_EINVAL, "GuidedFilter::filter: bad argument (filter size must be odd)");
    CLIM_TRY(check_img(ctx, in, CLIM_FMT_RGBA32F, "in"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, in, out, "GuidedFilter::filter"));
    // the one-launch form reads neighbouring tiles' input: in place takes the passes
    if (clim::guided_filter_fused(filter_size) && in->data != out->data) {
        clim::launch_guided_filter_fused(view(in), view(out), filter_size, host_params(eps, 3), ctx->stream);
    } else {
        clim::GuidedScratch g{};
        CLIM_TRY(guided_scratch(ctx, ctx->op_guided_imgs, in->width, in->height, &g));
        clim::launch_guided_filter(view(in), view(out), filter_size, host_params(eps, 3), g, ctx->stream);
    }
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_reassemble_image(clim_ctx* ctx, const clim_image* denoised0, const clim_image* image1,
                                             const clim_image* denoised1, float sharpening, const float nlf[2],
                                             clim_image* out) {
    if (!ctx || !nlf) return ctx_fail(ctx, CLIM_EINVAL, "reassembleImage: null nlf");
    CLIM_TRY(check_img(ctx, denoised0, CLIM_FMT_RGBA32F, "denoised0"));
    CLIM_TRY(check_img(ctx, image1, CLIM_FMT_RGBA32F, "image1"));
    CLIM_TRY(check_img(ctx, denoised1, CLIM_FMT_RGBA32F, "denoised1"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_RGBA32F, "out"));
    CLIM_TRY(check_same(ctx, denoised0, out, "reassembleImage"));
    CLIM_TRY(check_same(ctx, image1, denoised1, "reassembleImage"));
    const float p[6] = {nlf[0], 0.0f, 0.0f, nlf[1], 0.0f, 0.0f};
    clim::launch_reassemble(view(denoised0), view(image1), view(denoised1), view(out), sharpening, host_params(p, 6),
                            ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_convert_to_srgb(clim_ctx* ctx, const clim_image* linear, const clim_image* ltm_mask,
                                            clim_image* out_rgba8, clim_image* out_f32, const float transform[9],
                                            const clim_rgb_conversion_params* params) {
    if (!ctx || !transform || !params) return ctx_fail(ctx, CLIM_EINVAL, "convertTosRGB: null argument");
    CLIM_TRY(check_img(ctx, linear, CLIM_FMT_RGBA32F, "linear"));
    CLIM_TRY(check_img(ctx, out_rgba8, CLIM_FMT_RGBA8, "out_rgba8"));
    CLIM_TRY(check_same(ctx, linear, out_rgba8, "convertTosRGB"));
    Img of{nullptr, 0, 0, 0}, mask{nullptr, 0, 0, 0};
    if (out_f32) {
        CLIM_TRY(check_img(ctx, out_f32, CLIM_FMT_RGBA32F, "out_f32"));
        CLIM_TRY(check_same(ctx, linear, out_f32, "convertTosRGB"));
        of = view(out_f32);
    }
    // the mask is read only with localToneMapping set (cl/demosaic.cl:1456);
    // otherwise it may be null or a placeholder of any size
    if (params->localToneMapping) {
        if (!ltm_mask) return ctx_fail(ctx, CLIM_EINVAL, "convertTosRGB: localToneMapping needs ltm_mask");
        CLIM_TRY(check_img(ctx, ltm_mask, CLIM_FMT_R32F, "ltm_mask"));
        CLIM_TRY(check_same(ctx, linear, ltm_mask, "convertTosRGB"));
        mask = view(ltm_mask);
    }
    clim::launch_tone(view(linear), mask, view(out_rgba8), of, nullptr, transform, params->contrast,
                      params->saturation, params->toneCurveSlope, params->exposureBias, params->blacks, ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_local_tone_mapping_mask(clim_ctx* ctx, const clim_image* input, const clim_image* guide,
                                                    const clim_ltm_params* ltm, const float transform[9],
                                                    clim_image* out) {
    if (!ctx || !ltm || !transform) return ctx_fail(ctx, CLIM_EINVAL, "localToneMappingMask: null argument");
    CLIM_TRY(check_img(ctx, input, CLIM_FMT_RGBA32F, "input"));
    CLIM_TRY(check_img(ctx, guide, CLIM_FMT_RGBA32F, "guide"));
    CLIM_TRY(check_img(ctx, out, CLIM_FMT_R32F, "out"));
    CLIM_TRY(check_same(ctx, input, out, "localToneMappingMask"));
    const float p[4] = {ltm->guidedFilterEps, ltm->shadows, ltm->highlights, ltm->detail};
    clim::launch_ltm_mask(view(input), view(guide), p, transform, view(out), ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

// ------------------------------------------------------- host-side math

extern "C" void clim_default_params(clim_demosaic_params* p, int32_t bayer, float black_level, float white_level) {
    std::memset(p, 0, sizeof(*p));
    p->bayerPattern = bayer;
    p->black_level = black_level;
    p->white_level = white_level;
    p->exposure_multiplier = 1.0f;
    const float sm = 65535.0f / (white_level - black_level);
    for (int c = 0; c < 4; c++) p->scale_mul[c] = sm;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p->rgb_cam[i][j] = i == j ? 1.0f : 0.0f;
    // "minimal noise model" of cpp/raw_pipeline.cpp:101-111
    const float raw[4] = {2.8482e-04f, 3.6848e-04f, 3.3079e-04f, 3.2821e-04f};
    const float pyr[5][6] = {{2.1e-04f, 1.0e-08f, 8.2e-06f, 2.6e-04f, 3.9e-05f, 9.0e-06f},
                             {1.6e-04f, 1.0e-08f, 1.6e-06f, 6.1e-04f, 1.8e-04f, 6.1e-05f},
                             {2.5e-04f, 1.0e-08f, 1.0e-08f, 6.0e-04f, 3.6e-04f, 1.3e-04f},
                             {4.2e-04f, 1.0e-08f, 1.0e-08f, 4.2e-04f, 8.8e-04f, 3.4e-04f},
                             {3.0e-04f, 1.0e-08f, 1.0e-08f, 1.0e-03f, 1.8e-03f, 4.2e-04f}};
    std::memcpy(p->noiseModel.rawNlf, raw, sizeof(raw));
    std::memcpy(p->noiseModel.pyramidNlf, pyr, sizeof(pyr));
    for (int i = 0; i < CLIM_MAX_LEVELS; i++) p->denoiseParameters[i] = {1.0f, 1.0f, 1.0f, 1.0f, 1.0f};
    p->noiseLevel = 0.0f;
    p->rgbConversionParameters = {1.05f, 1.0f, 3.5f, 0.0f, 0.0f, 0};
    p->ltmParameters = {0.01f, 1.25f, 1.0f, 1.1f};
    p->levels = 5;
    p->denoiseAlgorithm = CLIM_DENOISE_BILATERAL;
}

// ------------------------------------------------------------ profiling

static const char* kStageNames[CLIM_PROFILE_STAGES] = {
    "scale_raw", "raw_stats", "raw_fit", "green", "red_blue_ycc", "despeckle",
    "downsample", "level_stats", "level_fit", "denoise", "reassemble", "tone", "raw_denoise", "strength"};

static constexpr size_t kProfEvents = 4096;

ProfScope::ProfScope(clim_ctx* c, int stage, long long texels) : ctx(c) {
    if (!ctx->prof || !((ctx->prof_mask >> stage) & 1u)) return;
    // the pool grows while a profile is being collected; events are reused
    // after clim_profile_read
    while (ctx->ev_next + 2 > ctx->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        ctx->ev_pool.push_back(e);
    }
    ProfRec r{stage, texels, ctx->ev_pool[ctx->ev_next], ctx->ev_pool[ctx->ev_next + 1]};
    ctx->ev_next += 2;
    (void)hipEventRecord(r.a, ctx->stream);
    idx = (int)ctx->prof_recs.size();
    ctx->prof_recs.push_back(r);
}

ProfScope::~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(ctx->prof_recs[idx].b, ctx->stream);
}

extern "C" clim_status clim_profile_enable(clim_ctx* ctx, int32_t enable) {
    if (!ctx) return CLIM_EINVAL;
    if (enable && ctx->ev_pool.empty()) {
        ctx->ev_pool.resize(kProfEvents);
        for (auto& e : ctx->ev_pool) CLIM_TRY_HIP(ctx, hipEventCreate(&e));
        ctx->prof_recs.reserve(kProfEvents / 2);
    }
    ctx->prof = enable != 0;
    return CLIM_OK;
}

extern "C" clim_status clim_profile_select(clim_ctx* ctx, uint32_t stage_mask) {
    if (!ctx) return CLIM_EINVAL;
    ctx->prof_mask = stage_mask;
    return CLIM_OK;
}

extern "C" clim_status clim_profile_read(clim_ctx* ctx, clim_profile* out) {
    if (!ctx || !out) return CLIM_EINVAL;
    std::memset(out, 0, sizeof(*out));
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (const ProfRec& r : ctx->prof_recs) {
        float ms = 0.0f;
        CLIM_TRY_HIP(ctx, hipEventElapsedTime(&ms, r.a, r.b));
        out->ms[r.stage] += ms;
        out->launches[r.stage] += 1;
        out->texels[r.stage] += r.texels;
    }
    ctx->prof_recs.clear();
    ctx->ev_next = 0;
    return CLIM_OK;
}

extern "C" const char* clim_profile_stage_name(int32_t stage) {
    return stage >= 0 && stage < CLIM_PROFILE_STAGES ? kStageNames[stage] : "";
}
