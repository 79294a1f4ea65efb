// pipeline.cpp - RawConverter::demosaicImage and PyramidalDenoise::denoise as
// device-resident launch sequences (C ABI entry points clim_raw_to_rgb,
// clim_pyramidal_denoise and clim_generate_bayer).
//
// Differences from the reference orchestration (cpp/raw_converter.cpp:73-186,
// cpp/pyramidal_denoise.cpp:103-149), none of which changes a value:
//  - the raw and per-level noise-model fits run on the device; kernels that
//    need a fitted value read it from device memory, so the frame has no
//    host round trip (the reference maps six images per frame);
//  - blendHighlightsImage and the camera->YCbCr transformImage run inside
//    the interpolateRedBlue kernel; the YCbCr->camera transformImage and
//    convertTosRGB run inside the kernel that reassembles pyramid level 0
//    (same per-pixel operations, without the level-0 image round trip);
//  - work buffers follow the frame size (the reference allocates once and
//    never checks, cpp/raw_converter.cpp:27).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.h"
#include "pipeline.h"

using clim::Img;
using clim::ParamRef;

clim_status run_stats_fit(clim_ctx* ctx, clim::FitJobs& jobs, bool raw, const float* smul, float black);
clim_status run_fit_rest(clim_ctx* ctx, clim::FitJobs& jobs, bool raw);
clim_status fit_list_alloc(clim_ctx* ctx, OwnedImg& o, bool raw, int w, int h);
clim::FitJob make_fit_job(clim_ctx* ctx, int slot, const Img& src, const Img& m, const Img& v, const float* nlf_in,
                          const float* mult);
clim_status guided_scratch(clim_ctx* ctx, OwnedImg* imgs, int w, int h, clim::GuidedScratch* g);
int inverse3x3_scaled(const float mf[9], double scale, float out[9]);

static ParamRef dev_params(const float* p) {
    ParamRef r{};
    r.dev = p;
    return r;
}

// half_levels: the level images in FP16 (the mixed-precision pipeline; the
// stand-alone pyramid entry points take the caller's FP32 level 0)
static clim_status ensure_workspace(clim_ctx* ctx, int w, int h, int levels, int algorithm,
                                   bool scaled_plane = false, bool half_levels = false) {
    Workspace& ws = ctx->ws;
    const int lf = half_levels ? CLIM_FMT_RGBA16F : CLIM_FMT_RGBA32F;
    if (scaled_plane) CLIM_TRY(owned_alloc(ctx, ws.scaled, w, h, CLIM_FMT_R32F));
    CLIM_TRY(owned_alloc(ctx, ws.green, w, h, CLIM_FMT_R32F));
    CLIM_TRY(owned_alloc(ctx, ws.rgbA, w, h, lf));
    CLIM_TRY(owned_alloc(ctx, ws.rgbB, w, h, lf));
    CLIM_TRY(fit_list_alloc(ctx, ws.raw_mean, true, w, h));
    CLIM_TRY(fit_list_alloc(ctx, ws.raw_var, true, w, h));
    for (int i = 0, scale = 1; i < levels; i++, scale *= 2) {
        const int lw = w / scale, lh = h / scale;
        CLIM_TRY(fit_list_alloc(ctx, ws.stats[i], false, lw, lh));
        if (i > 0) {
            CLIM_TRY(owned_alloc(ctx, ws.pyr[i], lw, lh, lf));
            CLIM_TRY(owned_alloc(ctx, ws.den[i], lw, lh, lf));
        }
    }
    ws.w = w;
    ws.h = h;
    ws.levels = levels;
    ws.algorithm = algorithm;
    return CLIM_OK;
}

// The pyramid of cpp/pyramidal_denoise.cpp:103-149 on device images.
// level0: despeckled YCbCr; out: denoised level 0 (may not alias level0).
// Fit slots 1..levels receive the per-level noise models.
// With `tone`, level 0's reassembly hands its result to the tone map in the
// same kernel (out is then left holding level 0's denoised-only texels).
// ext_lvl / ext_den (nullable): the caller's images for levels 1..levels-1
// (the downsampled and the denoised pyramid) instead of the workspace's.
static clim_status pyramid(clim_ctx* ctx, const Img& level0, const Img& out, int levels, int algorithm,
                           const clim_denoise_params* dp, const float (*nlf_in)[6],
                           const clim::ToneLaunch* tone = nullptr, const Img* ext_lvl = nullptr,
                           const Img* ext_den = nullptr) {
    Workspace& ws = ctx->ws;
    hipStream_t st = ctx->stream;
    Img lvl[CLIM_MAX_LEVELS], den[CLIM_MAX_LEVELS];
    lvl[0] = level0;
    den[0] = out;
    for (int i = 1; i < levels; i++) {
        lvl[i] = ext_lvl ? ext_lvl[i] : ws.pyr[i].img;
        den[i] = ext_den ? ext_den[i] : ws.den[i].img;
    }
    // analysis: downsample and statistics of every level
    clim::FitJobs jobs{};
    for (int i = 0; i < levels; i++) {
        const float lm = dp[i].luma * dp[i].luma, cm = dp[i].chroma * dp[i].chroma;
        const float mult[8] = {lm, cm, cm, lm, cm, cm, 1.0f, 1.0f};
        float in8[8] = {0};
        std::memcpy(in8, nlf_in[i], 6 * sizeof(float));
        jobs.j[i] = make_fit_job(ctx, 1 + i, lvl[i], ws.stats[i].img, ws.stats[i].img, in8, mult);
    }
    jobs.njobs = levels;
    long long nlev = 0;
    for (int i = 0; i < levels; i++) nlev += (long long)lvl[i].w * lvl[i].h;
    // noiseStatistics of every level with the first regression pass, then the
    // refit.  The statistics kernels of levels 0 and 1 write the next level
    // from their tiles when the reduction is exactly 2x (a level is then read
    // once for both); the remaining levels follow their downsampling in one
    // launch.
    constexpr int kFusedLevels = 2;
    bool listed = true;
    int next = 0;  // the first level whose statistics are not launched yet
    while (next < kFusedLevels && next + 1 < levels && clim::downsample_exact2x(lvl[next], lvl[next + 1])) {
        ProfScope ps(ctx, ST_LEVEL_STATS, (long long)lvl[next].w * lvl[next].h);
        listed = listed && clim::launch_noise_stats_fit(jobs, next, next + 1, &lvl[next + 1], st);
        next++;
    }
    long long rest = 0;
    for (int i = next; i < levels; i++) rest += (long long)lvl[i].w * lvl[i].h;
    for (int i = next; i < levels - 1; i++) {
        ProfScope ps(ctx, ST_DOWNSAMPLE, (long long)lvl[i + 1].w * lvl[i + 1].h);
        clim::launch_downsample(lvl[i], lvl[i + 1], st);
    }
    if (next < levels) {
        ProfScope ps(ctx, ST_LEVEL_STATS, rest);
        listed = listed && clim::launch_noise_stats_fit(jobs, next, levels, nullptr, st);
    }
    if (!listed) return ctx_fail(ctx, CLIM_EINVAL, "pyramid: sample list smaller than a level");
    CLIM_TRY_HIP(ctx, hipGetLastError());
    {
        ProfScope ps(ctx, ST_LEVEL_FIT, nlev);
        CLIM_TRY(run_fit_rest(ctx, jobs, false));
    }
    // the learned strength of the level filters (off: sr leaves them plain)
    clim::StrengthRef sr{nullptr, 0, 0, nullptr, 0.0f};
    CLIM_TRY(strength_infer(ctx, lvl, levels, algorithm, &sr));

    // synthesis, coarse to fine.  The bilateral denoiser reads only the noisy
    // levels: level 0 and the coarsest are denoised in one launch ahead of
    // the reassembly chain, and each FP32 level between them is denoised in
    // the launch that reassembles it (its denoised texels stay on chip).
    // FP16 levels, and a level whose next level is not half its size, take
    // the one launch instead.
    bool fused[CLIM_MAX_LEVELS] = {false};
    if (algorithm == CLIM_DENOISE_BILATERAL) {
        clim::BilateralLevels bl{};
        bl.sr = sr;
        long long ntex = 0;
        for (int i = 0; i < levels; i++) {
            fused[i] = i > 0 && i < levels - 1 &&
                       clim::bilateral_reassemble_fits(lvl[i], lvl[i + 1], den[i + 1], den[i]);
            if (fused[i]) continue;
            const int e = bl.n++;
            bl.in[e] = lvl[i];
            bl.out[e] = den[i];
            bl.level[e] = i;
            bl.nlf[e] = dev_params(ctx->fit[1 + i].cal);
            bl.chromaBoost[e] = dp[i].chromaBoost;
            bl.gradientBoost[e] = dp[i].gradientBoost;
            ntex += (long long)lvl[i].w * lvl[i].h;
        }
        ProfScope ps(ctx, ST_DENOISE, ntex);
        clim::launch_bilateral_levels(bl, st);
    }
    for (int i = levels - 1; i >= 0; i--) {
        const ParamRef cal = dev_params(ctx->fit[1 + i].cal);
        if (fused[i]) {
            ProfScope pr(ctx, ST_REASSEMBLE, (long long)lvl[i].w * lvl[i].h);
            clim::launch_bilateral_reassemble(lvl[i], lvl[i + 1], den[i + 1], den[i], cal, dp[i].chromaBoost,
                                              dp[i].gradientBoost, dp[i].sharpening, sr, i, st);
            continue;
        }
        const bool bilateral = algorithm == CLIM_DENOISE_GUIDED_FAST && i == levels - 1;
        if (algorithm != CLIM_DENOISE_BILATERAL) {
            ProfScope ps(ctx, ST_DENOISE, (long long)lvl[i].w * lvl[i].h);
            if (algorithm == CLIM_DENOISE_GUIDED_PRECISE) {
                // GuidedPreciseDenoiser: filter size 5, var_b as eps (cpp/pyramidal_denoise.cpp:65)
                clim::launch_guided_filter_fused(lvl[i], den[i], 5, dev_params(ctx->fit[1 + i].cal + 3), st);
            } else if (bilateral) {
                clim::launch_bilateral(lvl[i], den[i], cal, dp[i].chromaBoost, dp[i].gradientBoost, st, &sr, i);
            } else {
                clim::launch_guided_fast(lvl[i], den[i], cal, st, &sr, i);
            }
        }
        if (i < levels - 1) {
            const bool fused = tone && i == 0;
            ProfScope pr(ctx, fused ? ST_TONE : ST_REASSEMBLE, (long long)lvl[i].w * lvl[i].h);
            clim::launch_reassemble(den[i], lvl[i + 1], den[i + 1], den[i], dp[i].sharpening, cal, st,
                                    fused ? tone : nullptr);
        }
    }
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

static clim_status check_params(clim_ctx* ctx, const clim_demosaic_params* p) {
    if (p->bayerPattern < 0 || p->bayerPattern > 3) return ctx_fail(ctx, CLIM_EINVAL, "bad bayerPattern");
    if (p->levels < 2 || p->levels > CLIM_MAX_LEVELS) return ctx_fail(ctx, CLIM_EINVAL, "levels must be 2..5");
    if (p->denoiseAlgorithm < 0 || p->denoiseAlgorithm > 2) return ctx_fail(ctx, CLIM_EINVAL, "bad denoiseAlgorithm");
    if (p->rgbConversionParameters.localToneMapping && p->levels < 4)
        return ctx_fail(ctx, CLIM_EINVAL, "local tone mapping reads pyramid level 3: levels must be 4 or 5");
    return CLIM_OK;
}

static clim_status check_image(clim_ctx* ctx, const clim_image* im, int format, const char* name) {
    if (!im || !im->data || im->format != format || im->width <= 0 || im->height <= 0 ||
        im->pitch_bytes < im->width * clim_format_texel_bytes(format))
        return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": bad image");
    if (format == CLIM_FMT_RGBA32F && im->pitch_bytes % 16)
        return ctx_fail(ctx, CLIM_EINVAL, std::string(name) + ": pitch must be a multiple of 16 bytes");
    return CLIM_OK;
}

static Img view(const clim_image* im) {
    return Img{(char*)im->data, im->width, im->height, im->pitch_bytes, im->format == CLIM_FMT_RGBA16F ? 1 : 0};
}

extern "C" clim_status clim_pyramidal_denoise(clim_ctx* ctx, const clim_image* image, int32_t levels,
                                              int32_t algorithm, const clim_denoise_params* params, clim_image* out,
                                              float* nlf_out) {
    if (!ctx || !params) return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: null argument");
    if (levels < 2 || levels > CLIM_MAX_LEVELS || algorithm < 0 || algorithm > 2)
        return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: bad levels or algorithm");
    CLIM_TRY(check_image(ctx, image, CLIM_FMT_RGBA32F, "image"));
    CLIM_TRY(check_image(ctx, out, CLIM_FMT_RGBA32F, "out"));
    if (image->width != out->width || image->height != out->height || image->data == out->data)
        return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: out must be a distinct image of the same size");
    CLIM_TRY(ensure_workspace(ctx, image->width, image->height, levels, algorithm));
    float nlf_in[CLIM_MAX_LEVELS][6];
    if (nlf_out) std::memcpy(nlf_in, nlf_out, sizeof(float) * 6 * levels);
    else std::memset(nlf_in, 0, sizeof(nlf_in));
    CLIM_TRY(pyramid(ctx, view(image), view(out), levels, algorithm, params, nlf_in));
    if (nlf_out) {
        CLIM_TRY_HIP(ctx, hipMemcpyAsync(ctx->fit_host + 1, ctx->fit + 1, sizeof(clim::FitState) * levels,
                                         hipMemcpyDeviceToHost, ctx->stream));
        CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < levels; i++) std::memcpy(nlf_out + 6 * i, ctx->fit_host[1 + i].out, 6 * sizeof(float));
    }
    return CLIM_OK;
}

extern "C" clim_status clim_pyramidal_denoise_into(clim_ctx* ctx, const clim_image* image, int32_t levels,
                                                   int32_t algorithm, const clim_denoise_params* params,
                                                   clim_image* const* image_pyramid,
                                                   clim_image* const* denoised_pyramid, float* nlf_out) {
    if (!ctx || !params || !image_pyramid || !denoised_pyramid)
        return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: null argument");
    if (levels < 2 || levels > CLIM_MAX_LEVELS || algorithm < 0 || algorithm > 2)
        return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: bad levels or algorithm");
    CLIM_TRY(check_image(ctx, image, CLIM_FMT_RGBA32F, "image"));
    const int w = image->width, h = image->height;
    if ((w >> (levels - 1)) < 1 || (h >> (levels - 1)) < 1)
        return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: image too small for the pyramid depth");
    Img lvl[CLIM_MAX_LEVELS], den[CLIM_MAX_LEVELS];
    for (int i = 0; i < levels; i++) {
        // level i is width / 2^i x height / 2^i (cpp/pyramidal_denoise.cpp:71-76)
        const clim_image* d = denoised_pyramid[i];
        CLIM_TRY(check_image(ctx, d, CLIM_FMT_RGBA32F, "denoised_pyramid"));
        if (d->width != (w >> i) || d->height != (h >> i) || d->data == image->data)
            return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: denoised level " + std::to_string(i) +
                                                  " must be a distinct image of width/2^i x height/2^i");
        den[i] = view(d);
        if (i > 0) {
            const clim_image* l = image_pyramid[i - 1];
            CLIM_TRY(check_image(ctx, l, CLIM_FMT_RGBA32F, "image_pyramid"));
            if (l->width != (w >> i) || l->height != (h >> i) || l->data == d->data)
                return ctx_fail(ctx, CLIM_EINVAL, "pyramidal denoise: image_pyramid[" + std::to_string(i - 1) +
                                                      "] must be a distinct image of width/2^i x height/2^i");
            lvl[i] = view(l);
        }
    }
    CLIM_TRY(ensure_workspace(ctx, w, h, levels, algorithm));
    float nlf_in[CLIM_MAX_LEVELS][6];
    if (nlf_out) std::memcpy(nlf_in, nlf_out, sizeof(float) * 6 * levels);
    else std::memset(nlf_in, 0, sizeof(nlf_in));
    CLIM_TRY(pyramid(ctx, view(image), den[0], levels, algorithm, params, nlf_in, nullptr, lvl, den));
    if (nlf_out) {
        CLIM_TRY_HIP(ctx, hipMemcpyAsync(ctx->fit_host + 1, ctx->fit + 1, sizeof(clim::FitState) * levels,
                                         hipMemcpyDeviceToHost, ctx->stream));
        CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < levels; i++) std::memcpy(nlf_out + 6 * i, ctx->fit_host[1 + i].out, 6 * sizeof(float));
    }
    return CLIM_OK;
}

clim_status frame_prepare(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8, clim_image* out_f32,
                          const clim_demosaic_params* p, FrameCall* fc) {
    if (!ctx || !p) return ctx_fail(ctx, CLIM_EINVAL, "raw_to_rgb: null argument");
    CLIM_TRY(check_params(ctx, p));
    CLIM_TRY(check_image(ctx, raw, CLIM_FMT_R16U, "raw"));
    CLIM_TRY(check_image(ctx, out_rgba8, CLIM_FMT_RGBA8, "out_rgba8"));
    const int w = raw->width, h = raw->height;
    if ((w & 1) || (h & 1)) return ctx_fail(ctx, CLIM_EINVAL, "raw_to_rgb: mosaic size must be even");
    if ((w >> (p->levels - 1)) < 1 || (h >> (p->levels - 1)) < 1)
        return ctx_fail(ctx, CLIM_EINVAL, "raw_to_rgb: image too small for the pyramid depth");
    if (out_rgba8->width != w || out_rgba8->height != h)
        return ctx_fail(ctx, CLIM_EINVAL, "raw_to_rgb: output size differs from the mosaic");
    FrameCall f{};
    f.p = p;
    f.w = w;
    f.h = h;
    f.out8 = view(out_rgba8);
    f.outf = Img{nullptr, 0, 0, 0};
    if (out_f32) {
        CLIM_TRY(check_image(ctx, out_f32, CLIM_FMT_RGBA32F, "out_f32"));
        if (out_f32->width != w || out_f32->height != h)
            return ctx_fail(ctx, CLIM_EINVAL, "raw_to_rgb: float output size differs from the mosaic");
        f.outf = view(out_f32);
    }
    std::memcpy(f.cam, p->rgb_cam, sizeof(f.cam));
    clim_cam_ycbcr(f.cam, f.c2y);
    if (inverse3x3_scaled(f.c2y, (double)p->exposure_multiplier, f.y2c))
        return ctx_fail(ctx, CLIM_ENUMERIC, "raw_to_rgb: camera->YCbCr matrix is singular");
    // cpp/raw_converter.cpp:82: the raw-domain pre-denoise of high-noise frames
    f.high_noise = p->noiseLevel > 0.6f;
    // scaleRawData (black level normalised as in cpp/raw_converter.cpp:105-106)
    // is applied on the fly by the kernels that read the mosaic
    f.mosaic = clim::MosaicArg{};
    f.mosaic.im = view(raw);
    f.mosaic.b = make_bayer(p->bayerPattern);
    f.mosaic.smul = make_float4(p->scale_mul[0], p->scale_mul[1], p->scale_mul[2], p->scale_mul[3]);
    f.mosaic.black = p->black_level / 65535.0f;
    f.mosaic.u16 = true;
    f.raw = f.mosaic;
    CLIM_TRY(frame_workspace(ctx, f));
    *fc = f;
    return CLIM_OK;
}

clim_status frame_workspace(clim_ctx* ctx, const FrameCall& f) {
    return ensure_workspace(ctx, f.w, f.h, f.p->levels, f.p->denoiseAlgorithm, f.high_noise,
                            ctx->precision == CLIM_PRECISION_MIXED);
}

clim_status frame_raw_fit(clim_ctx* ctx, FrameCall& f, clim::FitState* st_out) {
    Workspace& ws = ctx->ws;
    const clim_demosaic_params* p = f.p;
    const long long npx = (long long)f.w * f.h;
    f.mosaic = f.raw;
    // computeRawNoiseStatistics (cpp/raw_converter.cpp:108-109): rawNlf = B;
    // the input rawNlf is the fallback when no quad is in range
    clim::FitJobs rjobs{};
    {
        const float in8[8] = {0, 0, 0, 0, p->noiseModel.rawNlf[0], p->noiseModel.rawNlf[1], p->noiseModel.rawNlf[2],
                              p->noiseModel.rawNlf[3]};
        rjobs.njobs = 1;
        rjobs.j[0] = make_fit_job(ctx, 0, f.mosaic.im, ws.raw_mean.img, ws.raw_var.img, in8, nullptr);
        rjobs.j[0].b = f.mosaic.b;
        if (st_out) rjobs.j[0].st = st_out;
    }
    f.rawfit = rjobs.j[0].st->out;  // device: A[4], B[4]
    {
        ProfScope ps(ctx, ST_RAW_STATS, npx / 4);
        CLIM_TRY(run_stats_fit(ctx, rjobs, true, p->scale_mul, f.mosaic.black));
    }
    {
        ProfScope ps(ctx, ST_RAW_FIT, npx / 4);
        CLIM_TRY(run_fit_rest(ctx, rjobs, true));
    }
    if (f.high_noise) {
        // bayerToRawRGBA -> despeckleRawRGBAImage -> denoiseRawRGBAImage(rawNlf)
        // -> rawRGBAToBayer (cpp/raw_converter.cpp:111-123) in one kernel; the
        // demosaic then reads the denoised FP32 mosaic
        ProfScope ps(ctx, ST_RAW_DENOISE, npx / 4);
        clim::launch_raw_predenoise(f.mosaic, dev_params(f.rawfit + 4), ws.scaled.img, ctx->stream);
        f.mosaic.im = ws.scaled.img;
        f.mosaic.u16 = false;
    }
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

clim_status frame_demosaic(clim_ctx* ctx, const FrameCall& f, const Img& green, const Img& ycc, int rotate_180) {
    hipStream_t st = ctx->stream;
    const int bayer = f.p->bayerPattern;
    const long long npx = (long long)f.w * f.h;
    // interpolateGreen(rawNlf[1]) and interpolateRedBlue((rawNlf[0]+rawNlf[2])/2)
    // + blendHighlightsImage(clip = 1) + transformImage(cam_ycbcr(rgb_cam))
    {
        ProfScope ps(ctx, ST_GREEN, npx);
        clim::launch_green(f.mosaic, green, bayer, dev_params(f.rawfit + 5), st);
    }
    // despeckleImage with the caller's pyramidNlf[0] (cpp/raw_converter.cpp:144-145),
    // in the same kernel as interpolateRedBlue
    const float* np0 = f.p->noiseModel.pyramidNlf[0];
    if (rotate_180) {
        // cl/demosaic.cl:313-315: source texel (x, y) lands on (w - x, h - y)
        // when that is inside the image, so row 0 and column 0 are never
        // written; here they are zero (the reference leaves whatever its
        // image held).  The despeckle then runs on the rotated image.
        ProfScope ps(ctx, ST_RED_BLUE, npx);
        const Img rot = ctx->ws.rgbA.img;
        clim::launch_zero_row0_col0(rot, st);
        clim::launch_red_blue(f.mosaic, green, rot, bayer, f.mosaic.b, dev_params(f.rawfit), 1, true, 1.0f, f.c2y,
                              st);
        clim::launch_despeckle(rot, ycc, np0[0], np0[3], st);
    } else {
        ProfScope ps(ctx, ST_RED_BLUE, npx);
        clim::launch_rb_despeckle(f.mosaic, green, ycc, bayer, f.mosaic.b, dev_params(f.rawfit), 1.0f, f.c2y, np0[0],
                                  np0[3], st);
    }
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

clim_status frame_pyramid(clim_ctx* ctx, const FrameCall& f, const Img& ycc) {
    Workspace& ws = ctx->ws;
    const clim_demosaic_params* p = f.p;
    // transformImage(inverse(cam_to_ycbcr) * exposure) + convertTosRGB, in the
    // kernel that reassembles pyramid level 0, with localToneMappingMask of the
    // denoised image (guide: denoisedImagePyramid[3], cpp/raw_converter.cpp:164-167)
    const clim_rgb_conversion_params& rc = p->rgbConversionParameters;
    clim::ToneLaunch tone{f.out8,          f.outf,        f.y2c,              f.cam,
                          rc.contrast,     rc.saturation, rc.toneCurveSlope, rc.exposureBias,
                          rc.blacks,       Img{nullptr, 0, 0, 0}, {0, 0, 0, 0}};
    if (rc.localToneMapping) {
        tone.ltm_guide = ws.den[3].img;
        const clim_ltm_params& lp = p->ltmParameters;
        const float l4[4] = {lp.guidedFilterEps, lp.shadows, lp.highlights, lp.detail};
        std::memcpy(tone.ltm, l4, sizeof(l4));
    }
    CLIM_TRY(pyramid(ctx, ycc, ws.rgbA.img, p->levels, p->denoiseAlgorithm, p->denoiseParameters,
                     p->noiseModel.pyramidNlf, &tone));
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

static clim_status raw_to_rgb(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8, clim_image* out_f32,
                              const clim_demosaic_params* p, clim_noise_model* nlf_out, int rotate_180) {
    FrameCall f;
    CLIM_TRY(frame_prepare(ctx, raw, out_rgba8, out_f32, p, &f));
    CLIM_TRY(frame_raw_fit(ctx, f, nullptr));
    CLIM_TRY(frame_demosaic(ctx, f, ctx->ws.green.img, ctx->ws.rgbB.img, rotate_180));
    CLIM_TRY(frame_pyramid(ctx, f, ctx->ws.rgbB.img));
    if (nlf_out) {
        hipStream_t st = ctx->stream;
        CLIM_TRY_HIP(ctx, hipMemcpyAsync(ctx->fit_host, ctx->fit, sizeof(clim::FitState) * (1 + p->levels),
                                         hipMemcpyDeviceToHost, st));
        CLIM_TRY_HIP(ctx, hipStreamSynchronize(st));
        std::memcpy(nlf_out->rawNlf, ctx->fit_host[0].out + 4, 4 * sizeof(float));
        for (int i = 0; i < p->levels; i++)
            std::memcpy(nlf_out->pyramidNlf[i], ctx->fit_host[1 + i].out, 6 * sizeof(float));
    }
    return CLIM_OK;
}

extern "C" clim_status clim_raw_to_rgb(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8,
                                       clim_image* out_f32, const clim_demosaic_params* p, clim_noise_model* nlf_out) {
    return raw_to_rgb(ctx, raw, out_rgba8, out_f32, p, nlf_out, 0);
}

extern "C" clim_status clim_raw_to_rgb_ex(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8,
                                          clim_image* out_f32, const clim_demosaic_params* p,
                                          clim_noise_model* nlf_out, int32_t rotate_180) {
    return raw_to_rgb(ctx, raw, out_rgba8, out_f32, p, nlf_out, rotate_180 ? 1 : 0);
}

extern "C" clim_status clim_fast_demosaic(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8,
                                          clim_image* out_f32, const clim_demosaic_params* p) {
    if (!ctx || !p) return ctx_fail(ctx, CLIM_EINVAL, "fastDemosaicImage: null argument");
    if (p->bayerPattern < 0 || p->bayerPattern > 3) return ctx_fail(ctx, CLIM_EINVAL, "bad bayerPattern");
    CLIM_TRY(check_image(ctx, raw, CLIM_FMT_R16U, "raw"));
    CLIM_TRY(check_image(ctx, out_rgba8, CLIM_FMT_RGBA8, "out_rgba8"));
    const int w = raw->width, h = raw->height;
    if ((w & 1) || (h & 1)) return ctx_fail(ctx, CLIM_EINVAL, "fastDemosaicImage: mosaic size must be even");
    if (out_rgba8->width != w / 2 || out_rgba8->height != h / 2)
        return ctx_fail(ctx, CLIM_EINVAL, "fastDemosaicImage: output must be half the mosaic's size");
    Img outf{nullptr, 0, 0, 0};
    if (out_f32) {
        CLIM_TRY(check_image(ctx, out_f32, CLIM_FMT_RGBA32F, "out_f32"));
        if (out_f32->width != w / 2 || out_f32->height != h / 2)
            return ctx_fail(ctx, CLIM_EINVAL, "fastDemosaicImage: float output must be half the mosaic's size");
        outf = view(out_f32);
    }
    // scaleRawData (black normalised as cpp/raw_converter.cpp:201-202) applied
    // on the fly; localToneMapping is not applied (the reference hands this
    // path a 1x1 placeholder mask, cpp/raw_converter.cpp:69,212)
    clim::MosaicArg mosaic{};
    mosaic.im = view(raw);
    mosaic.b = make_bayer(p->bayerPattern);
    mosaic.smul = make_float4(p->scale_mul[0], p->scale_mul[1], p->scale_mul[2], p->scale_mul[3]);
    mosaic.black = p->black_level / 65535.0f;
    mosaic.u16 = true;
    float cam[9];
    std::memcpy(cam, p->rgb_cam, sizeof(cam));
    const clim_rgb_conversion_params& rc = p->rgbConversionParameters;
    {
        ProfScope ps(ctx, ST_TONE, (long long)(w / 2) * (h / 2));
        clim::launch_fast_preview(mosaic, view(out_rgba8), outf, cam, rc.contrast, rc.saturation, rc.toneCurveSlope,
                                  rc.exposureBias, rc.blacks, ctx->stream);
    }
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

extern "C" clim_status clim_generate_bayer(clim_ctx* ctx, clim_image* raw, clim_image* clean, int32_t bayer,
                                           uint32_t scene_seed, uint32_t noise_seed, float noise_a, float noise_b,
                                           float black_level, float white_level) {
    if (!ctx || bayer < 0 || bayer > 3 || !(noise_a > 0.0f) || noise_b < 0.0f)
        return ctx_fail(ctx, CLIM_EINVAL, "generate_bayer: bad argument");
    if (!raw && !clean) return ctx_fail(ctx, CLIM_EINVAL, "generate_bayer: no output");
    Img r{nullptr, 0, 0, 0}, c{nullptr, 0, 0, 0};
    if (raw) {
        CLIM_TRY(check_image(ctx, raw, CLIM_FMT_R16U, "raw"));
        r = view(raw);
    }
    if (clean) {
        CLIM_TRY(check_image(ctx, clean, CLIM_FMT_RGBA32F, "clean"));
        c = view(clean);
        if (raw && (clean->width != raw->width || clean->height != raw->height))
            return ctx_fail(ctx, CLIM_EINVAL, "generate_bayer: raw and clean sizes differ");
    }
    const int w = raw ? raw->width : clean->width, h = raw ? raw->height : clean->height;
    // the pinned staging buffer is reused: wait for the previous upload
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    clim::scene_draw_params(w, h, scene_seed, ctx->scene_host);
    CLIM_TRY_HIP(ctx, hipMemcpyAsync(ctx->scene, ctx->scene_host, sizeof(clim::SceneParams), hipMemcpyHostToDevice,
                                     ctx->stream));
    clim::launch_scene(r, c, ctx->scene, scene_colors(bayer), noise_seed, noise_a, noise_b, black_level, white_level,
                       ctx->stream);
    CLIM_TRY_HIP(ctx, hipGetLastError());
    return CLIM_OK;
}

// Several frames, each as clim_raw_to_rgb, alternating between the context's
// stream and a second lane's, so that one frame's small launches and launch
// tails overlap the other's kernels.  The second lane starts after the work
// already queued on the context's stream, and the context's stream waits for
// the second lane at the end: later work on the context sees every output.
extern "C" clim_status clim_raw_to_rgb_batch(clim_ctx* ctx, int32_t n, const clim_image* const* raws,
                                             clim_image* const* out_rgba8, clim_image* const* out_f32,
                                             const clim_demosaic_params* params, clim_noise_model* nlf_out) {
    if (!ctx || n < 0 || (n > 0 && (!raws || !out_rgba8 || !params)))
        return ctx_fail(ctx, CLIM_EINVAL, "raw_to_rgb_batch: null argument");
    if (n == 0) return CLIM_OK;
    // lanes in use: one per frame up to kBatchLanes, created on first use
    const int nl = std::min(n, (int)clim_ctx::kBatchLanes);
    for (int k = 1; k < nl; k++) {
        if (ctx->lanes[k - 1]) continue;
        clim_ctx* lane = nullptr;
        const clim_status s = clim_create(&lane, ctx->device);
        if (s != CLIM_OK) return ctx_fail(ctx, s, "raw_to_rgb_batch: lane " + std::to_string(k));
        ctx->lanes[k - 1] = lane;
    }
    for (int k = 0; k < nl && nl > 1; k++)
        if (!ctx->lane_ev[k]) CLIM_TRY_HIP(ctx, hipEventCreateWithFlags(&ctx->lane_ev[k], hipEventDisableTiming));
    int max_levels = 0;
    for (int i = 0; i < n; i++) max_levels = std::max(max_levels, (int)params[i].levels);
    const size_t per = 1 + (size_t)std::min(std::max(max_levels, 1), CLIM_MAX_LEVELS);
    if (nlf_out && ctx->fit_hist_n < per * n) {
        if (ctx->fit_hist) (void)hipFree(ctx->fit_hist);
        ctx->fit_hist = nullptr;
        ctx->fit_hist_n = 0;
        CLIM_TRY_HIP(ctx, hipMalloc(&ctx->fit_hist, sizeof(clim::FitState) * per * n));
        ctx->fit_hist_n = per * n;
    }
    if (nl > 1) {
        CLIM_TRY_HIP(ctx, hipEventRecord(ctx->lane_ev[0], ctx->stream));
        for (int k = 1; k < nl; k++)
            CLIM_TRY_HIP(ctx, hipStreamWaitEvent(ctx->lanes[k - 1]->stream, ctx->lane_ev[0], 0));
    }
    // the lanes run the learned strength as this context does, on its
    // weights, and at its precision
    for (int k = 1; k < nl; k++) {
        ctx->lanes[k - 1]->sn.mode = ctx->sn.mode;
        ctx->lanes[k - 1]->sn.use = ctx->sn.use;
        ctx->lanes[k - 1]->precision = ctx->precision;
    }
    clim_status st = CLIM_OK;
    for (int i = 0; i < n && st == CLIM_OK; i++) {
        const int k = i % nl;
        clim_ctx* c = k ? ctx->lanes[k - 1] : ctx;
        st = clim_raw_to_rgb(c, raws[i], out_rgba8[i], out_f32 ? out_f32[i] : nullptr, &params[i], nullptr);
        if (st != CLIM_OK) {
            const std::string msg = c->last_error;
            ctx_fail(ctx, st, "raw_to_rgb_batch frame " + std::to_string(i) + ": " + msg);
            break;
        }
        if (nlf_out) {
            const hipError_t e = hipMemcpyAsync(ctx->fit_hist + per * i, c->fit,
                                                sizeof(clim::FitState) * (1 + params[i].levels),
                                                hipMemcpyDeviceToDevice, c->stream);
            if (e != hipSuccess) st = ctx_hip(ctx, e, "raw_to_rgb_batch: fit copy");
        }
    }
    for (int k = 1; k < nl; k++) {
        CLIM_TRY_HIP(ctx, hipEventRecord(ctx->lane_ev[k], ctx->lanes[k - 1]->stream));
        CLIM_TRY_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->lane_ev[k], 0));
    }
    if (st != CLIM_OK || !nlf_out) return st;
    CLIM_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    std::vector<clim::FitState> fits(per * n);
    CLIM_TRY_HIP(ctx, hipMemcpy(fits.data(), ctx->fit_hist, sizeof(clim::FitState) * per * n,
                                hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        std::memcpy(nlf_out[i].rawNlf, fits[per * i].out + 4, 4 * sizeof(float));
        for (int l = 0; l < params[i].levels; l++)
            std::memcpy(nlf_out[i].pyramidNlf[l], fits[per * i + 1 + l].out, 6 * sizeof(float));
    }
    return CLIM_OK;
}

