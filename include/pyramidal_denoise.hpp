// pyramidal_denoise.hpp - the per-level denoiser plug-in and the pyramid.
//
// ImageDenoiser, PyramidalDenoise<levels>, computeNoiseStatistics and
// computeRawNoiseStatistics keep the interface of cpp/pyramidal_denoise.hpp:24-58.
// With the built-in denoisers the whole pyramid (downsampling, per-level
// statistics and FP64 fits, filtering, reassembly) runs as one device-resident
// sequence (clim_pyramidal_denoise).  A caller-supplied ImageDenoiser is run
// level by level exactly as the reference orchestrates it
// (cpp/pyramidal_denoise.cpp:103-149).
#pragma once

#include <array>
#include <memory>

#include "demosaic.hpp"

struct ImageDenoiser {
    ImageDenoiser(gls::HIPContext*, int, int) {}
    virtual ~ImageDenoiser() {}
    virtual void denoise(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                         const std::array<float, 3>& sigma_a, const std::array<float, 3>& sigma_b, float chromaBoost,
                         float gradientBoost, int pyramidLevel,
                         gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) = 0;
    // clim_denoise_algorithm of a built-in denoiser, -1 for a caller's plug-in
    virtual int builtin() const { return -1; }
};

struct BilateralDenoiser : ImageDenoiser {
    using ImageDenoiser::ImageDenoiser;
    void denoise(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& in,
                 const std::array<float, 3>& var_a, const std::array<float, 3>& var_b, float chromaBoost,
                 float gradientBoost, int, gls::hip_image_2d<gls::rgba_pixel_float>* out) override {
        ::denoiseImage(ctx, in, var_a, var_b, chromaBoost, gradientBoost, out);
    }
    int builtin() const override { return CLIM_DENOISE_BILATERAL; }
};

struct GuidedFastDenoiser : ImageDenoiser {
    using ImageDenoiser::ImageDenoiser;
    void denoise(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& in,
                 const std::array<float, 3>& var_a, const std::array<float, 3>& var_b, float, float, int,
                 gls::hip_image_2d<gls::rgba_pixel_float>* out) override {
        ::denoiseImageGuided(ctx, in, var_a, var_b, out);
    }
    int builtin() const override { return CLIM_DENOISE_GUIDED_FAST; }
};

struct GuidedPreciseDenoiser : ImageDenoiser {
    GuidedFilter guidedFilter;
    GuidedPreciseDenoiser(gls::HIPContext* ctx, int w, int h) : ImageDenoiser(ctx, w, h), guidedFilter(ctx, w, h) {}
    void denoise(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& in,
                 const std::array<float, 3>&, const std::array<float, 3>& var_b, float, float, int,
                 gls::hip_image_2d<gls::rgba_pixel_float>* out) override {
        guidedFilter.filter(ctx, in, /*filterSize=*/5, var_b, out);
    }
    int builtin() const override { return CLIM_DENOISE_GUIDED_PRECISE; }
};

// computeNoiseStatistics, cpp/pyramidal_denoise.cpp:161-256 (blocking)
inline std::array<float, 6> computeNoiseStatistics(gls::HIPContext* ctx,
                                                   const gls::hip_image_2d<gls::rgba_pixel_float>& image) {
    std::array<float, 6> out{};
    ctx->check(clim_compute_noise_statistics(ctx->ctx(), image.clim(), out.data()), "computeNoiseStatistics");
    return out;
}

// computeRawNoiseStatistics, cpp/pyramidal_denoise.cpp:259-362 (blocking)
inline std::array<float, 8> computeRawNoiseStatistics(gls::HIPContext* ctx,
                                                      const gls::hip_image_2d<gls::luma_pixel_float>& rawImage,
                                                      BayerPattern bayerPattern) {
    std::array<float, 8> out{};
    ctx->check(clim_compute_raw_noise_statistics(ctx->ctx(), rawImage.clim(), bayerPattern, out.data()),
               "computeRawNoiseStatistics");
    return out;
}

template <size_t levels>
struct PyramidalDenoise {
    enum DenoiseAlgorithm { Bilateral, GuidedFast, GuidedPrecise };

    typedef gls::hip_image_2d<gls::rgba_pixel_float> imageType;
    std::array<typename imageType::unique_ptr, levels - 1> imagePyramid;
    std::array<typename imageType::unique_ptr, levels> denoisedImagePyramid;
    std::array<std::unique_ptr<ImageDenoiser>, levels> denoiser;

    PyramidalDenoise(gls::HIPContext* ctx, int width, int height, DenoiseAlgorithm algorithm = Bilateral)
        : _algorithm(algorithm) {
        for (size_t i = 0, scale = 2; i < levels - 1; i++, scale *= 2)
            imagePyramid[i] = std::make_unique<imageType>(ctx, width / scale, height / scale);
        for (size_t i = 0, scale = 1; i < levels; i++, scale *= 2) {
            denoisedImagePyramid[i] = std::make_unique<imageType>(ctx, width / scale, height / scale);
            if (algorithm == GuidedPrecise)
                denoiser[i] = std::make_unique<GuidedPreciseDenoiser>(ctx, width / scale, height / scale);
            else if (algorithm == GuidedFast && i < levels - 1)
                denoiser[i] = std::make_unique<GuidedFastDenoiser>(ctx, width / scale, height / scale);
            else
                denoiser[i] = std::make_unique<BilateralDenoiser>(ctx, width / scale, height / scale);
        }
    }

    // nlfParameters: in, the fallback model; out, the measured per-level model
    imageType* denoise(gls::HIPContext* ctx, std::array<DenoiseParameters, levels>* denoiseParameters,
                       imageType* image, const Matrix3x3& /*rgb_cam*/, const gls::rectangle* /*gmb_position*/,
                       bool /*rotate_180*/, std::array<std::array<float, 6>, levels>* nlfParameters) {
        if (builtin_only()) {
            clim_denoise_params dp[5];
            for (size_t i = 0; i < levels; i++) {
                const DenoiseParameters& d = (*denoiseParameters)[i];
                dp[i] = {d.luma, d.chroma, d.chromaBoost, d.gradientBoost, d.sharpening};
            }
            float nlf[5 * 6];
            for (size_t i = 0; i < levels; i++)
                for (int k = 0; k < 6; k++) nlf[i * 6 + k] = (*nlfParameters)[i][k];
            // one device-resident call that leaves every level in this object's images
            clim_image* lvl[levels - 1];
            clim_image* den[levels];
            for (size_t i = 0; i < levels; i++) {
                den[i] = denoisedImagePyramid[i]->clim();
                if (i + 1 < levels) lvl[i] = imagePyramid[i]->clim();
            }
            ctx->check(clim_pyramidal_denoise_into(ctx->ctx(), image->clim(), (int)levels, (int)_algorithm, dp, lvl,
                                                   den, nlf),
                       "PyramidalDenoise::denoise");
            for (size_t i = 0; i < levels; i++)
                for (int k = 0; k < 6; k++) (*nlfParameters)[i][k] = nlf[i * 6 + k];
            return denoisedImagePyramid[0].get();
        }
        // caller plug-ins: the reference's level-by-level orchestration
        std::array<std::array<float, 6>, levels> calibrated;
        for (size_t i = 0; i < levels; i++) {
            if (i < levels - 1) resampleImage(ctx, "downsampleImage", i == 0 ? *image : *imagePyramid[i - 1],
                                              imagePyramid[i].get());
            try {
                (*nlfParameters)[i] = computeNoiseStatistics(ctx, i == 0 ? *image : *imagePyramid[i - 1]);
            } catch (const gls::clim_error& e) {
                // no sample in range: keep the caller's model for this level
                if (e.status != CLIM_ENUMERIC) throw;
            }
            const DenoiseParameters& d = (*denoiseParameters)[i];
            const float lm = d.luma * d.luma, cm = d.chroma * d.chroma;
            const float mul[6] = {lm, cm, cm, lm, cm, cm};
            for (int k = 0; k < 6; k++) calibrated[i][k] = (*nlfParameters)[i][k] * mul[k];
        }
        for (int i = (int)levels - 1; i >= 0; i--) {
            const imageType& in = i == 0 ? *image : *imagePyramid[i - 1];
            const auto& np = calibrated[i];
            const DenoiseParameters& d = (*denoiseParameters)[i];
            denoiser[i]->denoise(ctx, in, {np[0], np[1], np[2]}, {np[3], np[4], np[5]}, d.chromaBoost,
                                 d.gradientBoost, i, denoisedImagePyramid[i].get());
            if (i < (int)levels - 1)
                reassembleImage(ctx, *denoisedImagePyramid[i], *imagePyramid[i], *denoisedImagePyramid[i + 1],
                                d.sharpening, {np[0], np[3]}, denoisedImagePyramid[i].get());
        }
        return denoisedImagePyramid[0].get();
    }

   private:
    DenoiseAlgorithm _algorithm;
    bool builtin_only() const {
        for (size_t i = 0; i < levels; i++) {
            const int b = denoiser[i]->builtin();
            const int want = _algorithm == GuidedPrecise ? CLIM_DENOISE_GUIDED_PRECISE
                             : (_algorithm == GuidedFast && i < levels - 1) ? CLIM_DENOISE_GUIDED_FAST
                                                                            : CLIM_DENOISE_BILATERAL;
            if (b != want) return false;
        }
        return true;
    }
};
