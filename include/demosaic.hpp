// demosaic.hpp - parameter types and the operator API of the RAW->RGB path.
//
// Types follow cpp/demosaic.hpp:23-88 field for field (same defaults); the
// operator free functions follow cpp/demosaic_cl.hpp:23-129 (same names,
// argument order and meaning) with gls::HIPContext / gls::hip_image_2d in
// place of gls::OpenCLContext / gls::cl_image_2d.  Each call is asynchronous
// on the context's stream; failures throw gls::clim_error.
#pragma once

#include <array>
#include <cstring>
#include <string>

#include "climage.h"
#include "gls_hip_image.hpp"

enum BayerPattern { grbg = 0, gbrg = 1, rggb = 2, bggr = 3 };

struct DenoiseParameters {
    float luma = 1.0;
    float chroma = 1.0;
    float chromaBoost = 1.0;
    float gradientBoost = 1.0;
    float sharpening = 1.0;
};

struct RGBConversionParameters {
    float contrast = 1.05;
    float saturation = 1.0;
    float toneCurveSlope = 3.5;
    float exposureBias = 0;
    float blacks = 0;
    int localToneMapping = 0;
};

struct NoiseModel {
    std::array<float, 4> rawNlf{};                  // NLF for raw data
    std::array<std::array<float, 6>, 5> pyramidNlf{};  // per level: A_Y, A_Cb, A_Cr, B_Y, B_Cb, B_Cr
};

struct LTMParameters {
    float guidedFilterEps = 0.01;
    float shadows = 1.25;
    float highlights = 1.0;
    float detail = 1.1;
};

struct DemosaicParameters {
    BayerPattern bayerPattern;
    float black_level;
    float white_level;
    float exposure_multiplier;
    std::array<float, 4> scale_mul;
    std::array<std::array<float, 3>, 3> rgb_cam;
    NoiseModel noiseModel;
    std::array<DenoiseParameters, 5> denoiseParameters;
    float noiseLevel;
    RGBConversionParameters rgbConversionParameters;
    LTMParameters ltmParameters;
};

// The OpenCL table (cl/demosaic.cl:29-34), (x, y) of raw_red, raw_green,
// raw_blue, raw_green2.  The reference's C++ table (cpp/demosaic.hpp:83-88)
// swaps the two greens for rggb/bggr; the GPU path uses this one.
constexpr gls::point bayerOffsets[4][4] = {
    {{1, 0}, {0, 0}, {0, 1}, {1, 1}},  // grbg
    {{0, 1}, {0, 0}, {1, 0}, {1, 1}},  // gbrg
    {{0, 0}, {0, 1}, {1, 1}, {1, 0}},  // rggb
    {{1, 1}, {0, 1}, {0, 0}, {1, 0}}   // bggr
};

// Matrix<3,3> helpers (row-major float[9])
typedef std::array<std::array<float, 3>, 3> Matrix3x3;

inline std::array<float, 9> flat(const Matrix3x3& m) {
    return {m[0][0], m[0][1], m[0][2], m[1][0], m[1][1], m[1][2], m[2][0], m[2][1], m[2][2]};
}

// cam_ycbcr, cpp/demosaic_utils.cpp:809-822 (FP64 inside, float result)
inline Matrix3x3 cam_ycbcr(const Matrix3x3& rgb_cam) {
    std::array<float, 9> in = flat(rgb_cam), out{};
    clim_cam_ycbcr(in.data(), out.data());
    return {{{out[0], out[1], out[2]}, {out[3], out[4], out[5]}, {out[6], out[7], out[8]}}};
}

// inverse, src/gls_linalg.hpp:544-557 (FP64 inside); throws like the reference
// (std::range_error("null determinant")) on a singular matrix
inline Matrix3x3 inverse(const Matrix3x3& m) {
    std::array<float, 9> in = flat(m), out{};
    if (clim_inverse3x3(in.data(), out.data()) != CLIM_OK) throw std::range_error("null determinant");
    return {{{out[0], out[1], out[2]}, {out[3], out[4], out[5]}, {out[6], out[7], out[8]}}};
}

inline clim_demosaic_params to_clim(const DemosaicParameters& p, int levels = 5, int algorithm = 0) {
    clim_demosaic_params c;
    std::memset(&c, 0, sizeof(c));
    c.bayerPattern = p.bayerPattern;
    c.black_level = p.black_level;
    c.white_level = p.white_level;
    c.exposure_multiplier = p.exposure_multiplier;
    for (int i = 0; i < 4; i++) c.scale_mul[i] = p.scale_mul[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c.rgb_cam[i][j] = p.rgb_cam[i][j];
    for (int i = 0; i < 4; i++) c.noiseModel.rawNlf[i] = p.noiseModel.rawNlf[i];
    for (int l = 0; l < 5; l++)
        for (int k = 0; k < 6; k++) c.noiseModel.pyramidNlf[l][k] = p.noiseModel.pyramidNlf[l][k];
    for (int l = 0; l < 5; l++) {
        const DenoiseParameters& d = p.denoiseParameters[l];
        c.denoiseParameters[l] = {d.luma, d.chroma, d.chromaBoost, d.gradientBoost, d.sharpening};
    }
    c.noiseLevel = p.noiseLevel;
    const RGBConversionParameters& r = p.rgbConversionParameters;
    c.rgbConversionParameters = {r.contrast, r.saturation, r.toneCurveSlope, r.exposureBias, r.blacks,
                                 r.localToneMapping};
    c.ltmParameters = {p.ltmParameters.guidedFilterEps, p.ltmParameters.shadows, p.ltmParameters.highlights,
                       p.ltmParameters.detail};
    c.levels = levels;
    c.denoiseAlgorithm = algorithm;
    return c;
}

// ------------------------------------------------------------ operators
// (cpp/demosaic_cl.hpp:28-129)

inline void scaleRawData(gls::HIPContext* ctx, const gls::hip_image_2d<gls::luma_pixel_16>& rawImage,
                         gls::hip_image_2d<gls::luma_pixel_float>* scaledRawImage, BayerPattern bayerPattern,
                         std::array<float, 4> scaleMul, float blackLevel) {
    ctx->check(clim_scale_raw_data(ctx->ctx(), rawImage.clim(), scaledRawImage->clim(), bayerPattern, scaleMul.data(),
                                   blackLevel),
               "scaleRawData");
}

inline void interpolateGreen(gls::HIPContext* ctx, const gls::hip_image_2d<gls::luma_pixel_float>& rawImage,
                             gls::hip_image_2d<gls::luma_pixel_float>* greenImage, BayerPattern bayerPattern,
                             float lumaVariance) {
    ctx->check(clim_interpolate_green(ctx->ctx(), rawImage.clim(), greenImage->clim(), bayerPattern, lumaVariance),
               "interpolateGreen");
}

inline void interpolateRedBlue(gls::HIPContext* ctx, const gls::hip_image_2d<gls::luma_pixel_float>& rawImage,
                               const gls::hip_image_2d<gls::luma_pixel_float>& greenImage,
                               gls::hip_image_2d<gls::rgba_pixel_float>* rgbImage, BayerPattern bayerPattern,
                               float chromaVariance, bool rotate_180) {
    ctx->check(clim_interpolate_red_blue(ctx->ctx(), rawImage.clim(), greenImage.clim(), rgbImage->clim(),
                                         bayerPattern, chromaVariance, rotate_180 ? 1 : 0),
               "interpolateRedBlue");
}

inline void rawNoiseStatistics(gls::HIPContext* ctx, const gls::hip_image_2d<gls::luma_pixel_float>& rawImage,
                               BayerPattern bayerPattern, gls::hip_image_2d<gls::rgba_pixel_float>* meanImage,
                               gls::hip_image_2d<gls::rgba_pixel_float>* varImage) {
    ctx->check(clim_raw_noise_statistics(ctx->ctx(), rawImage.clim(), bayerPattern, meanImage->clim(),
                                         varImage->clim()),
               "rawNoiseStatistics");
}

// applyKernel(ctx, "noiseStatistics", in, out) (cpp/demosaic_cl.cpp:130-145)
inline void applyKernel(gls::HIPContext* ctx, const std::string& kernelName,
                        const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                        gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    if (kernelName != "noiseStatistics") throw gls::clim_error(CLIM_EINVAL, "applyKernel: unsupported " + kernelName);
    ctx->check(clim_noise_statistics(ctx->ctx(), inputImage.clim(), outputImage->clim()), "noiseStatistics");
}

// resampleImage(ctx, "downsampleImage", in, out) (cpp/demosaic_cl.cpp:157-178)
inline void resampleImage(gls::HIPContext* ctx, const std::string& kernelName,
                          const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                          gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    if (kernelName != "downsampleImage") throw gls::clim_error(CLIM_EINVAL, "resampleImage: unsupported " + kernelName);
    ctx->check(clim_downsample_image(ctx->ctx(), inputImage.clim(), outputImage->clim()), "downsampleImage");
}

inline void reassembleImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImageDenoised0,
                            const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage1,
                            const gls::hip_image_2d<gls::rgba_pixel_float>& inputImageDenoised1, float sharpening,
                            std::array<float, 2> nlf, gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_reassemble_image(ctx->ctx(), inputImageDenoised0.clim(), inputImage1.clim(),
                                     inputImageDenoised1.clim(), sharpening, nlf.data(), outputImage->clim()),
               "reassembleImage");
}

inline void transformImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& linearImage,
                           gls::hip_image_2d<gls::rgba_pixel_float>* rgbImage, const Matrix3x3& transform) {
    auto m = flat(transform);
    ctx->check(clim_transform_image(ctx->ctx(), linearImage.clim(), rgbImage->clim(), m.data()), "transformImage");
}

// cpp/demosaic_cl.hpp:69-73.  ltmMaskImage is read only with
// rgbConversionParameters.localToneMapping set; otherwise a 1x1 placeholder
// will do, as in cpp/raw_converter.cpp:36.
inline void convertTosRGB(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& linearImage,
                          const gls::hip_image_2d<gls::luma_pixel_float>& ltmMaskImage,
                          gls::hip_image_2d<gls::rgba_pixel>* rgbImage, const DemosaicParameters& demosaicParameters) {
    auto m = flat(demosaicParameters.rgb_cam);
    const RGBConversionParameters& r = demosaicParameters.rgbConversionParameters;
    clim_rgb_conversion_params p{r.contrast, r.saturation, r.toneCurveSlope, r.exposureBias, r.blacks,
                                 r.localToneMapping};
    ctx->check(clim_convert_to_srgb(ctx->ctx(), linearImage.clim(), ltmMaskImage.clim(), rgbImage->clim(), nullptr,
                                    m.data(), &p),
               "convertTosRGB");
}

// cpp/demosaic_cl.hpp:44-47: the half-resolution (R, (G+G2)/2, B) debayer
inline void fasteDebayer(gls::HIPContext* ctx, const gls::hip_image_2d<gls::luma_pixel_float>& rawImage,
                         gls::hip_image_2d<gls::rgba_pixel_float>* rgbImage, BayerPattern bayerPattern) {
    ctx->check(clim_fast_debayer(ctx->ctx(), rawImage.clim(), rgbImage->clim(), bayerPattern), "fasteDebayer");
}

// The high-noise raw branch (cpp/demosaic_cl.hpp:102-119): an RGBA raw image
// holds one (R, G, B, G2) quad per texel, half the mosaic's size.
inline void bayerToRawRGBA(gls::HIPContext* ctx, const gls::hip_image_2d<gls::luma_pixel_float>& rawImage,
                           gls::hip_image_2d<gls::rgba_pixel_float>* rgbaImage, BayerPattern bayerPattern) {
    ctx->check(clim_bayer_to_raw_rgba(ctx->ctx(), rawImage.clim(), rgbaImage->clim(), bayerPattern), "bayerToRawRGBA");
}

inline void rawRGBAToBayer(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& rgbaImage,
                           gls::hip_image_2d<gls::luma_pixel_float>* rawImage, BayerPattern bayerPattern) {
    ctx->check(clim_raw_rgba_to_bayer(ctx->ctx(), rgbaImage.clim(), rawImage->clim(), bayerPattern), "rawRGBAToBayer");
}

inline void denoiseRawRGBAImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                                const std::array<float, 4> rawVariance,
                                gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_denoise_raw_rgba(ctx->ctx(), inputImage.clim(), rawVariance.data(), outputImage->clim()),
               "denoiseRawRGBAImage");
}

inline void despeckleRawRGBAImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                                  gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_despeckle_raw_rgba(ctx->ctx(), inputImage.clim(), outputImage->clim()), "despeckleRawRGBAImage");
}

// cpp/demosaic_cl.hpp:91-95.  inputImage: the denoised YCbCr image; guideImage:
// pyramid level 3 in cpp/raw_converter.cpp:165; transform: YCbCr -> camera RGB.
inline void localToneMappingMask(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                                 const gls::hip_image_2d<gls::rgba_pixel_float>& guideImage,
                                 const LTMParameters& ltmParameters, const Matrix3x3& transform,
                                 gls::hip_image_2d<gls::luma_pixel_float>* outputImage) {
    auto m = flat(transform);
    const clim_ltm_params lp{ltmParameters.guidedFilterEps, ltmParameters.shadows, ltmParameters.highlights,
                             ltmParameters.detail};
    ctx->check(clim_local_tone_mapping_mask(ctx->ctx(), inputImage.clim(), guideImage.clim(), &lp, m.data(),
                                            outputImage->clim()),
               "localToneMappingMask");
}

inline void despeckleImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                           const std::array<float, 3>& var_a, const std::array<float, 3>& var_b,
                           gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_despeckle_image(ctx->ctx(), inputImage.clim(), var_a.data(), var_b.data(), outputImage->clim()),
               "despeckleImage");
}

inline void denoiseImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                         const std::array<float, 3>& var_a, const std::array<float, 3>& var_b, float chromaBoost,
                         float gradientBoost, gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_denoise_image(ctx->ctx(), inputImage.clim(), var_a.data(), var_b.data(), chromaBoost,
                                  gradientBoost, outputImage->clim()),
               "denoiseImage");
}

inline void denoiseImageGuided(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                               const std::array<float, 3>& var_a, const std::array<float, 3>& var_b,
                               gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_denoise_image_guided(ctx->ctx(), inputImage.clim(), var_a.data(), var_b.data(),
                                         outputImage->clim()),
               "denoiseImageGuided");
}

inline void blendHighlightsImage(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage,
                                 float clip, gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
    ctx->check(clim_blend_highlights(ctx->ctx(), inputImage.clim(), clip, outputImage->clim()),
               "blendHighlightsImage");
}

// GuidedFilter::filter (cpp/guided_filter.hpp:35-55); scratch lives in the context
struct GuidedFilter {
    GuidedFilter(gls::HIPContext*, int, int) {}
    void filter(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& inputImage, int filterSize,
                const std::array<float, 3>& eps, gls::hip_image_2d<gls::rgba_pixel_float>* outputImage) {
        ctx->check(clim_guided_filter(ctx->ctx(), inputImage.clim(), filterSize, eps.data(), outputImage->clim()),
                   "GuidedFilter::filter");
    }
};

// DNG ingest (read_dng_file, unpackDNGMetadata)
#include "gls_dng.hpp"
