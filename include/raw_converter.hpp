// raw_converter.hpp - RawConverter (cpp/raw_converter.hpp:22-62) on gfx950.
//
// demosaicImage uploads the mosaic (blocking, like copyPixelsFrom at
// cpp/raw_converter.cpp:101), runs the whole path on the device and writes
// the measured noise model back into demosaicParameters->noiseModel, as the
// reference does.  The returned image is owned by the converter and valid
// until the next call.  Work buffers follow the frame size.
#pragma once

#include <memory>
#include <vector>

#include "pyramidal_denoise.hpp"

class RawConverter {
   public:
    explicit RawConverter(gls::HIPContext* ctx, int levels = 5,
                          PyramidalDenoise<5>::DenoiseAlgorithm algorithm = PyramidalDenoise<5>::Bilateral)
        : _ctx(ctx), _levels(levels), _algorithm(algorithm) {}

    // gmb_position is accepted and ignored, as in the reference: its only
    // consumer, PyramidalDenoise::denoise, calibrates the noise model on
    // every frame whatever the chart position (cpp/pyramidal_denoise.cpp:109,
    // "calibrate_nlf = true; // gmb_position != nullptr").  rotate_180 turns
    // the image by 180 degrees at interpolateRedBlue with the reference's
    // (width - x, height - y) target (cl/demosaic.cl:313-315); row 0 and
    // column 0, which no texel reaches, are zero (clim_raw_to_rgb_ex).
    gls::hip_image_2d<gls::rgba_pixel>* demosaicImage(const gls::image<gls::luma_pixel_16>& rawImage,
                                                      DemosaicParameters* demosaicParameters,
                                                      const gls::rectangle* /*gmb_position*/, bool rotate_180) {
        allocateTextures(rawImage.width, rawImage.height);
        _raw->copyPixelsFrom(rawImage);
        return demosaicImage(*_raw, demosaicParameters, rotate_180);
    }

    // device-resident input: no upload
    gls::hip_image_2d<gls::rgba_pixel>* demosaicImage(const gls::hip_image_2d<gls::luma_pixel_16>& raw,
                                                      DemosaicParameters* demosaicParameters,
                                                      bool rotate_180 = false) {
        allocateTextures(raw.width, raw.height);
        clim_demosaic_params p = to_clim(*demosaicParameters, _levels, _algorithm);
        clim_noise_model nm;
        _ctx->check(clim_raw_to_rgb_ex(_ctx->ctx(), raw.clim(), _srgb->clim(), nullptr, &p, &nm, rotate_180 ? 1 : 0),
                    "demosaicImage");
        for (int c = 0; c < 4; c++) demosaicParameters->noiseModel.rawNlf[c] = nm.rawNlf[c];
        for (int l = 0; l < _levels; l++)
            for (int k = 0; k < 6; k++) demosaicParameters->noiseModel.pyramidNlf[l][k] = nm.pyramidNlf[l][k];
        return _srgb.get();
    }

    // Several device-resident frames in one call (clim_raw_to_rgb_batch):
    // frame i from raws[i] into outs[i] (the caller's images, of the mosaic's
    // size); demosaicParameters[i]->noiseModel receives frame i's measured
    // noise model.  The frames are spread over three streams of the context.
    void demosaicImages(const std::vector<const gls::hip_image_2d<gls::luma_pixel_16>*>& raws,
                        const std::vector<gls::hip_image_2d<gls::rgba_pixel>*>& outs,
                        const std::vector<DemosaicParameters*>& demosaicParameters) {
        const size_t n = raws.size();
        if (outs.size() != n || demosaicParameters.size() != n)
            throw gls::clim_error(CLIM_EINVAL, "demosaicImages: lists of different lengths");
        std::vector<const clim_image*> r(n);
        std::vector<clim_image*> o(n);
        std::vector<clim_demosaic_params> p(n);
        std::vector<clim_noise_model> nm(n);
        for (size_t i = 0; i < n; i++) {
            r[i] = raws[i]->clim();
            o[i] = outs[i]->clim();
            p[i] = to_clim(*demosaicParameters[i], _levels, _algorithm);
        }
        _ctx->check(clim_raw_to_rgb_batch(_ctx->ctx(), (int32_t)n, r.data(), o.data(), nullptr, p.data(), nm.data()),
                    "demosaicImages");
        for (size_t i = 0; i < n; i++) {
            auto& model = demosaicParameters[i]->noiseModel;
            for (int c = 0; c < 4; c++) model.rawNlf[c] = nm[i].rawNlf[c];
            for (int l = 0; l < _levels; l++)
                for (int k = 0; k < 6; k++) model.pyramidNlf[l][k] = nm[i].pyramidNlf[l][k];
        }
    }

    // cpp/raw_converter.cpp:188-224: the half-resolution preview (one kernel);
    // the returned image is owned by the converter and valid until the next call
    gls::hip_image_2d<gls::rgba_pixel>* fastDemosaicImage(const gls::image<gls::luma_pixel_16>& rawImage,
                                                          const DemosaicParameters& demosaicParameters) {
        if (!_raw || _raw->width != rawImage.width || _raw->height != rawImage.height)
            _raw = std::make_unique<gls::hip_image_2d<gls::luma_pixel_16>>(_ctx, rawImage.width, rawImage.height);
        if (!_fast || _fast->width != rawImage.width / 2 || _fast->height != rawImage.height / 2)
            _fast = std::make_unique<gls::hip_image_2d<gls::rgba_pixel>>(_ctx, rawImage.width / 2, rawImage.height / 2);
        _raw->copyPixelsFrom(rawImage);
        clim_demosaic_params p = to_clim(demosaicParameters, _levels, _algorithm);
        _ctx->check(clim_fast_demosaic(_ctx->ctx(), _raw->clim(), _fast->clim(), nullptr, &p), "fastDemosaicImage");
        return _fast.get();
    }

    // cpp/raw_converter.cpp:226-237
    static gls::image<gls::rgb_pixel>::unique_ptr convertToRGBImage(
        const gls::hip_image_2d<gls::rgba_pixel>& rgbaImage) {
        auto rgba = rgbaImage.mapImage();
        auto rgb = std::make_unique<gls::image<gls::rgb_pixel>>(rgbaImage.width, rgbaImage.height);
        for (int y = 0; y < rgbaImage.height; y++)
            for (int x = 0; x < rgbaImage.width; x++) {
                const auto& p = rgba[y][x];
                (*rgb)[y][x] = {p.red, p.green, p.blue};
            }
        return rgb;
    }

   private:
    void allocateTextures(int width, int height) {
        if (!_raw || _raw->width != width || _raw->height != height)
            _raw = std::make_unique<gls::hip_image_2d<gls::luma_pixel_16>>(_ctx, width, height);
        if (!_srgb || _srgb->width != width || _srgb->height != height)
            _srgb = std::make_unique<gls::hip_image_2d<gls::rgba_pixel>>(_ctx, width, height);
    }

    gls::HIPContext* _ctx;
    int _levels;
    int _algorithm;
    std::unique_ptr<gls::hip_image_2d<gls::luma_pixel_16>> _raw;
    std::unique_ptr<gls::hip_image_2d<gls::rgba_pixel>> _srgb;
    std::unique_ptr<gls::hip_image_2d<gls::rgba_pixel>> _fast;
};
