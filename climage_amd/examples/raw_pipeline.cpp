// raw_pipeline.cpp - drives the C++ operator API the way the reference's
// RawPipeline does (cpp/raw_pipeline.cpp:216-256), on a seeded synthetic
// mosaic instead of a DNG file.
//
//   raw_pipeline W H BAYER LEVELS OUT_PREFIX [LTM [NOISE_LEVEL]]
//   raw_pipeline --dng FILE OUT_PREFIX [AWB]        (or --dng-params: no GPU)
//
// LTM = 1 turns on local tone mapping (rgbConversionParameters.localToneMapping);
// NOISE_LEVEL > 0.6 the raw-domain pre-denoise of high-noise frames.
// Writes OUT_PREFIX.fused.rgb (RawConverter::demosaicImage + convertToRGBImage),
// OUT_PREFIX.stages.rgb (the same path spelled out operator by operator, with a
// caller-supplied ImageDenoiser so that PyramidalDenoise runs level by level
// as in cpp/pyramidal_denoise.cpp:103-149) and OUT_PREFIX.nlf (noise models).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>

#include "raw_converter.hpp"

// A caller plug-in: the built-in hard sigma filter, but not recognised as
// built-in, which forces the level-by-level orchestration.
struct MyBilateral : ImageDenoiser {
    using ImageDenoiser::ImageDenoiser;
    void denoise(gls::HIPContext* ctx, const gls::hip_image_2d<gls::rgba_pixel_float>& in,
                 const std::array<float, 3>& a, const std::array<float, 3>& b, float cb, float gb, int,
                 gls::hip_image_2d<gls::rgba_pixel_float>* out) override {
        ::denoiseImage(ctx, in, a, b, cb, gb, out);
    }
};

static DemosaicParameters synthetic_params(BayerPattern bayer, float black, float white) {
    clim_demosaic_params c;
    clim_default_params(&c, bayer, black, white);
    DemosaicParameters p{};
    p.bayerPattern = bayer;
    p.black_level = black;
    p.white_level = white;
    p.exposure_multiplier = 1;
    for (int i = 0; i < 4; i++) p.scale_mul[i] = c.scale_mul[i];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) p.rgb_cam[i][j] = c.rgb_cam[i][j];
    for (int i = 0; i < 4; i++) p.noiseModel.rawNlf[i] = c.noiseModel.rawNlf[i];
    for (int l = 0; l < 5; l++)
        for (int k = 0; k < 6; k++) p.noiseModel.pyramidNlf[l][k] = c.noiseModel.pyramidNlf[l][k];
    p.noiseLevel = 0;
    return p;
}

static void write_rgb(const std::string& path, const gls::image<gls::rgb_pixel>& im) {
    std::ofstream f(path, std::ios::binary);
    for (int y = 0; y < im.height; y++) f.write(reinterpret_cast<const char*>(im[y]), im.width * 3);
}

// A DNG file the way the reference's RawPipeline takes one
// (cpp/raw_pipeline.cpp:94-104): read_dng_file, unpackDNGMetadata, then
// RawConverter::demosaicImage.  Writes OUT_PREFIX.params (the parameters the
// metadata gave) and, unless PARAMS_ONLY, OUT_PREFIX.dng.rgb.
static int run_dng(const std::string& path, const std::string& out, bool awb, bool params_only) {
    gls::tiff_metadata dng_metadata, exif_metadata;
    const auto rawImage = gls::image<gls::luma_pixel_16>::read_dng_file(path, &dng_metadata, &exif_metadata);
    DemosaicParameters p = synthetic_params(rggb, 0, 0xffff);
    const float exposure = unpackDNGMetadata(*rawImage, &dng_metadata, &p, awb, nullptr, false);
    {
        std::ofstream f(out + ".params");
        f.precision(9);
        f << rawImage->width << ' ' << rawImage->height << ' ' << (int)p.bayerPattern << ' ' << p.black_level << ' '
          << p.white_level << ' ' << p.exposure_multiplier << ' ' << exposure;
        for (float v : p.scale_mul) f << ' ' << v;
        for (const auto& r : p.rgb_cam)
            for (float v : r) f << ' ' << v;
        f << '\n';
    }
    if (params_only) return 0;
    gls::HIPContext ctx(0);
    RawConverter rc(&ctx);
    auto rgb = RawConverter::convertToRGBImage(*rc.demosaicImage(*rawImage, &p, nullptr, false));
    write_rgb(out + ".dng.rgb", *rgb);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 4 && (std::string(argv[1]) == "--dng" || std::string(argv[1]) == "--dng-params")) {
        try {
            return run_dng(argv[2], argv[3], argc >= 5 && std::atoi(argv[4]) != 0,
                           std::string(argv[1]) == "--dng-params");
        } catch (const std::exception& e) {
            std::cerr << "raw_pipeline: " << e.what() << std::endl;
            return 1;
        }
    }
    if (argc < 6 || argc > 8) {
        std::cerr << "usage: raw_pipeline W H BAYER LEVELS OUT_PREFIX [LTM [NOISE_LEVEL]]\n";
        return 2;
    }
    const int ltm = argc >= 7 ? std::atoi(argv[6]) : 0;
    const float noise_level = argc >= 8 ? (float)std::atof(argv[7]) : 0.0f;
    const int w = std::atoi(argv[1]), h = std::atoi(argv[2]);
    const BayerPattern bayer = (BayerPattern)std::atoi(argv[3]);
    const int levels = std::atoi(argv[4]);
    const std::string out = argv[5];
    if (levels != 5) {
        std::cerr << "this driver instantiates PyramidalDenoise<5>\n";
        return 2;
    }
    try {
        gls::HIPContext ctx(0);
        const float black = 64, white = 4095;
        gls::hip_image_2d<gls::luma_pixel_16> rawDev(&ctx, w, h);
        ctx.check(clim_generate_bayer(ctx.ctx(), rawDev.clim(), nullptr, bayer, 1000, 2000, 0.002f, 1e-5f, black,
                                      white),
                  "generate");
        auto rawImage = rawDev.toImage();

        // 1. the fused pipeline through RawConverter
        DemosaicParameters p1 = synthetic_params(bayer, black, white);
        p1.rgbConversionParameters.localToneMapping = ltm;
        p1.noiseLevel = noise_level;
        RawConverter rc(&ctx);
        auto rgb1 = RawConverter::convertToRGBImage(*rc.demosaicImage(*rawImage, &p1, nullptr, false));
        write_rgb(out + ".fused.rgb", *rgb1);

        // 2. operator by operator (cpp/raw_converter.cpp:101-174)
        DemosaicParameters p2 = synthetic_params(bayer, black, white);
        p2.rgbConversionParameters.localToneMapping = ltm;
        p2.noiseLevel = noise_level;
        gls::hip_image_2d<gls::luma_pixel_16> raw(&ctx, *rawImage);
        gls::hip_image_2d<gls::luma_pixel_float> scaled(&ctx, w, h), green(&ctx, w, h);
        gls::hip_image_2d<gls::rgba_pixel_float> A(&ctx, w, h), B(&ctx, w, h);
        gls::hip_image_2d<gls::rgba_pixel> srgb(&ctx, w, h);
        scaleRawData(&ctx, raw, &scaled, bayer, p2.scale_mul, p2.black_level / 0xffff);
        const auto rawNLF = computeRawNoiseStatistics(&ctx, scaled, bayer);
        p2.noiseModel.rawNlf = {rawNLF[4], rawNLF[5], rawNLF[6], rawNLF[7]};
        if (p2.noiseLevel > 0.6f) {
            // cpp/raw_converter.cpp:111-123
            gls::hip_image_2d<gls::rgba_pixel_float> q0(&ctx, w / 2, h / 2), q1(&ctx, w / 2, h / 2);
            bayerToRawRGBA(&ctx, scaled, &q0, bayer);
            despeckleRawRGBAImage(&ctx, q0, &q1);
            denoiseRawRGBAImage(&ctx, q1, p2.noiseModel.rawNlf, &q0);
            rawRGBAToBayer(&ctx, q0, &scaled, bayer);
        }
        interpolateGreen(&ctx, scaled, &green, bayer, p2.noiseModel.rawNlf[1]);
        interpolateRedBlue(&ctx, scaled, green, &A, bayer, (p2.noiseModel.rawNlf[0] + p2.noiseModel.rawNlf[2]) / 2,
                           false);
        blendHighlightsImage(&ctx, A, 1.0, &A);
        const auto cam_to_ycbcr = cam_ycbcr(p2.rgb_cam);
        transformImage(&ctx, A, &A, cam_to_ycbcr);
        const auto& np = p2.noiseModel.pyramidNlf[0];
        despeckleImage(&ctx, A, {np[0], np[1], np[2]}, {np[3], np[4], np[5]}, &B);
        // the pyramid twice: with a caller plug-in (level by level, as the
        // reference orchestrates it) and with the built-in filter (one fused
        // call that fills the same level images)
        const Matrix3x3 inv = inverse(cam_to_ycbcr);  // exposure_multiplier = 1
        DemosaicParameters p3 = p2;
        for (int plugin = 1; plugin >= 0; plugin--) {
            DemosaicParameters& pp = plugin ? p2 : p3;
            PyramidalDenoise<5> pyramid(&ctx, w, h);
            if (plugin)
                for (int i = 0; i < 5; i++) pyramid.denoiser[i] = std::make_unique<MyBilateral>(&ctx, w >> i, h >> i);
            // (the pyramid reads its input image, it does not write it)
            auto* den = pyramid.denoise(&ctx, &pp.denoiseParameters, &B, pp.rgb_cam, nullptr, false,
                                        &pp.noiseModel.pyramidNlf);
            // local tone mapping mask on the denoised YCbCr image, guided by level 3
            // (cpp/raw_converter.cpp:164-167); a 1x1 placeholder when off
            gls::hip_image_2d<gls::luma_pixel_float> ltmMask(&ctx, ltm ? w : 1, ltm ? h : 1);
            if (ltm)
                localToneMappingMask(&ctx, *den, *pyramid.denoisedImagePyramid[3], pp.ltmParameters, inv, &ltmMask);
            transformImage(&ctx, *den, den, inv);
            convertTosRGB(&ctx, *den, ltmMask, &srgb, pp);
            auto rgb2 = RawConverter::convertToRGBImage(srgb);
            write_rgb(out + (plugin ? ".stages.rgb" : ".builtin.rgb"), *rgb2);
        }

        // 3. the half-resolution preview (cpp/raw_converter.cpp:188-224), fused
        //    and operator by operator
        {
            DemosaicParameters pf = synthetic_params(bayer, black, white);
            auto prev = RawConverter::convertToRGBImage(*rc.fastDemosaicImage(*rawImage, pf));
            write_rgb(out + ".preview.rgb", *prev);
            gls::hip_image_2d<gls::luma_pixel_float> sc(&ctx, w, h);
            gls::hip_image_2d<gls::rgba_pixel_float> lin(&ctx, w / 2, h / 2);
            gls::hip_image_2d<gls::rgba_pixel> prgb(&ctx, w / 2, h / 2);
            gls::hip_image_2d<gls::luma_pixel_float> placeholder(&ctx, 1, 1);
            scaleRawData(&ctx, raw, &sc, bayer, pf.scale_mul, pf.black_level / 0xffff);
            fasteDebayer(&ctx, sc, &lin, bayer);
            blendHighlightsImage(&ctx, lin, 1.0, &lin);
            convertTosRGB(&ctx, lin, placeholder, &prgb, pf);
            write_rgb(out + ".preview_stages.rgb", *RawConverter::convertToRGBImage(prgb));
        }

        // 4. two frames in one RawConverter::demosaicImages call (the second
        //    on another of the context's streams): each equals the fused result
        DemosaicParameters p4 = synthetic_params(bayer, black, white), p5 = p4;
        for (DemosaicParameters* pp : {&p4, &p5}) {
            pp->rgbConversionParameters.localToneMapping = ltm;
            pp->noiseLevel = noise_level;
        }
        {
            gls::hip_image_2d<gls::rgba_pixel> b0(&ctx, w, h), b1(&ctx, w, h);
            rc.demosaicImages({&rawDev, &rawDev}, {&b0, &b1}, {&p4, &p5});
            write_rgb(out + ".batch0.rgb", *RawConverter::convertToRGBImage(b0));
            write_rgb(out + ".batch1.rgb", *RawConverter::convertToRGBImage(b1));
        }

        std::ofstream f(out + ".nlf");
        f.precision(9);
        for (const auto* p : {&p1, &p2, &p3, &p4, &p5}) {
            for (float v : p->noiseModel.rawNlf) f << v << ' ';
            for (const auto& row : p->noiseModel.pyramidNlf)
                for (float v : row) f << v << ' ';
            f << '\n';
        }
        std::cout << "raw_pipeline: " << w << "x" << h << " done" << std::endl;
    } catch (const std::exception& e) {
        std::cerr << "raw_pipeline: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
