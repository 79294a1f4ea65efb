// gls_dng.hpp - DNG ingest for the C++ API on the library's host decoder:
// gls::image<gls::luma_pixel_16>::read_dng_file (src/gls_image.hpp:429-443,
// src/gls_image_tiff.cpp:255-458) and unpackDNGMetadata
// (cpp/demosaic_utils.cpp:280-375), with the reference's getVector/getValue
// accessors of the metadata map (src/gls_tiff_metadata.hpp).  No GPU needed.
#pragma once

#include <cmath>
#include <fstream>
#include <iterator>
#include <type_traits>

#include "climage.h"
#include "demosaic.hpp"
#include "gls_hip_image.hpp"
#include "gls_image.hpp"

// the DNG 1.4 tags the path reads
#ifndef TIFFTAG_CFAPATTERN
#define TIFFTAG_ORIENTATION 274
#define TIFFTAG_CFAPATTERN 33422
#define TIFFTAG_BLACKLEVEL 50714
#define TIFFTAG_WHITELEVEL 50717
#define TIFFTAG_COLORMATRIX1 50721
#define TIFFTAG_COLORMATRIX2 50722
#define TIFFTAG_ASSHOTNEUTRAL 50728
#define TIFFTAG_BASELINEEXPOSURE 50730
#endif

namespace gls {

// the values of a tag as T (converted from whichever vector type holds them)
template <typename T>
std::vector<T> getVector(const tiff_metadata& metadata, uint16_t key) {
    const auto entry = metadata.find(key);
    if (entry == metadata.end()) return {};
    return std::visit(
        [](const auto& v) -> std::vector<T> {
            using V = std::decay_t<decltype(v)>;
            if constexpr (std::is_same_v<V, std::string>) {
                return {};
            } else if constexpr (std::is_arithmetic_v<V>) {
                return {(T)v};
            } else {
                return std::vector<T>(v.begin(), v.end());
            }
        },
        entry->second);
}

template <typename T>
bool getValue(const tiff_metadata& metadata, uint16_t key, T* value) {
    const auto v = getVector<T>(metadata, key);
    if (v.empty()) return false;
    *value = v[0];
    return true;
}

template <>
inline image<luma_pixel_16>::unique_ptr image<luma_pixel_16>::read_dng_file(const std::string& filename,
                                                                            tiff_metadata* dng_metadata,
                                                                            tiff_metadata* /*exif_metadata*/) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw clim_error(CLIM_EINVAL, "Couldn't read dng file.");
    const std::vector<uint8_t> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    clim_dng_info info;
    if (clim_dng_decode(data.data(), (int64_t)data.size(), &info, nullptr, 0) != CLIM_OK)
        throw clim_error(CLIM_EINVAL, std::string("read_dng_file: ") + clim_dng_last_error());
    auto img = std::make_unique<image<luma_pixel_16>>(info.width, info.height);
    static_assert(sizeof(luma_pixel_16) == sizeof(uint16_t));
    if (clim_dng_decode(data.data(), (int64_t)data.size(), &info, reinterpret_cast<uint16_t*>(img->pixels().data()),
                        img->stride) != CLIM_OK)
        throw clim_error(CLIM_EINVAL, std::string("read_dng_file: ") + clim_dng_last_error());
    if (dng_metadata) {
        // insert: values already in the map win, as std::map::insert in the reference
        static const uint8_t kCfa[4][4] = {{1, 0, 2, 1}, {1, 2, 0, 1}, {0, 1, 1, 2}, {2, 1, 1, 0}};
        dng_metadata->insert({TIFFTAG_BLACKLEVEL, std::vector<float>{info.black_level}});
        dng_metadata->insert({TIFFTAG_WHITELEVEL, std::vector<uint32_t>{(uint32_t)info.white_level}});
        dng_metadata->insert({TIFFTAG_CFAPATTERN, std::vector<uint8_t>(kCfa[info.bayerPattern], kCfa[info.bayerPattern] + 4)});
        dng_metadata->insert({TIFFTAG_ORIENTATION, (uint16_t)info.orientation});
        if (info.has_color_matrix1)
            dng_metadata->insert({TIFFTAG_COLORMATRIX1, std::vector<float>(info.color_matrix1, info.color_matrix1 + 9)});
        if (info.has_color_matrix2)
            dng_metadata->insert({TIFFTAG_COLORMATRIX2, std::vector<float>(info.color_matrix2, info.color_matrix2 + 9)});
        if (info.has_as_shot_neutral)
            dng_metadata->insert({TIFFTAG_ASSHOTNEUTRAL,
                                  std::vector<float>(info.as_shot_neutral, info.as_shot_neutral + 3)});
        dng_metadata->insert({TIFFTAG_BASELINEEXPOSURE, info.baseline_exposure});
    }
    return img;
}

}  // namespace gls

// cpp/demosaic_utils.cpp:280-375.  Sets bayerPattern, black_level, white_level,
// exposure_multiplier, scale_mul and rgb_cam; with auto_white_balance, writes the
// measured white balance back as AsShotNeutral (:348) and the highlight fraction
// to *highlights.  Returns the exposure multiplier.  ColorChecker calibration
// (gmb_position) is not supported.
inline float unpackDNGMetadata(const gls::image<gls::luma_pixel_16>& rawImage, gls::tiff_metadata* dng_metadata,
                               DemosaicParameters* demosaicParameters, bool auto_white_balance,
                               const gls::rectangle* gmb_position, bool /*rotate_180*/, float* highlights = nullptr) {
    if (gmb_position) throw gls::clim_error(CLIM_EINVAL, "unpackDNGMetadata: ColorChecker calibration is not supported");
    clim_dng_info info{};
    info.width = rawImage.width;
    info.height = rawImage.height;
    const auto cfa = gls::getVector<uint8_t>(*dng_metadata, TIFFTAG_CFAPATTERN);
    info.bayerPattern = cfa.size() >= 4 && !std::memcmp(cfa.data(), "\0\1\1\2", 4)   ? CLIM_RGGB
                        : cfa.size() >= 4 && !std::memcmp(cfa.data(), "\2\1\1\0", 4) ? CLIM_BGGR
                        : cfa.size() >= 4 && !std::memcmp(cfa.data(), "\1\0\2\1", 4) ? CLIM_GRBG
                                                                                     : CLIM_GBRG;
    const auto black = gls::getVector<float>(*dng_metadata, TIFFTAG_BLACKLEVEL);
    const auto white = gls::getVector<uint32_t>(*dng_metadata, TIFFTAG_WHITELEVEL);
    info.black_level = black.empty() ? 0.0f : black[0];
    info.white_level = white.empty() ? 65535.0f : (float)white[0];
    auto vec = [&](uint16_t tag, float* dst, int n, int32_t* has) {
        const auto v = gls::getVector<float>(*dng_metadata, tag);
        *has = (int)v.size() >= n;
        if (*has) std::copy(v.begin(), v.begin() + n, dst);
    };
    vec(TIFFTAG_COLORMATRIX1, info.color_matrix1, 9, &info.has_color_matrix1);
    vec(TIFFTAG_COLORMATRIX2, info.color_matrix2, 9, &info.has_color_matrix2);
    vec(TIFFTAG_ASSHOTNEUTRAL, info.as_shot_neutral, 3, &info.has_as_shot_neutral);
    gls::getValue(*dng_metadata, TIFFTAG_BASELINEEXPOSURE, &info.baseline_exposure);
    clim_demosaic_params p;
    clim_default_params(&p, info.bayerPattern, info.black_level, info.white_level);
    float exposure = 1.0f, hl = 0.0f, asn[3];
    if (clim_unpack_dng_metadata(&info, reinterpret_cast<const uint16_t*>(&rawImage[0][0]), rawImage.stride,
                                 auto_white_balance, &p, &exposure, &hl, asn) != CLIM_OK)
        throw gls::clim_error(CLIM_EINVAL, std::string("unpackDNGMetadata: ") + clim_dng_last_error());
    demosaicParameters->bayerPattern = (BayerPattern)p.bayerPattern;
    demosaicParameters->black_level = p.black_level;
    demosaicParameters->white_level = p.white_level;
    demosaicParameters->exposure_multiplier = p.exposure_multiplier;
    for (int c = 0; c < 4; c++) demosaicParameters->scale_mul[c] = p.scale_mul[c];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) demosaicParameters->rgb_cam[i][j] = p.rgb_cam[i][j];
    if (auto_white_balance) {
        (*dng_metadata)[TIFFTAG_ASSHOTNEUTRAL] = std::vector<float>(asn, asn + 3);
        if (highlights) *highlights = hl;
    }
    return exposure;
}
