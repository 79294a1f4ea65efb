// gls_image.hpp - host image and pixel types of the RAW->RGB path.
//
// Source-compatible subset of the reference's gls::image<T> and pixel
// typedefs (src/gls_image.hpp:55-457) that the path's operator signatures
// use.  float_type is FP32: this build implements the reference's
// USE_FP16_FLOATS=0 configuration (src/gls_image.hpp:32,189-210).
#pragma once

#include <array>
#include <cassert>
#include <cstdint>
#include <map>
#include <memory>
#include <span>
#include <string>
#include <variant>
#include <vector>

namespace gls {

struct point {
    int x, y;
};

struct size {
    int width, height;
};

struct rectangle {
    int x, y, width, height;
};

template <typename T>
struct basic_luma_pixel {
    static constexpr int channels = 1;
    static constexpr int bit_depth = 8 * sizeof(T);
    T luma;
    T& operator[](int) { return luma; }
    const T& operator[](int) const { return luma; }
};

template <typename T>
struct basic_luma_alpha_pixel {
    static constexpr int channels = 2;
    static constexpr int bit_depth = 8 * sizeof(T);
    T luma, alpha;
    T& operator[](int c) { return (&luma)[c]; }
    const T& operator[](int c) const { return (&luma)[c]; }
};

template <typename T>
struct basic_rgb_pixel {
    static constexpr int channels = 3;
    static constexpr int bit_depth = 8 * sizeof(T);
    T red, green, blue;
    T& operator[](int c) { return (&red)[c]; }
    const T& operator[](int c) const { return (&red)[c]; }
};

template <typename T>
struct basic_rgba_pixel {
    static constexpr int channels = 4;
    static constexpr int bit_depth = 8 * sizeof(T);
    T red, green, blue, alpha;
    T& operator[](int c) { return (&red)[c]; }
    const T& operator[](int c) const { return (&red)[c]; }
};

typedef basic_luma_pixel<uint8_t> luma_pixel;
typedef basic_rgb_pixel<uint8_t> rgb_pixel;
typedef basic_rgba_pixel<uint8_t> rgba_pixel;
typedef basic_luma_pixel<uint16_t> luma_pixel_16;
typedef basic_rgb_pixel<uint16_t> rgb_pixel_16;
typedef basic_luma_pixel<float> luma_pixel_fp32;
typedef basic_rgba_pixel<float> rgba_pixel_fp32;
typedef basic_luma_alpha_pixel<float> pixel_fp32_2;
typedef basic_rgba_pixel<float> pixel_fp32_4;

typedef float float_type;
typedef basic_luma_pixel<float_type> luma_pixel_float;
typedef basic_rgba_pixel<float_type> rgba_pixel_float;
typedef basic_rgba_pixel<float_type> pixel_float4;

// TIFF / DNG metadata, tag -> values (src/gls_tiff_metadata.hpp); read with
// getVector / getValue (gls_dng.hpp)
typedef std::variant<std::vector<uint8_t>, std::vector<uint16_t>, std::vector<uint32_t>, std::vector<float>,
                     std::vector<double>, std::string, float, uint16_t, uint32_t>
    tiff_metadata_value;
typedef std::map<uint16_t, tiff_metadata_value> tiff_metadata;

// Row-major host image with a stride (in pixels); owns its storage or wraps
// the caller's (src/gls_image.hpp:230-300).
template <typename T>
class image {
   public:
    const int width;
    const int height;
    const int stride;
    typedef std::unique_ptr<image<T>> unique_ptr;

    image(int w, int h, int s) : width(w), height(h), stride(s), _store(std::make_shared<std::vector<T>>((size_t)s * h)),
                                 _data(_store->data(), _store->size()) {}
    image(int w, int h) : image(w, h, w) {}
    image(int w, int h, int s, std::span<T> data) : width(w), height(h), stride(s), _data(data) {
        assert((size_t)s * h <= data.size());
    }

    // Image factory from a DNG file (src/gls_image.hpp:429-443); defined for
    // luma_pixel_16 in gls_dng.hpp.  EXIF metadata is not read.
    static unique_ptr read_dng_file(const std::string& filename, tiff_metadata* dng_metadata = nullptr,
                                    tiff_metadata* exif_metadata = nullptr);

    T* operator[](int row) { return _data.data() + (size_t)row * stride; }
    const T* operator[](int row) const { return _data.data() + (size_t)row * stride; }
    std::span<T> pixels() { return _data; }
    std::span<const T> pixels() const { return {_data.data(), _data.size()}; }

    template <typename F>
    void apply(F f) {
        for (int y = 0; y < height; y++)
            for (int x = 0; x < width; x++) f(&(*this)[y][x], x, y);
    }

   private:
    std::shared_ptr<std::vector<T>> _store;
    std::span<T> _data;
};

}  // namespace gls
