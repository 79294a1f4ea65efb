// gls_hip_image.hpp - device images and the device context on gfx950.
//
// gls::HIPContext stands in for gls::OpenCLContext (src/gls_cl.hpp:47-85) and
// gls::hip_image_2d<T> for gls::cl_image_2d<T> (src/gls_cl_image.hpp:37-107):
// same member names (width, height, copyPixelsFrom/To, mapImage/unmapImage,
// toImage, fromImage), over pitch-linear device memory owned through the C ABI.
#pragma once

#include <stdexcept>
#include <string>

#include "climage.h"
#include "gls_image.hpp"

namespace gls {

// Thrown where the reference would throw cl::Error (src/gls_cl.hpp:21).
struct clim_error : std::runtime_error {
    clim_status status;
    clim_error(clim_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

class HIPContext {
   public:
    explicit HIPContext(int device = 0) {
        clim_status s = clim_create(&_ctx, device);
        if (s != CLIM_OK) throw clim_error(s, std::string("HIPContext: ") + clim_status_string(s));
    }
    ~HIPContext() { clim_destroy(_ctx); }
    HIPContext(const HIPContext&) = delete;
    HIPContext& operator=(const HIPContext&) = delete;

    clim_ctx* ctx() const { return _ctx; }
    void check(clim_status s, const char* what) const {
        if (s != CLIM_OK) throw clim_error(s, std::string(what) + ": " + clim_last_error(_ctx));
    }
    void finish() const { check(clim_sync(_ctx), "finish"); }

   private:
    clim_ctx* _ctx = nullptr;
};

template <typename T>
struct hip_format;
template <> struct hip_format<luma_pixel_16> { static constexpr int value = CLIM_FMT_R16U; };
template <> struct hip_format<luma_pixel_float> { static constexpr int value = CLIM_FMT_R32F; };
template <> struct hip_format<rgba_pixel_float> { static constexpr int value = CLIM_FMT_RGBA32F; };
template <> struct hip_format<rgba_pixel> { static constexpr int value = CLIM_FMT_RGBA8; };
template <> struct hip_format<pixel_fp32_2> { static constexpr int value = CLIM_FMT_RG32F; };

template <typename T>
class hip_image_2d {
   public:
    typedef std::unique_ptr<hip_image_2d<T>> unique_ptr;
    const int width;
    const int height;

    hip_image_2d(HIPContext* ctx, int w, int h) : width(w), height(h), _ctx(ctx) {
        _ctx->check(clim_image_alloc(_ctx->ctx(), w, h, hip_format<T>::value, &_img), "hip_image_2d");
    }
    hip_image_2d(HIPContext* ctx, const image<T>& other) : hip_image_2d(ctx, other.width, other.height) {
        copyPixelsFrom(other);
    }
    ~hip_image_2d() { clim_image_free(_ctx->ctx(), &_img); }
    hip_image_2d(const hip_image_2d&) = delete;
    hip_image_2d& operator=(const hip_image_2d&) = delete;

    static unique_ptr fromImage(HIPContext* ctx, const image<T>& other) {
        return std::make_unique<hip_image_2d<T>>(ctx, other);
    }

    // blocking, like the reference's enqueueWriteImage(..., true, ...)
    void copyPixelsFrom(const image<T>& other) const {
        if (other.width != width || other.height != height) throw clim_error(CLIM_EINVAL, "copyPixelsFrom: size");
        _ctx->check(clim_copy_from_host(_ctx->ctx(), other.pixels().data(), (int64_t)other.stride * sizeof(T),
                                        const_cast<clim_image*>(&_img)),
                    "copyPixelsFrom");
    }
    void copyPixelsTo(image<T>* other) const {
        if (other->width != width || other->height != height) throw clim_error(CLIM_EINVAL, "copyPixelsTo: size");
        _ctx->check(clim_copy_to_host(_ctx->ctx(), &_img, other->pixels().data(), (int64_t)other->stride * sizeof(T)),
                    "copyPixelsTo");
    }
    typename image<T>::unique_ptr toImage() const {
        auto out = std::make_unique<image<T>>(width, height);
        copyPixelsTo(out.get());
        return out;
    }
    // mapImage returns a host copy; unmapImage writes it back (the reference
    // maps CL_MAP_READ | CL_MAP_WRITE, src/gls_cl_image.hpp:92-104)
    image<T> mapImage() const {
        image<T> host(width, height);
        copyPixelsTo(&host);
        return host;
    }
    void unmapImage(const image<T>& mapped) const { copyPixelsFrom(mapped); }

    const clim_image* clim() const { return &_img; }
    clim_image* clim() { return &_img; }
    HIPContext* context() const { return _ctx; }

   private:
    HIPContext* _ctx;
    clim_image _img{};
};

}  // namespace gls
