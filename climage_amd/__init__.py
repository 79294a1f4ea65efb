"""climage_amd - MI355X (gfx950) RAW->RGB path of CLImage behind a C ABI.

This module is a thin ctypes front end of climage_amd/lib/libclimage_hip.so
(include/climage.h).  It mirrors the reference's operator names
(cpp/demosaic_cl.hpp, cpp/pyramidal_denoise.hpp, cpp/raw_converter.hpp) so
that tests and tools read like the reference's call sites.  There is no CPU
fallback: every call runs the HIP kernels, and importing raises if the
library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libclimage_hip.so")

# clim_status
OK, EINVAL, ENOMEM, EHIP, ENUMERIC, ENODEV = range(6)
# clim_format
FMT_R16U, FMT_R32F, FMT_RGBA32F, FMT_RGBA8, FMT_RG32F, FMT_RGBA16F = 1, 2, 3, 4, 5, 6
# BayerPattern (cpp/demosaic.hpp:23-28)
GRBG, GBRG, RGGB, BGGR = 0, 1, 2, 3
# PyramidalDenoise::DenoiseAlgorithm (cpp/pyramidal_denoise.hpp:38)
BILATERAL, GUIDED_FAST, GUIDED_PRECISE = 0, 1, 2

_NP = {FMT_R16U: (np.uint16, 1), FMT_R32F: (np.float32, 1), FMT_RGBA32F: (np.float32, 4),
       FMT_RGBA8: (np.uint8, 4), FMT_RG32F: (np.float32, 2), FMT_RGBA16F: (np.float16, 4)}


class ClimError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


class Image(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("pitch_bytes", C.c_int32), ("format", C.c_int32)]


class DenoiseParameters(C.Structure):
    _fields_ = [("luma", C.c_float), ("chroma", C.c_float), ("chromaBoost", C.c_float),
                ("gradientBoost", C.c_float), ("sharpening", C.c_float)]


class RGBConversionParameters(C.Structure):
    _fields_ = [("contrast", C.c_float), ("saturation", C.c_float), ("toneCurveSlope", C.c_float),
                ("exposureBias", C.c_float), ("blacks", C.c_float), ("localToneMapping", C.c_int32)]


class NoiseModel(C.Structure):
    _fields_ = [("rawNlf", C.c_float * 4), ("pyramidNlf", (C.c_float * 6) * 5)]


class LTMParameters(C.Structure):
    _fields_ = [("guidedFilterEps", C.c_float), ("shadows", C.c_float), ("highlights", C.c_float),
                ("detail", C.c_float)]


class DemosaicParameters(C.Structure):
    _fields_ = [("bayerPattern", C.c_int32), ("black_level", C.c_float), ("white_level", C.c_float),
                ("exposure_multiplier", C.c_float), ("scale_mul", C.c_float * 4),
                ("rgb_cam", (C.c_float * 3) * 3), ("noiseModel", NoiseModel),
                ("denoiseParameters", DenoiseParameters * 5), ("noiseLevel", C.c_float),
                ("rgbConversionParameters", RGBConversionParameters), ("ltmParameters", LTMParameters),
                ("levels", C.c_int32), ("denoiseAlgorithm", C.c_int32)]


class DngInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("bayerPattern", C.c_int32),
                ("orientation", C.c_int32), ("bits_per_sample", C.c_int32), ("compression", C.c_int32),
                ("tiled", C.c_int32), ("black_level", C.c_float), ("white_level", C.c_float),
                ("color_matrix1", C.c_float * 9), ("color_matrix2", C.c_float * 9),
                ("has_color_matrix1", C.c_int32), ("has_color_matrix2", C.c_int32),
                ("as_shot_neutral", C.c_float * 3), ("has_as_shot_neutral", C.c_int32),
                ("baseline_exposure", C.c_float)]


_lib = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"climage_amd: {LIB_PATH} is missing; build it with "
                              f"`python climage_amd/build.py` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        L.clim_last_error.restype = C.c_char_p
        L.clim_status_string.restype = C.c_char_p
        L.clim_stream.restype = C.c_void_p
        L.clim_burst_destroy.restype = None
        _lib = L
    return _lib


PROFILE_STAGES = 14  # CLIM_PROFILE_STAGES


def exported_symbols():
    return [n for n in ("clim_create", "clim_destroy", "clim_last_error", "clim_status_string", "clim_sync",
                        "clim_stream", "clim_image_alloc", "clim_image_free", "clim_copy_from_host",
                        "clim_copy_to_host", "clim_format_texel_bytes", "clim_scale_raw_data",
                        "clim_raw_noise_statistics", "clim_compute_raw_noise_statistics",
                        "clim_interpolate_green", "clim_interpolate_red_blue", "clim_blend_highlights",
                        "clim_transform_image", "clim_despeckle_image", "clim_downsample_image",
                        "clim_noise_statistics", "clim_compute_noise_statistics", "clim_denoise_image",
                        "clim_denoise_image_guided", "clim_guided_filter", "clim_reassemble_image",
                        "clim_convert_to_srgb", "clim_local_tone_mapping_mask", "clim_bayer_to_raw_rgba",
                        "clim_raw_rgba_to_bayer", "clim_despeckle_raw_rgba", "clim_denoise_raw_rgba",
                        "clim_fast_debayer", "clim_fast_demosaic", "clim_dng_decode", "clim_unpack_dng_metadata",
                        "clim_dng_last_error",
                        "clim_cam_ycbcr", "clim_inverse3x3", "clim_default_params",
                        "clim_pyramidal_denoise", "clim_pyramidal_denoise_into", "clim_raw_to_rgb", "clim_raw_to_rgb_ex", "clim_raw_to_rgb_batch",
                        "clim_generate_bayer", "clim_compute_raw_noise_statistics_noalloc",
                        "clim_strength_set_weights", "clim_strength_get_weights", "clim_strength_set_mode",
                        "clim_strength_gate_read", "clim_strength_map_read", "clim_strength_loss_grad",
                        "clim_strength_train", "clim_burst_create", "clim_burst_launch", "clim_burst_frame_times",
                        "clim_burst_tuning_read", "clim_burst_destroy", "clim_set_precision")]


# learned denoise strength (include/climage.h, clim_strength_*)
STRENGTH_PARAMS = 26612
STRENGTH_OFF, STRENGTH_GATED, STRENGTH_FORCE_PLAIN, STRENGTH_FORCE_LEARNED = 0, 1, 2, 3
STRENGTH_WEIGHTS = os.path.join(HERE, "weights", "strength.f32")


class StrengthTrainConfig(C.Structure):
    _fields_ = [("frames", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("crop", C.c_int32),
                ("epochs", C.c_int32), ("batch", C.c_int32), ("lr", C.c_float), ("seed", C.c_uint32),
                ("bayer", C.c_int32), ("noise_levels", C.c_int32), ("verbose", C.c_int32)]


def load_strength_weights(path=STRENGTH_WEIGHTS):
    """The trained network parameters shipped with the package (float32,
    the canonical layout of clim_strength_set_weights)."""
    p = np.fromfile(path, dtype="<f4")
    if p.size != STRENGTH_PARAMS:
        raise ValueError(f"{path}: {p.size} parameters, expected {STRENGTH_PARAMS}")
    return p


def _f(vals):
    vals = [float(v) for v in np.asarray(vals, np.float32).ravel()]
    return (C.c_float * len(vals))(*vals)


def dng_decode(data):
    """read_dng_file from the file's bytes (host only, no GPU): (raw uint16 (H, W), DngInfo)."""
    L = lib()
    L.clim_dng_last_error.restype = C.c_char_p
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    info = DngInfo()
    st = L.clim_dng_decode(buf, C.c_int64(len(data)), C.byref(info), None, C.c_int64(0))
    if st != OK:
        raise ClimError(st, "dng_decode: " + L.clim_dng_last_error().decode())
    raw = np.zeros((info.height, info.width), np.uint16)
    st = L.clim_dng_decode(buf, C.c_int64(len(data)), C.byref(info), raw.ctypes.data_as(C.c_void_p),
                           C.c_int64(info.width))
    if st != OK:
        raise ClimError(st, "dng_decode: " + L.clim_dng_last_error().decode())
    return raw, info


def unpack_dng_metadata(info, raw=None, auto_white_balance=False, params=None):
    """unpackDNGMetadata into DemosaicParameters (host only).  Returns
    (params, exposure multiplier, highlights fraction, as-shot neutral used)."""
    L = lib()
    L.clim_dng_last_error.restype = C.c_char_p
    p = params if params is not None else default_params(info.bayerPattern)
    exp, hl = C.c_float(), C.c_float()
    asn = (C.c_float * 3)()
    rp = None
    pitch = 0
    if raw is not None:
        raw = np.ascontiguousarray(raw, np.uint16)
        rp, pitch = raw.ctypes.data_as(C.c_void_p), raw.shape[1]
    st = L.clim_unpack_dng_metadata(C.byref(info), rp, C.c_int64(pitch), int(bool(auto_white_balance)), C.byref(p),
                                    C.byref(exp), C.byref(hl), asn)
    if st != OK:
        raise ClimError(st, "unpackDNGMetadata: " + L.clim_dng_last_error().decode())
    return p, exp.value, hl.value, np.array(asn[:], np.float32)


def default_params(bayer=RGGB, black=64.0, white=4095.0, levels=5, algorithm=BILATERAL):
    """DemosaicParameters for a synthetic identity-colour camera (clim_default_params)."""
    p = DemosaicParameters()
    lib().clim_default_params(C.byref(p), C.c_int32(bayer), C.c_float(black), C.c_float(white))
    p.levels = levels
    p.denoiseAlgorithm = algorithm
    return p


def cam_ycbcr(rgb_cam):
    out = (C.c_float * 9)()
    lib().clim_cam_ycbcr(_f(rgb_cam), out)
    return np.array(out[:], np.float32).reshape(3, 3)


def inverse3x3(m):
    out = (C.c_float * 9)()
    st = lib().clim_inverse3x3(_f(m), out)
    if st != OK:
        raise ClimError(st, "inverse3x3: singular matrix")
    return np.array(out[:], np.float32).reshape(3, 3)


class DeviceImage:
    """A device image owned by a Context (cl_image_2d counterpart)."""

    def __init__(self, ctx, width, height, fmt):
        self.ctx = ctx
        self.im = Image()
        ctx._check(lib().clim_image_alloc(ctx.h, width, height, fmt, C.byref(self.im)), "image_alloc")

    @property
    def width(self):
        return self.im.width

    @property
    def height(self):
        return self.im.height

    @property
    def format(self):
        return self.im.format

    def ref(self):
        return C.byref(self.im)

    def upload(self, arr):
        dt, ch = _NP[self.format]
        arr = np.ascontiguousarray(arr, dt)
        shape = (self.height, self.width) if ch == 1 else (self.height, self.width, ch)
        if arr.shape != shape:
            raise ValueError(f"upload: shape {arr.shape} != {shape}")
        self.ctx._check(lib().clim_copy_from_host(self.ctx.h, arr.ctypes.data_as(C.c_void_p),
                                                  C.c_int64(arr.strides[0]), self.ref()), "copy_from_host")

    def download(self):
        dt, ch = _NP[self.format]
        shape = (self.height, self.width) if ch == 1 else (self.height, self.width, ch)
        out = np.empty(shape, dt)
        self.ctx._check(lib().clim_copy_to_host(self.ctx.h, self.ref(), out.ctypes.data_as(C.c_void_p),
                                                C.c_int64(out.strides[0])), "copy_to_host")
        return out

    def free(self):
        if self.im.data:
            lib().clim_image_free(self.ctx.h, C.byref(self.im))


BURST_FRAME_TIMES = 1  # CLIM_BURST_FRAME_TIMES
BURST_CROSS_FRAME, BURST_PER_FRAME = 2, 4  # CLIM_BURST_CROSS_FRAME, CLIM_BURST_PER_FRAME
PRECISION_FP32, PRECISION_MIXED = 0, 1  # clim_precision


class BurstTuning(C.Structure):
    _fields_ = [("waves", C.c_int32 * 2), ("blocks_per_cu", C.c_int32 * 2), ("grab", C.c_int32 * 2),
                ("us", C.c_float * 2), ("us_first", C.c_float * 2), ("candidates", C.c_int32 * 2),
                ("tiles", C.c_int32 * 2), ("cross_frames", C.c_int32), ("cached", C.c_int32),
                ("form", C.c_int32), ("form_ms", C.c_float * 2)]


class Burst:
    """A burst recorded once and replayed (clim_burst_*): frame i from raws[i]
    into outs8[i] (and outs_f32[i]) with params[i].  The images must outlive
    the burst.  form: None (both recorded forms timed, the faster kept),
    "cross_frame" or "per_frame" (that form only)."""

    def __init__(self, ctx, raws, outs8, params, outs_f32=None, frame_times=False, form=None):
        n = len(raws)
        if len(outs8) != n or len(params) != n or (outs_f32 is not None and len(outs_f32) != n):
            raise ValueError("Burst: lists of different lengths")
        self.ctx, self.n = ctx, n
        self._keep = (list(raws), list(outs8), list(outs_f32 or []))
        R = (C.POINTER(Image) * n)(*[C.pointer(r.im) for r in raws])
        O8 = (C.POINTER(Image) * n)(*[C.pointer(o.im) for o in outs8])
        OF = None
        if outs_f32 is not None:
            OF = (C.POINTER(Image) * n)(*[C.pointer(o.im) if o is not None else None for o in outs_f32])
        P = (DemosaicParameters * n)(*params)
        self.h = C.c_void_p()
        flags = (BURST_FRAME_TIMES if frame_times else 0) | \
            {None: 0, "cross_frame": BURST_CROSS_FRAME, "per_frame": BURST_PER_FRAME}[form]
        ctx._check(lib().clim_burst_create(ctx.h, n, R, O8, OF, P, C.c_uint32(flags), C.byref(self.h)),
                   "burst_create")

    def launch(self):
        self.ctx._check(lib().clim_burst_launch(self.h), "burst_launch")

    def frame_times(self):
        """ms from the last launch's start to each frame's end (blocks)."""
        ms = (C.c_float * self.n)()
        self.ctx._check(lib().clim_burst_frame_times(self.h, ms), "burst_frame_times")
        return np.array(ms[:], np.float32)

    def tuning(self):
        t = BurstTuning()
        self.ctx._check(lib().clim_burst_tuning_read(self.h, C.byref(t)), "burst_tuning_read")
        names = ("green", "red_blue_despeckle")
        return {names[k]: {"waves_per_simd": t.waves[k], "blocks_per_cu": t.blocks_per_cu[k], "grab": t.grab[k],
                           "us": round(t.us[k], 2), "us_first_candidate": round(t.us_first[k], 2),
                           "candidates": t.candidates[k], "tiles": t.tiles[k]} for k in range(2)} | \
            {"cross_frames": t.cross_frames, "cached": bool(t.cached),
             "form": "cross-frame launches" if t.form == 0 else "per-frame kernels",
             "replay_ms": {"cross_frame": round(t.form_ms[0], 3), "per_frame": round(t.form_ms[1], 3)}}

    def close(self):
        if self.h:
            lib().clim_burst_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Context:
    """One HIP stream plus work buffers on one gfx950 device (OpenCLContext counterpart)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        st = lib().clim_create(C.byref(self.h), device)
        if st != OK:
            raise ClimError(st, f"clim_create(device={device}): {lib().clim_status_string(st).decode()}")

    def close(self):
        if self.h:
            lib().clim_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st != OK:
            raise ClimError(st, f"{what}: {lib().clim_last_error(self.h).decode()}")

    def image(self, width, height, fmt):
        return DeviceImage(self, width, height, fmt)

    def upload(self, arr, fmt):
        arr = np.asarray(arr)
        im = self.image(arr.shape[1], arr.shape[0], fmt)
        im.upload(arr)
        return im

    def sync(self):
        self._check(lib().clim_sync(self.h), "sync")

    def set_precision(self, precision):
        """CLIM_PRECISION_FP32 (default) or CLIM_PRECISION_MIXED (FP16 level
        storage, packed-FP16 bilateral weights with FP32 sums)."""
        self._check(lib().clim_set_precision(self.h, int(precision)), "set_precision")

    # ---- operators (cpp/demosaic_cl.hpp names) ---------------------------
    def scaleRawData(self, raw, scaled, bayer, scale_mul, black_level):
        self._check(lib().clim_scale_raw_data(self.h, raw.ref(), scaled.ref(), bayer, _f(scale_mul),
                                              C.c_float(black_level)), "scaleRawData")

    def rawNoiseStatistics(self, scaled, bayer, mean, var):
        self._check(lib().clim_raw_noise_statistics(self.h, scaled.ref(), bayer, mean.ref(), var.ref()),
                    "rawNoiseStatistics")

    def computeRawNoiseStatistics(self, scaled, bayer):
        out = (C.c_float * 8)()
        self._check(lib().clim_compute_raw_noise_statistics(self.h, scaled.ref(), bayer, out),
                    "computeRawNoiseStatistics")
        return np.array(out[:], np.float32)

    def computeRawNoiseStatisticsNoAlloc(self, scaled, bayer, nlf_dev, status_dev=None):
        """clim_compute_raw_noise_statistics_noalloc: asynchronous, no device
        allocation; nlf_dev (a device image of at least 8 floats) receives
        A[4], B[4], status_dev (nullable, a device image) the fit status."""
        self._check(lib().clim_compute_raw_noise_statistics_noalloc(
            self.h, scaled.ref(), bayer, C.c_void_p(nlf_dev.im.data),
            C.c_void_p(status_dev.im.data) if status_dev is not None else None),
            "computeRawNoiseStatisticsNoAlloc")

    def interpolateGreen(self, scaled, green, bayer, luma_variance):
        self._check(lib().clim_interpolate_green(self.h, scaled.ref(), green.ref(), bayer,
                                                 C.c_float(luma_variance)), "interpolateGreen")

    def interpolateRedBlue(self, scaled, green, rgb, bayer, chroma_variance, rotate_180=False):
        self._check(lib().clim_interpolate_red_blue(self.h, scaled.ref(), green.ref(), rgb.ref(), bayer,
                                                    C.c_float(chroma_variance), int(bool(rotate_180))),
                    "interpolateRedBlue")

    def blendHighlightsImage(self, inp, clip, out):
        self._check(lib().clim_blend_highlights(self.h, inp.ref(), C.c_float(clip), out.ref()), "blendHighlights")

    def transformImage(self, inp, out, transform):
        self._check(lib().clim_transform_image(self.h, inp.ref(), out.ref(), _f(transform)), "transformImage")

    def despeckleImage(self, inp, var_a, var_b, out):
        self._check(lib().clim_despeckle_image(self.h, inp.ref(), _f(var_a), _f(var_b), out.ref()),
                    "despeckleImage")

    def downsampleImage(self, inp, out):
        self._check(lib().clim_downsample_image(self.h, inp.ref(), out.ref()), "downsampleImage")

    def noiseStatistics(self, inp, out):
        self._check(lib().clim_noise_statistics(self.h, inp.ref(), out.ref()), "noiseStatistics")

    def computeNoiseStatistics(self, inp):
        out = (C.c_float * 6)()
        self._check(lib().clim_compute_noise_statistics(self.h, inp.ref(), out), "computeNoiseStatistics")
        return np.array(out[:], np.float32)

    def denoiseImage(self, inp, var_a, var_b, chroma_boost, gradient_boost, out):
        self._check(lib().clim_denoise_image(self.h, inp.ref(), _f(var_a), _f(var_b), C.c_float(chroma_boost),
                                             C.c_float(gradient_boost), out.ref()), "denoiseImage")

    def denoiseImageGuided(self, inp, var_a, var_b, out):
        self._check(lib().clim_denoise_image_guided(self.h, inp.ref(), _f(var_a), _f(var_b), out.ref()),
                    "denoiseImageGuided")

    def guidedFilter(self, inp, filter_size, eps, out):
        self._check(lib().clim_guided_filter(self.h, inp.ref(), int(filter_size), _f(eps), out.ref()),
                    "guidedFilter")

    def reassembleImage(self, d0, n1, d1, sharpening, nlf, out):
        self._check(lib().clim_reassemble_image(self.h, d0.ref(), n1.ref(), d1.ref(), C.c_float(sharpening),
                                                _f(nlf), out.ref()), "reassembleImage")

    def convertTosRGB(self, linear, out8, transform, params, out_f32=None, ltm_mask=None):
        self._check(lib().clim_convert_to_srgb(self.h, linear.ref(),
                                               ltm_mask.ref() if ltm_mask is not None else None, out8.ref(),
                                               out_f32.ref() if out_f32 is not None else None, _f(transform),
                                               C.byref(params)), "convertTosRGB")

    def fasteDebayer(self, raw, rgba, bayer):
        self._check(lib().clim_fast_debayer(self.h, raw.ref(), rgba.ref(), int(bayer)), "fasteDebayer")

    def fast_demosaic(self, raw, out8, params, out_f32=None):
        """RawConverter::fastDemosaicImage: half-resolution RGBA8 (and float)."""
        self._check(lib().clim_fast_demosaic(self.h, raw.ref(), out8.ref(),
                                             out_f32.ref() if out_f32 is not None else None, C.byref(params)),
                    "fastDemosaicImage")

    def bayerToRawRGBA(self, raw, rgba, bayer):
        self._check(lib().clim_bayer_to_raw_rgba(self.h, raw.ref(), rgba.ref(), int(bayer)), "bayerToRawRGBA")

    def rawRGBAToBayer(self, rgba, raw, bayer):
        self._check(lib().clim_raw_rgba_to_bayer(self.h, rgba.ref(), raw.ref(), int(bayer)), "rawRGBAToBayer")

    def despeckleRawRGBAImage(self, inp, out):
        self._check(lib().clim_despeckle_raw_rgba(self.h, inp.ref(), out.ref()), "despeckleRawRGBAImage")

    def denoiseRawRGBAImage(self, inp, raw_variance, out):
        self._check(lib().clim_denoise_raw_rgba(self.h, inp.ref(), _f(raw_variance), out.ref()),
                    "denoiseRawRGBAImage")

    def localToneMappingMask(self, inp, guide, ltm, transform, out):
        """ltm: LTMParameters; transform: YCbCr -> camera RGB (3x3)."""
        self._check(lib().clim_local_tone_mapping_mask(self.h, inp.ref(), guide.ref(), C.byref(ltm),
                                                       _f(transform), out.ref()), "localToneMappingMask")

    def pyramidalDenoise(self, image, levels, algorithm, params, out, nlf=None):
        dp = (DenoiseParameters * 5)(*params)
        buf = _f(np.zeros(30) if nlf is None else np.asarray(nlf, np.float32).ravel().tolist()
                 + [0.0] * (30 - 6 * levels))
        self._check(lib().clim_pyramidal_denoise(self.h, image.ref(), levels, algorithm, dp, out.ref(), buf),
                    "pyramidalDenoise")
        return np.array(buf[:6 * levels], np.float32).reshape(levels, 6)

    def pyramidalDenoiseInto(self, image, levels, algorithm, params, image_pyramid, denoised_pyramid, nlf=None):
        """clim_pyramidal_denoise_into: image_pyramid (levels - 1 images, level
        i + 1 at index i) and denoised_pyramid (levels images) are filled."""
        dp = (DenoiseParameters * 5)(*params)
        buf = _f(np.zeros(30) if nlf is None else np.asarray(nlf, np.float32).ravel().tolist()
                 + [0.0] * (30 - 6 * levels))
        lv = (C.POINTER(Image) * len(image_pyramid))(*[C.pointer(im.im) for im in image_pyramid])
        dn = (C.POINTER(Image) * len(denoised_pyramid))(*[C.pointer(im.im) for im in denoised_pyramid])
        self._check(lib().clim_pyramidal_denoise_into(self.h, image.ref(), levels, algorithm, dp, lv, dn, buf),
                    "pyramidalDenoiseInto")
        return np.array(buf[:6 * levels], np.float32).reshape(levels, 6)

    def raw_to_rgb(self, raw, out8, params, out_f32=None, writeback=True, rotate_180=False):
        """RawConverter::demosaicImage (rotate_180: clim_raw_to_rgb_ex).  With
        writeback, blocks and returns the measured noise model (rawNlf,
        pyramidNlf)."""
        nm = NoiseModel() if writeback else None
        self._check(lib().clim_raw_to_rgb_ex(self.h, raw.ref(), out8.ref(),
                                             out_f32.ref() if out_f32 is not None else None, C.byref(params),
                                             C.byref(nm) if writeback else None, int(bool(rotate_180))),
                    "raw_to_rgb")
        if not writeback:
            return None
        return (np.array(nm.rawNlf[:], np.float32),
                np.array([nm.pyramidNlf[i][:] for i in range(params.levels)], np.float32))

    def raw_to_rgb_batch(self, raws, outs8, params, outs_f32=None, writeback=False):
        """Several frames (clim_raw_to_rgb_batch): raws, outs8, params (and
        outs_f32) are lists of equal length.  With writeback, blocks and
        returns each frame's measured noise model as (rawNlf, pyramidNlf)."""
        n = len(raws)
        if len(outs8) != n or len(params) != n or (outs_f32 is not None and len(outs_f32) != n):
            raise ValueError("raw_to_rgb_batch: lists of different lengths")
        R = (C.POINTER(Image) * n)(*[C.pointer(r.im) for r in raws])
        O8 = (C.POINTER(Image) * n)(*[C.pointer(o.im) for o in outs8])
        OF = None
        if outs_f32 is not None:
            OF = (C.POINTER(Image) * n)(*[C.pointer(o.im) if o is not None else None for o in outs_f32])
        P = (DemosaicParameters * n)(*params)
        NM = (NoiseModel * n)() if writeback else None
        self._check(lib().clim_raw_to_rgb_batch(self.h, n, R, O8, OF, P, NM), "raw_to_rgb_batch")
        if not writeback:
            return None
        return [(np.array(NM[i].rawNlf[:], np.float32),
                 np.array([NM[i].pyramidNlf[l][:] for l in range(params[i].levels)], np.float32))
                for i in range(n)]

    # ---- learned denoise strength ----------------------------------------
    def strength_set_weights(self, params):
        p = np.ascontiguousarray(params, np.float32)
        self._check(lib().clim_strength_set_weights(self.h, p.ctypes.data_as(C.c_void_p), C.c_int64(p.size)),
                    "strength_set_weights")

    def strength_get_weights(self):
        p = np.zeros(STRENGTH_PARAMS, np.float32)
        self._check(lib().clim_strength_get_weights(self.h, p.ctypes.data_as(C.c_void_p), C.c_int64(p.size)),
                    "strength_get_weights")
        return p

    def strength_set_mode(self, mode):
        self._check(lib().clim_strength_set_mode(self.h, int(mode)), "strength_set_mode")

    def strength_gate_read(self):
        o = (C.c_int32 * 5)()
        m = (C.c_float * 5)()
        self._check(lib().clim_strength_gate_read(self.h, o, m), "strength_gate_read")
        return np.array(o[:], np.int32), np.array(m[:], np.float32)

    def strength_map_read(self):
        gw, gh = C.c_int32(), C.c_int32()
        self._check(lib().clim_strength_map_read(self.h, None, C.c_int64(0), C.byref(gw), C.byref(gh)),
                    "strength_map_read")
        m = np.zeros((5, gh.value, gw.value), np.float32)
        self._check(lib().clim_strength_map_read(self.h, m.ctypes.data_as(C.c_void_p), C.c_int64(m.size),
                                                 C.byref(gw), C.byref(gh)), "strength_map_read")
        return m

    def strength_loss_grad(self, params, feat, curves, norm):
        """(loss (2,), gradient (STRENGTH_PARAMS,), raw outputs (n, h, w, 32))."""
        p = np.ascontiguousarray(params, np.float32)
        f = np.ascontiguousarray(feat, np.float32)
        n, h, w, _ = f.shape
        cv = np.ascontiguousarray(curves, np.float32).reshape(n, h, w, 70)
        nm = np.ascontiguousarray(norm, np.float32).reshape(n, 10)
        out = np.zeros((n, h, w, 32), np.float32)
        g = np.zeros(STRENGTH_PARAMS, np.float32)
        loss = (C.c_double * 2)()
        ptr = lambda a: a.ctypes.data_as(C.c_void_p)
        self._check(lib().clim_strength_loss_grad(self.h, ptr(p), ptr(f), ptr(cv), ptr(nm), n, h, w, ptr(out),
                                                  ptr(g), loss), "strength_loss_grad")
        return np.array(loss[:]), g, out

    def strength_train(self, frames, width=4032, height=3024, crop=32, epochs=10, batch=32, lr=1e-3, seed=10000,
                       bayer=RGGB, noise_levels=6, verbose=False):
        """clim_strength_train; returns the (epochs, 2) sampled losses."""
        cfg = StrengthTrainConfig(frames, width, height, crop, epochs, batch, lr, seed, bayer, noise_levels,
                                  int(verbose))
        log = np.zeros(2 * epochs, np.float32)
        self._check(lib().clim_strength_train(self.h, C.byref(cfg), log.ctypes.data_as(C.c_void_p),
                                              C.c_int64(log.size)), "strength_train")
        return log.reshape(epochs, 2)

    def generate_bayer(self, raw, bayer, scene_seed, noise_seed, a, b, black, white, clean=None):
        self._check(lib().clim_generate_bayer(self.h, raw.ref() if raw is not None else None,
                                              clean.ref() if clean is not None else None, bayer,
                                              C.c_uint32(scene_seed), C.c_uint32(noise_seed), C.c_float(a),
                                              C.c_float(b), C.c_float(black), C.c_float(white)), "generate_bayer")
