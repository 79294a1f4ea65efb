"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

`Oracle(double=False)` exposes the FP32 restatement (o32_*), `double=True`
the FP64 one (o64_*).  Images are numpy arrays: luma (H, W), RGBA (H, W, 4).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# CLIM_ORACLE_LIB: another build of the same sources (the sanitizer build of
# `make -C oracle sanitize`, loaded by tests/test_oracle_sanitize.py)
LIB_PATH = os.environ.get("CLIM_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = C.CDLL(LIB_PATH)
    return _lib


class DenoiseParams(C.Structure):
    _fields_ = [("luma", C.c_float), ("chroma", C.c_float), ("chromaBoost", C.c_float),
                ("gradientBoost", C.c_float), ("sharpening", C.c_float)]


class PipelineParams(C.Structure):
    _fields_ = [("bayer", C.c_int), ("black_level", C.c_float), ("white_level", C.c_float),
                ("exposure_multiplier", C.c_float), ("scale_mul", C.c_float * 4),
                ("rgb_cam", C.c_float * 9), ("rawNlf", C.c_float * 4),
                ("pyramidNlf", C.c_float * 30), ("dp", DenoiseParams * 5),
                ("contrast", C.c_float), ("saturation", C.c_float), ("toneCurveSlope", C.c_float),
                ("exposureBias", C.c_float), ("blacks", C.c_float), ("levels", C.c_int),
                ("algorithm", C.c_int), ("localToneMapping", C.c_int), ("ltm", C.c_float * 4),
                ("noiseLevel", C.c_float), ("rotate_180", C.c_int), ("fp16_storage", C.c_int)]


# pyramidNlf row 0 of cpp/raw_pipeline.cpp:105-111 ("minimal noise model")
DEFAULT_PYRAMID_NLF = [
    [2.1e-04, 1.0e-08, 8.2e-06, 2.6e-04, 3.9e-05, 9.0e-06],
    [1.6e-04, 1.0e-08, 1.6e-06, 6.1e-04, 1.8e-04, 6.1e-05],
    [2.5e-04, 1.0e-08, 1.0e-08, 6.0e-04, 3.6e-04, 1.3e-04],
    [4.2e-04, 1.0e-08, 1.0e-08, 4.2e-04, 8.8e-04, 3.4e-04],
    [3.0e-04, 1.0e-08, 1.0e-08, 1.0e-03, 1.8e-03, 4.2e-04],
]
DEFAULT_RAW_NLF = [2.8482e-04, 3.6848e-04, 3.3079e-04, 3.2821e-04]


def default_params(bayer=2, black=64.0, white=4095.0, levels=5, algorithm=0):
    p = PipelineParams()
    p.bayer = bayer
    p.black_level = black
    p.white_level = white
    p.exposure_multiplier = 1.0
    sm = np.float32(65535.0) / np.float32(white - black)
    for c in range(4):
        p.scale_mul[c] = float(sm)
    for k in range(9):
        p.rgb_cam[k] = 1.0 if k in (0, 4, 8) else 0.0
    for c in range(4):
        p.rawNlf[c] = DEFAULT_RAW_NLF[c]
    for i in range(5):
        for k in range(6):
            p.pyramidNlf[i * 6 + k] = DEFAULT_PYRAMID_NLF[i][k]
        p.dp[i] = DenoiseParams(1.0, 1.0, 1.0, 1.0, 1.0)
    p.contrast, p.saturation, p.toneCurveSlope, p.exposureBias, p.blacks = 1.05, 1.0, 3.5, 0.0, 0.0
    p.levels = levels
    p.algorithm = algorithm
    # LTMParameters defaults, cpp/demosaic.hpp:54-59
    p.localToneMapping = 0
    for k, v in enumerate((0.01, 1.25, 1.0, 1.1)):
        p.ltm[k] = v
    return p


def _f(a):
    return (C.c_float * len(a))(*[float(x) for x in a])


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def generate_bayer(w, h, bayer, scene_seed, noise_seed, a, b, black, white, clean=False):
    """Seeded synthetic mosaic (uint16 (H, W)) and optionally the clean RGBA scene."""
    L = lib()
    raw = np.zeros((h, w), np.uint16)
    cl = np.zeros((h, w, 4), np.float32) if clean else None
    L.oscene_generate(C.c_int(w), C.c_int(h), C.c_int(bayer), C.c_uint32(scene_seed),
                      C.c_uint32(noise_seed), C.c_float(a), C.c_float(b), C.c_float(black),
                      C.c_float(white), _ptr(raw), _ptr(cl) if clean else None)
    return (raw, cl) if clean else raw


class Oracle:
    """lib_path: another build of the same sources (the speed-built CPU
    baseline of `make -C oracle fast`)."""

    def __init__(self, double=False, lib_path=None):
        self.double = double
        self.dt = np.float64 if double else np.float32
        self.pfx = "o64_" if double else "o32_"
        self.L = C.CDLL(lib_path) if lib_path else lib()

    def _fn(self, name):
        return getattr(self.L, self.pfx + name)

    def _new(self, *shape):
        return np.zeros(shape, self.dt)

    def _chk(self, a, shape=None):
        a = np.ascontiguousarray(a, dtype=self.dt)
        if shape is not None:
            assert a.shape == shape, (a.shape, shape)
        return a

    # --- stages -----------------------------------------------------------
    def scale_raw(self, raw, bayer, scale_mul, black_norm):
        h, w = raw.shape
        raw = np.ascontiguousarray(raw, np.uint16)
        out = self._new(h, w)
        self._fn("scale_raw")(_ptr(raw), w, h, bayer, _f(scale_mul), C.c_float(black_norm), _ptr(out))
        return out

    def raw_noise_statistics(self, scaled, bayer):
        h, w = scaled.shape
        scaled = self._chk(scaled)
        mean = self._new(h // 2, w // 2, 4)
        var = self._new(h // 2, w // 2, 4)
        self._fn("raw_noise_statistics")(_ptr(scaled), w, h, bayer, _ptr(mean), _ptr(var))
        return mean, var

    def nlf_fit_raw(self, mean, var):
        mean, var = self._chk(mean), self._chk(var)
        out = (C.c_float * 8)()
        rc = self._fn("nlf_fit")(4, 4, _ptr(mean), 4, _ptr(var), 4, C.c_long(mean.shape[0] * mean.shape[1]), out)
        return rc, np.array(out[:], np.float32)

    def nlf_fit_level(self, stats):
        stats = self._chk(stats)
        out = (C.c_float * 6)()
        base = stats.ctypes.data
        rc = self._fn("nlf_fit")(3, 1, C.c_void_p(base), 4, C.c_void_p(base + stats.itemsize), 4,
                                 C.c_long(stats.shape[0] * stats.shape[1]), out)
        return rc, np.array(out[:], np.float32)

    def interpolate_green(self, scaled, bayer, luma_var):
        h, w = scaled.shape
        scaled = self._chk(scaled)
        out = self._new(h, w)
        self._fn("interpolate_green")(_ptr(scaled), w, h, bayer, C.c_float(luma_var), _ptr(out))
        return out

    def interpolate_red_blue(self, scaled, green, bayer, chroma_var, rotate_180=0):
        h, w = scaled.shape
        scaled, green = self._chk(scaled), self._chk(green)
        out = self._new(h, w, 4)
        self._fn("interpolate_red_blue")(_ptr(scaled), _ptr(green), w, h, bayer, C.c_float(chroma_var),
                                         int(rotate_180), _ptr(out))
        return out

    def blend_highlights(self, img, clip=1.0):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("blend_highlights")(_ptr(img), w, h, C.c_float(clip), _ptr(out))
        return out

    def transform(self, img, m):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("transform")(_ptr(img), w, h, _f(np.asarray(m).ravel()), _ptr(out))
        return out

    def despeckle(self, img, var_a, var_b):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("despeckle")(_ptr(img), w, h, _f(var_a), _f(var_b), _ptr(out))
        return out

    def downsample(self, img, ow, oh):
        ih, iw, _ = img.shape
        img = self._chk(img)
        out = self._new(oh, ow, 4)
        self._fn("downsample")(_ptr(img), iw, ih, _ptr(out), ow, oh)
        return out

    def noise_statistics(self, img):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("noise_statistics")(_ptr(img), w, h, _ptr(out))
        return out

    def denoise_bilateral(self, img, var_a, var_b, chroma_boost=1.0, gradient_boost=1.0):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("denoise_bilateral")(_ptr(img), w, h, _f(var_a), _f(var_b), C.c_float(chroma_boost),
                                      C.c_float(gradient_boost), _ptr(out))
        return out

    def denoise_guided_fast(self, img, var_a, var_b):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("denoise_guided_fast")(_ptr(img), w, h, _f(var_a), _f(var_b), _ptr(out))
        return out

    def guided_filter(self, img, fs, eps):
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        self._fn("guided_filter")(_ptr(img), w, h, int(fs), _f(eps), _ptr(out))
        return out

    def reassemble(self, d0, n1, d1, sharpening, nlf):
        h, w, _ = d0.shape
        h1, w1, _ = n1.shape
        d0, n1, d1 = self._chk(d0), self._chk(n1), self._chk(d1)
        out = self._new(h, w, 4)
        self._fn("reassemble")(_ptr(d0), _ptr(n1), _ptr(d1), w, h, w1, h1, C.c_float(sharpening),
                               _f(nlf), _ptr(out))
        return out

    def convert_to_srgb(self, img, transform, contrast=1.05, saturation=1.0, slope=3.5, exposure_bias=0.0,
                        blacks=0.0, ltm_mask=None):
        h, w, _ = img.shape
        img = self._chk(img)
        mask = self._chk(ltm_mask, (h, w)) if ltm_mask is not None else None
        out8 = np.zeros((h, w, 4), np.uint8)
        outf = self._new(h, w, 4)
        self._fn("convert_to_srgb")(_ptr(img), w, h, _f(np.asarray(transform).ravel()), C.c_float(contrast),
                                    C.c_float(saturation), C.c_float(slope), C.c_float(exposure_bias),
                                    C.c_float(blacks), _ptr(mask) if mask is not None else None, _ptr(out8),
                                    _ptr(outf))
        return out8, outf

    def fast_debayer(self, raw, bayer):
        h, w = raw.shape
        raw = self._chk(raw)
        out = self._new(h // 2, w // 2, 4)
        self._fn("fast_debayer")(_ptr(raw), w, h, bayer, _ptr(out))
        return out

    def fast_demosaic(self, raw, params):
        """RawConverter::fastDemosaicImage: (rgba8, float) at half resolution."""
        h, w = raw.shape
        raw = np.ascontiguousarray(raw, np.uint16)
        out8 = np.zeros((h // 2, w // 2, 4), np.uint8)
        outf = self._new(h // 2, w // 2, 4)
        self._fn("fast_demosaic")(_ptr(raw), w, h, C.byref(params), _ptr(out8), _ptr(outf))
        return out8, outf

    def bayer_to_raw_rgba(self, raw, bayer):
        h, w = raw.shape
        raw = self._chk(raw)
        out = self._new(h // 2, w // 2, 4)
        self._fn("bayer_to_raw_rgba")(_ptr(raw), w, h, bayer, _ptr(out))
        return out

    def raw_rgba_to_bayer(self, rgba, bayer):
        qh, qw, _ = rgba.shape
        rgba = self._chk(rgba)
        out = self._new(2 * qh, 2 * qw)
        self._fn("raw_rgba_to_bayer")(_ptr(rgba), qw, qh, bayer, _ptr(out))
        return out

    def despeckle_raw_rgba(self, rgba):
        qh, qw, _ = rgba.shape
        rgba = self._chk(rgba)
        out = self._new(qh, qw, 4)
        self._fn("despeckle_raw_rgba")(_ptr(rgba), qw, qh, _ptr(out))
        return out

    def denoise_raw_rgba(self, rgba, raw_variance):
        qh, qw, _ = rgba.shape
        rgba = self._chk(rgba)
        out = self._new(qh, qw, 4)
        self._fn("denoise_raw_rgba")(_ptr(rgba), qw, qh, _f(raw_variance), _ptr(out))
        return out

    def ltm_mask(self, img, guide, ltm, transform, var_eps=False):
        """localToneMappingMask: img the denoised YCbCr image (H, W, 4), guide
        (gh, gw, 4); ltm = (guidedFilterEps, shadows, highlights, detail).
        With var_eps, also the per-pixel denominator var + eps."""
        h, w, _ = img.shape
        gh, gw, _ = guide.shape
        img, guide = self._chk(img), self._chk(guide)
        out = self._new(h, w)
        ve = self._new(h, w) if var_eps else None
        self._fn("ltm_mask")(_ptr(img), w, h, _ptr(guide), gw, gh, _f(ltm), _f(np.asarray(transform).ravel()),
                             _ptr(out), _ptr(ve) if var_eps else None)
        return (out, ve) if var_eps else out

    def cam_ycbcr(self, rgb_cam):
        out = (C.c_float * 9)()
        self._fn("cam_ycbcr")(_f(np.asarray(rgb_cam).ravel()), out)
        return np.array(out[:], np.float32).reshape(3, 3)

    def inverse3x3(self, m, scale=1.0):
        out = (C.c_float * 9)()
        self._fn("inverse3x3").argtypes = [C.c_void_p, C.c_double, C.c_void_p]
        rc = self._fn("inverse3x3")(_f(np.asarray(m).ravel()), C.c_double(scale), out)
        return rc, np.array(out[:], np.float32).reshape(3, 3)

    def half_round(self, arr):
        """Values rounded to FP16 storage (the mixed-precision form's level images)."""
        a = np.array(arr, dtype=self.dt, copy=True)
        self._fn("half_round_img")(_ptr(a), C.c_size_t(a.size))
        return a

    def pyramidal_denoise(self, img, levels, algorithm, params, nlf_in, level3=False):
        """Returns (fallback mask, denoised level 0, fitted NLF) and, with
        level3, the reassembled level 3 (the local tone mapping guide)."""
        h, w, _ = img.shape
        img = self._chk(img)
        out = self._new(h, w, 4)
        g = self._new(h >> 3, w >> 3, 4) if level3 and levels > 3 else None
        dp = (DenoiseParams * 5)(*params)
        nlf = _f(np.asarray(nlf_in, np.float32).ravel())
        rc = self._fn("pyramidal_denoise")(_ptr(img), w, h, levels, algorithm, dp, nlf, _ptr(out),
                                           _ptr(g) if g is not None else None)
        res = (rc, out, np.array(nlf[:], np.float32).reshape(-1, 6))
        return res + (g,) if level3 else res

    def raw_to_rgb(self, raw, params, dumps=False, strength=None):
        """strength (optional): (map (5, gh, gw) float32, open (5,) ints), the
        learned-strength multipliers of the level filters (plane l: level l)."""
        h, w = raw.shape
        raw = np.ascontiguousarray(raw, np.uint16)
        out8 = np.zeros((h, w, 4), np.uint8)
        outf = self._new(h, w, 4)
        nlf = (C.c_float * 34)()
        ycc = self._new(h, w, 4) if dumps else None
        den = self._new(h, w, 4) if dumps else None
        if strength is None:
            rc = self._fn("raw_to_rgb")(_ptr(raw), w, h, C.byref(params), _ptr(out8), _ptr(outf), nlf,
                                        _ptr(ycc) if dumps else None, _ptr(den) if dumps else None)
        else:
            smap = np.ascontiguousarray(strength[0], np.float32)
            sopen = np.ascontiguousarray(strength[1], np.int32)
            assert smap.ndim == 3 and smap.shape[0] == 5 and sopen.shape == (5,)
            rc = self._fn("raw_to_rgb_strength")(_ptr(raw), w, h, C.byref(params), _ptr(out8), _ptr(outf), nlf,
                                                 _ptr(ycc) if dumps else None, _ptr(den) if dumps else None,
                                                 _ptr(smap), smap.shape[2], smap.shape[1], _ptr(sopen))
        nl = np.array(nlf[:], np.float32)
        res = {"rc": rc, "rgba8": out8, "rgbf": outf, "rawNlf": nl[:4],
               "pyramidNlf": nl[4:4 + 6 * params.levels].reshape(params.levels, 6)}
        if dumps:
            res["ycc"] = ycc
            res["den0"] = den
        return res
