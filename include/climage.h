/*
 * climage.h - C ABI of the MI355X (gfx950) RAW->RGB path.
 *
 * This is the boundary the reference's OpenCL operator layer is replaced at.
 * Every entry point below names the reference function it stands in for
 * (paths relative to the reference tree, cpp/ = CLImage/app/src/main/cpp/,
 * cl/ = cpp/OpenCL/).  The C++ typed wrappers in gls_hip_image.hpp,
 * demosaic.hpp, pyramidal_denoise.hpp and raw_converter.hpp call these and
 * keep the reference's names and argument order.
 *
 * Conventions
 *  - Images are device-resident, pitch-linear (clim_image).  Texel formats
 *    follow the reference textures (src/gls_cl_image.hpp:232-292): R16 raw
 *    reads as value/65535 (UNORM16); luma and RGBA float images are FP32
 *    (the reference's USE_FP16_FLOATS=0 build); RGBA8 is UNORM8.
 *  - All calls are asynchronous on the context's HIP stream, except those
 *    documented as blocking (copies, the *_nlf fits that return host values,
 *    and clim_raw_to_rgb when it is asked to write the measured noise model
 *    back).  clim_sync() waits for the stream.
 *  - Errors are status codes, never exceptions; clim_last_error() gives the
 *    message of the last failing call on a context.
 *  - One context per host thread; a context is not thread-safe.
 *  - Image borders: every stencil clamps coordinates to the image (the
 *    reference leaves sampler-less out-of-bounds reads undefined).
 */
#ifndef CLIMAGE_H
#define CLIMAGE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct clim_ctx clim_ctx;

typedef enum {
    CLIM_OK = 0,
    CLIM_EINVAL = 1,     /* bad argument: null pointer, size or format mismatch */
    CLIM_ENOMEM = 2,     /* device or pinned-host allocation failed */
    CLIM_EHIP = 3,       /* a HIP runtime call failed */
    CLIM_ENUMERIC = 4,   /* singular matrix, empty noise-model sample */
    CLIM_ENODEV = 5      /* no usable gfx950 device */
} clim_status;

typedef enum {
    CLIM_FMT_R16U = 1,      /* uint16 raw mosaic, read as UNORM16 */
    CLIM_FMT_R32F = 2,      /* float luma */
    CLIM_FMT_RGBA32F = 3,   /* float4 */
    CLIM_FMT_RGBA8 = 4,     /* UNORM8 x4 */
    CLIM_FMT_RG32F = 5,     /* float2 (guided-filter scratch) */
    CLIM_FMT_RGBA16F = 6    /* half4: the mixed-precision pyramid's level storage */
} clim_format;

/* device image, pitch-linear; pitch_bytes >= width * texel size */
typedef struct {
    void* data;
    int32_t width;
    int32_t height;
    int32_t pitch_bytes;
    int32_t format;
} clim_image;

/* cpp/demosaic.hpp:23-28, identical numbering */
typedef enum { CLIM_GRBG = 0, CLIM_GBRG = 1, CLIM_RGGB = 2, CLIM_BGGR = 3 } clim_bayer_pattern;

/* cpp/pyramidal_denoise.hpp:38 */
typedef enum {
    CLIM_DENOISE_BILATERAL = 0,
    CLIM_DENOISE_GUIDED_FAST = 1,
    CLIM_DENOISE_GUIDED_PRECISE = 2
} clim_denoise_algorithm;

/* cpp/demosaic.hpp:32-38 (defaults 1,1,1,1,1) */
typedef struct {
    float luma;
    float chroma;
    float chromaBoost;
    float gradientBoost;
    float sharpening;
} clim_denoise_params;

/* cpp/demosaic.hpp:40-47, layout of cl/demosaic.cl:1423-1430
 * (defaults contrast 1.05, saturation 1, toneCurveSlope 3.5, 0, 0, 0) */
typedef struct {
    float contrast;
    float saturation;
    float toneCurveSlope;
    float exposureBias;
    float blacks;
    int32_t localToneMapping;
} clim_rgb_conversion_params;

/* cpp/demosaic.hpp:49-52 */
typedef struct {
    float rawNlf[4];
    float pyramidNlf[5][6];   /* per level: A_Y, A_Cb, A_Cr, B_Y, B_Cb, B_Cr */
} clim_noise_model;

/* cpp/demosaic.hpp:54-59 */
typedef struct {
    float guidedFilterEps;
    float shadows;
    float highlights;
    float detail;
} clim_ltm_params;

#define CLIM_MAX_LEVELS 5

/* cpp/demosaic.hpp:61-81, plus the two knobs the reference fixes at compile
 * time: the pyramid depth (PyramidalDenoise<5>, cpp/raw_converter.hpp:33) and
 * the per-level filter (constructor argument, cpp/pyramidal_denoise.hpp:45). */
typedef struct {
    int32_t bayerPattern;
    float black_level;                 /* DN */
    float white_level;                 /* DN */
    float exposure_multiplier;
    float scale_mul[4];
    float rgb_cam[3][3];
    clim_noise_model noiseModel;       /* pyramidNlf[0] is read by despeckle */
    clim_denoise_params denoiseParameters[CLIM_MAX_LEVELS];
    float noiseLevel;                  /* > 0.6: raw-domain pre-denoise (cpp/raw_converter.cpp:82) */
    clim_rgb_conversion_params rgbConversionParameters;
    clim_ltm_params ltmParameters;
    int32_t levels;                    /* 2..5 */
    int32_t denoiseAlgorithm;          /* clim_denoise_algorithm */
} clim_demosaic_params;

/* ------------------------------------------------------------ context */
clim_status clim_create(clim_ctx** ctx, int device);
void clim_destroy(clim_ctx* ctx);
const char* clim_last_error(const clim_ctx* ctx);
const char* clim_status_string(clim_status s);
clim_status clim_sync(clim_ctx* ctx);
/* the context's stream, as a hipStream_t cast to void* */
void* clim_stream(clim_ctx* ctx);

/* ---------------------------------------------------------- precision
 * CLIM_PRECISION_FP32 (default): every stage in FP32 (the FP64 fits and 3x3
 * inversions aside), bit for bit the FP32 restatement of the reference.
 * CLIM_PRECISION_MIXED: the level images of the pyramid (the despeckled
 * YCbCr level 0, the downsampled and the denoised levels) are stored as FP16
 * (CLIM_FMT_RGBA16F), as the reference's ARM build stores every intermediate
 * image (src/gls_image.hpp:32,189-210), and the bilateral denoiser's range
 * test and weights run in packed FP16 with the weighted sums accumulated in
 * FP32; the noise-model regressions (FP64 over FP32 statistics), the colour
 * matrices (FP64) and the other stages' arithmetic stay as in the FP32 form.
 * Applies to the pipeline calls (clim_raw_to_rgb*, bursts created later) on
 * this context. */
typedef enum { CLIM_PRECISION_FP32 = 0, CLIM_PRECISION_MIXED = 1 } clim_precision;
clim_status clim_set_precision(clim_ctx* ctx, int32_t precision);

/* -------------------------------------------------- images and copies
 * cl_image_2d constructor / copyPixelsFrom / copyPixelsTo
 * (src/gls_cl_image.hpp:53-90).  Copies are blocking like the reference's. */
clim_status clim_image_alloc(clim_ctx* ctx, int32_t width, int32_t height, int32_t format, clim_image* out);
clim_status clim_image_free(clim_ctx* ctx, clim_image* img);
clim_status clim_copy_from_host(clim_ctx* ctx, const void* host, int64_t host_pitch_bytes, clim_image* dst);
clim_status clim_copy_to_host(clim_ctx* ctx, const clim_image* src, void* host, int64_t host_pitch_bytes);
int32_t clim_format_texel_bytes(int32_t format);

/* -------------------------------------------------------- operators */

/* scaleRawData, cpp/demosaic_cl.cpp:26-44, cl/demosaic.cl:69-78.
 * black_level is normalised (DN / 65535). */
clim_status clim_scale_raw_data(clim_ctx* ctx, const clim_image* raw, clim_image* scaled,
                                int32_t bayer, const float scale_mul[4], float black_level);

/* rawNoiseStatistics, cpp/demosaic_cl.cpp:107-128, cl/demosaic.cl:1288-1324.
 * mean, var: (W/2)x(H/2) RGBA32F, channels in (R, G, B, G2) order of the
 * OpenCL Bayer table (cl/demosaic.cl:29-34). */
clim_status clim_raw_noise_statistics(clim_ctx* ctx, const clim_image* scaled, int32_t bayer,
                                      clim_image* mean, clim_image* var);

/* computeRawNoiseStatistics, cpp/pyramidal_denoise.cpp:259-362: statistics
 * plus the three-pass FP64 regression, all on the device.  Blocking; writes
 * A[4], B[4]. */
clim_status clim_compute_raw_noise_statistics(clim_ctx* ctx, const clim_image* scaled, int32_t bayer,
                                              float nlf_out[8]);

/* interpolateGreen, cpp/demosaic_cl.cpp:46-63, cl/demosaic.cl:160-227 */
clim_status clim_interpolate_green(clim_ctx* ctx, const clim_image* scaled, clim_image* green,
                                   int32_t bayer, float luma_variance);

/* interpolateRedBlue, cpp/demosaic_cl.cpp:65-85, cl/demosaic.cl:229-318 */
clim_status clim_interpolate_red_blue(clim_ctx* ctx, const clim_image* scaled, const clim_image* green,
                                      clim_image* rgb, int32_t bayer, float chroma_variance,
                                      int32_t rotate_180);

/* blendHighlightsImage, cpp/demosaic_cl.cpp:506-522, cl/demosaic.cl:338-380
 * (in-place allowed) */
clim_status clim_blend_highlights(clim_ctx* ctx, const clim_image* in, float clip, clim_image* out);

/* transformImage, cpp/demosaic_cl.cpp:210-234, cl/demosaic.cl:728-733
 * (row-major 3x3; in-place allowed) */
clim_status clim_transform_image(clim_ctx* ctx, const clim_image* in, clim_image* out,
                                 const float transform[9]);

/* despeckleImage, cpp/demosaic_cl.cpp:267-287, cl/demosaic.cl:384-479 */
clim_status clim_despeckle_image(clim_ctx* ctx, const clim_image* in, const float var_a[3],
                                 const float var_b[3], clim_image* out);

/* resampleImage("downsampleImage"), cpp/demosaic_cl.cpp:157-178,
 * cl/demosaic.cl:853-865 */
clim_status clim_downsample_image(clim_ctx* ctx, const clim_image* in, clim_image* out);

/* applyKernel("noiseStatistics"), cpp/demosaic_cl.cpp:130-145,
 * cl/demosaic.cl:1267-1286: out = (mean Y, var Y, var Cb, var Cr) */
clim_status clim_noise_statistics(clim_ctx* ctx, const clim_image* in, clim_image* out);

/* computeNoiseStatistics, cpp/pyramidal_denoise.cpp:161-256 (device-side
 * fit).  Blocking; writes A_Y, A_Cb, A_Cr, B_Y, B_Cb, B_Cr. */
clim_status clim_compute_noise_statistics(clim_ctx* ctx, const clim_image* in, float nlf_out[6]);

/* denoiseImage, cpp/demosaic_cl.cpp:292-315, cl/demosaic.cl:735-765 */
clim_status clim_denoise_image(clim_ctx* ctx, const clim_image* in, const float var_a[3],
                               const float var_b[3], float chroma_boost, float gradient_boost,
                               clim_image* out);

/* denoiseImageGuided, cpp/demosaic_cl.cpp:317-337, cl/demosaic.cl:767-851 */
clim_status clim_denoise_image_guided(clim_ctx* ctx, const clim_image* in, const float var_a[3],
                                      const float var_b[3], clim_image* out);

/* GuidedFilter::filter, cpp/guided_filter.cpp:195-225, cl/guided_filter.cl.
 * Scratch comes from the context. */
clim_status clim_guided_filter(clim_ctx* ctx, const clim_image* in, int32_t filter_size,
                               const float eps[3], clim_image* out);

/* reassembleImage, cpp/demosaic_cl.cpp:180-208, cl/demosaic.cl:867-896
 * (out may alias denoised0) */
clim_status clim_reassemble_image(clim_ctx* ctx, const clim_image* denoised0, const clim_image* image1,
                                  const clim_image* denoised1, float sharpening, const float nlf[2],
                                  clim_image* out);

/* convertTosRGB, cpp/demosaic_cl.cpp:236-265, cl/demosaic.cl:1380-1474.
 * out_rgba8 receives the UNORM8 result; out_f32 (nullable) the same values
 * before quantisation.  ltm_mask (R32F, the size of linear) is read only when
 * params->localToneMapping is set; otherwise it may be null. */
clim_status clim_convert_to_srgb(clim_ctx* ctx, const clim_image* linear, const clim_image* ltm_mask,
                                 clim_image* out_rgba8, clim_image* out_f32, const float transform[9],
                                 const clim_rgb_conversion_params* params);

/* ------------------------------------------------------------ DNG ingest
 * Host only: no context and no GPU needed.  Errors carry a message in
 * clim_dng_last_error() (per thread). */

/* What read_dng_file (src/gls_image_tiff.cpp:255-458) and unpackDNGMetadata
 * (cpp/demosaic_utils.cpp:280-375) read from a DNG. */
typedef struct {
    int32_t width, height;           /* of the raw image after DefaultCrop / ActiveArea */
    int32_t bayerPattern;            /* from CFAPattern (gbrg when absent, as the reference) */
    int32_t orientation;
    int32_t bits_per_sample, compression, tiled;
    float black_level;               /* BlackLevel[0], 0 when absent */
    float white_level;               /* WhiteLevel[0], 65535 when absent */
    float color_matrix1[9], color_matrix2[9];
    int32_t has_color_matrix1, has_color_matrix2;
    float as_shot_neutral[3];
    int32_t has_as_shot_neutral;
    float baseline_exposure;         /* 0 when absent */
} clim_dng_info;

/* read_dng_file, src/gls_image_tiff.cpp:255-458, from the file's bytes:
 * the raw IFD (IFD0, or the first full-resolution SubIFD), uncompressed
 * (8..16-bit) or lossless JPEG (compression 7) strips or tiles, cropped by
 * DefaultCropOrigin + ActiveArea and DefaultCropSize.  With pixels null only
 * info is filled; else pixels receives info->height rows of pitch_elems
 * uint16 samples. */
clim_status clim_dng_decode(const uint8_t* data, int64_t size, clim_dng_info* info, uint16_t* pixels,
                            int64_t pitch_elems);

/* unpackDNGMetadata, cpp/demosaic_utils.cpp:280-375, without ColorChecker
 * calibration: sets bayerPattern, black_level, white_level,
 * exposure_multiplier, scale_mul and rgb_cam (cam_xyz_coeff, :25-52) of p.
 * With auto_white_balance the raw image (host, pitch_elems per row) is
 * analysed (autoWhiteBalance, :967-1026).  exposure_out (nullable): 2^BaselineExposure;
 * highlights_out (nullable): the highlight pixel fraction of the white balance;
 * as_shot_neutral_out (nullable): the white balance used, as AsShotNeutral. */
clim_status clim_unpack_dng_metadata(const clim_dng_info* info, const uint16_t* raw, int64_t pitch_elems,
                                     int32_t auto_white_balance, clim_demosaic_params* p, float* exposure_out,
                                     float* highlights_out, float* as_shot_neutral_out);
const char* clim_dng_last_error(void);

/* ---------------------------------------------- half-resolution preview */

/* fasteDebayer, cpp/demosaic_cl.cpp:87-105, cl/demosaic.cl:320-334
 * (raw: R32F 2w x 2h scaled mosaic; rgba: RGBA32F w x h, (R, (G+G2)/2, B, 0)) */
clim_status clim_fast_debayer(clim_ctx* ctx, const clim_image* raw, clim_image* rgba, int32_t bayer);

/* RawConverter::fastDemosaicImage, cpp/raw_converter.cpp:188-224: scaleRawData,
 * fasteDebayer, blendHighlightsImage(clip 1) and convertTosRGB(rgb_cam) at half
 * resolution, one kernel.  raw: R16U w x h (even); out_rgba8: w/2 x h/2;
 * out_f32 (nullable): the same values before quantisation.  localToneMapping
 * is not applied (the reference passes this path a 1x1 placeholder mask). */
clim_status clim_fast_demosaic(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8, clim_image* out_f32,
                               const clim_demosaic_params* p);

/* ------------------------------------------------ high-noise raw branch
 * cpp/raw_converter.cpp:111-123 (noiseLevel > 0.6).  An RGBA raw image holds
 * one (R, G, B, G2) quad per texel and is half the mosaic's size. */

/* bayerToRawRGBA, cpp/demosaic_cl.cpp:372-390, cl/demosaic.cl:898-912
 * (raw: R32F 2w x 2h; rgba: RGBA32F w x h) */
clim_status clim_bayer_to_raw_rgba(clim_ctx* ctx, const clim_image* raw, clim_image* rgba, int32_t bayer);

/* rawRGBAToBayer, cpp/demosaic_cl.cpp:392-410, cl/demosaic.cl:914-929 */
clim_status clim_raw_rgba_to_bayer(clim_ctx* ctx, const clim_image* rgba, clim_image* raw, int32_t bayer);

/* despeckleRawRGBAImage, cpp/demosaic_cl.cpp:430-443, cl/demosaic.cl:530-553,1195-1201
 * (out must differ from in) */
clim_status clim_despeckle_raw_rgba(clim_ctx* ctx, const clim_image* in, clim_image* out);

/* denoiseRawRGBAImage, cpp/demosaic_cl.cpp:412-428, cl/demosaic.cl:932-949,1187-1193.
 * raw_variance: the raw noise slopes (rawNlf, R, G, B, G2); out must differ from in. */
clim_status clim_denoise_raw_rgba(clim_ctx* ctx, const clim_image* in, const float raw_variance[4],
                                  clim_image* out);

/* localToneMappingMask, cpp/demosaic_cl.cpp:339-370, cl/demosaic.cl:1024-1107.
 * input: the denoised YCbCr image (RGBA32F); guide: RGBA32F, read through its
 * first channel (the pipeline passes pyramid level 3, cpp/raw_converter.cpp:165);
 * transform: YCbCr -> camera RGB (row-major, inverse(cam_to_ycbcr) times
 * the exposure multiplier); out: R32F, the size of input. */
clim_status clim_local_tone_mapping_mask(clim_ctx* ctx, const clim_image* input, const clim_image* guide,
                                         const clim_ltm_params* ltm, const float transform[9], clim_image* out);

/* ------------------------------------------------------ host-side math */

/* cam_ycbcr, cpp/demosaic_utils.cpp:809-822 (computed in FP64, rounded to float) */
void clim_cam_ycbcr(const float rgb_cam[9], float out[9]);
/* inverse, src/gls_linalg.hpp:544-557, in FP64; CLIM_ENUMERIC when singular */
clim_status clim_inverse3x3(const float m[9], float out[9]);
/* DemosaicParameters defaults of cpp/demosaic.hpp:32-81 for a synthetic
 * identity-colour camera: rgb_cam = I, scale_mul = 65535/(white-black),
 * exposure 1, noiseLevel 0, 5 Bilateral levels, pyramidNlf row 0 of
 * cpp/raw_pipeline.cpp:105-111. */
void clim_default_params(clim_demosaic_params* p, int32_t bayer, float black_level, float white_level);

/* -------------------------------------------------------------- pipeline */

/* PyramidalDenoise<levels>::denoise, cpp/pyramidal_denoise.cpp:103-149.
 * `image` is the despeckled YCbCr level 0; the result is written to `out`.
 * Per-level noise models are measured on the device; if nlf_out is not
 * null the call blocks and writes levels x 6 floats. */
clim_status clim_pyramidal_denoise(clim_ctx* ctx, const clim_image* image, int32_t levels,
                                   int32_t algorithm, const clim_denoise_params* params,
                                   clim_image* out, float* nlf_out);

/* The same, in the caller's level images: PyramidalDenoise's imagePyramid
 * (levels - 1 images, level i + 1 at index i) and denoisedImagePyramid
 * (levels images, level 0 the result), cpp/pyramidal_denoise.hpp:37-51,
 * each width / 2^i x height / 2^i RGBA32F.  They end up holding what the
 * reference's members hold after denoise(): the downsampled levels and the
 * reassembled denoised levels (level 3 is the local tone mapping guide,
 * cpp/raw_converter.cpp:165). */
clim_status clim_pyramidal_denoise_into(clim_ctx* ctx, const clim_image* image, int32_t levels,
                                        int32_t algorithm, const clim_denoise_params* params,
                                        clim_image* const* image_pyramid, clim_image* const* denoised_pyramid,
                                        float* nlf_out);

/* RawConverter::demosaicImage, cpp/raw_converter.cpp:73-186: raw mosaic to
 * RGBA8 (and optionally the FP32 values before quantisation).  Work
 * buffers are owned by the context and resized when the frame size changes.
 * If nlf_out is not null the call blocks at the end and writes the measured
 * noise model (rawNlf = B of the raw fit, pyramidNlf per level), as the
 * reference writes it back into DemosaicParameters::noiseModel. */
clim_status clim_raw_to_rgb(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8,
                            clim_image* out_f32, const clim_demosaic_params* params,
                            clim_noise_model* nlf_out);

/* The same with RawConverter::demosaicImage's rotate_180 argument
 * (cpp/raw_converter.hpp:54-55): non-zero rotates the image by 180 degrees
 * where interpolateRedBlue writes it, cl/demosaic.cl:313-315, with the
 * reference's target (width - x, height - y) - one texel off a true
 * rotation; row 0 and column 0, which no source texel reaches, are zero
 * before the despeckle.  The rest of the path runs on the rotated image.
 * clim_raw_to_rgb is this call with rotate_180 = 0. */
clim_status clim_raw_to_rgb_ex(clim_ctx* ctx, const clim_image* raw, clim_image* out_rgba8,
                               clim_image* out_f32, const clim_demosaic_params* params,
                               clim_noise_model* nlf_out, int32_t rotate_180);

/* n frames, each as clim_raw_to_rgb(raws[i], out_rgba8[i], out_f32 ? out_f32[i]
 * : null, &params[i]) - RawConverter::demosaicImage called once per frame of a
 * burst (cpp/raw_converter.cpp:73-186), the frames of any sizes and CFA
 * patterns.  No reference counterpart as one call: for callers with several
 * frames at hand.  The frames are spread over the context's stream and two
 * further streams with work buffers of their own (created on first use), so
 * that one frame's small launches overlap the others'; the results are those of
 * the single-frame calls.  Work queued on the context before the call is
 * complete before any frame starts, and work queued after it sees every
 * output.  out_f32 may be null (or hold nulls).  nlf_out (nullable, n
 * entries): the call blocks at the end and writes each frame's measured
 * noise model. */
clim_status clim_raw_to_rgb_batch(clim_ctx* ctx, int32_t n, const clim_image* const* raws,
                                  clim_image* const* out_rgba8, clim_image* const* out_f32,
                                  const clim_demosaic_params* params, clim_noise_model* nlf_out);

/* ------------------------------------------------- bursts (recorded once)
 * n frames, each as clim_raw_to_rgb(raws[i], out_rgba8[i], out_f32 ? out_f32[i]
 * : null, &params[i]) - RawConverter::demosaicImage once per frame of a burst
 * (cpp/raw_converter.cpp:73-186), frames of any sizes and CFA patterns - set
 * up once and then run any number of times with one call each.  No reference
 * counterpart: for callers that convert bursts of the same geometry again
 * and again (new mosaics in the same input images).
 *
 * clim_burst_create validates the frames, allocates the burst's own work
 * buffers (lanes of its own, each frame's green plane and YCbCr image),
 * runs the burst once, times the candidate shapes of its two cross-frame
 * persistent launches (the green and the red/blue + despeckle stages of
 * every frame, one launch each) on these frames and keeps the fastest, then
 * records the whole burst into a HIP graph twice - with the cross-frame
 * launches, and with every frame's stages as the per-frame kernels on the
 * burst's lanes - times a replay of each and keeps the faster graph (the
 * choices are remembered in the process for bursts of the same frame
 * sizes).  Work queued on the
 * context before the call is complete first.  The caller's images must
 * outlive the burst; the context's learned-strength weights and mode are
 * copied in at creation.
 *
 * clim_burst_launch replays the graph on the context's stream (one graph
 * launch, no allocation, no host work per frame): it reads the current
 * contents of raws and writes the outputs, equal bit for bit to the
 * single-frame calls'.  Launches of one burst must not overlap (they share
 * its work buffers); the context's stream orders them.
 *
 * With CLIM_BURST_FRAME_TIMES the graph also records, after a launch,
 * when each frame's last kernel finished: clim_burst_frame_times waits for
 * the context's stream and writes ms[i], frame i's completion time
 * measured from the launch's start. */
typedef struct clim_burst clim_burst;
#define CLIM_BURST_FRAME_TIMES 1u
/* record only one form instead of timing both: the cross-frame launches
 * (ignored when no frame can take them) or the per-frame kernels */
#define CLIM_BURST_CROSS_FRAME 2u
#define CLIM_BURST_PER_FRAME 4u

typedef struct {
    int32_t waves[2];          /* register budget of the chosen build, as waves per SIMD: [0] green, [1] red/blue */
    int32_t blocks_per_cu[2];  /* persistent blocks per CU */
    int32_t grab[2];           /* tiles a block takes from its queue at a time */
    float us[2];               /* the chosen shape's time per launch (us), at creation */
    float us_first[2];         /* the first candidate's (one block per CU, one tile per take) */
    int32_t candidates[2];     /* shapes timed */
    int32_t tiles[2];          /* tiles per launch (all frames) */
    int32_t cross_frames;      /* frames the cross-frame launches can take (the high-noise ones demosaic on their own) */
    int32_t cached;            /* 1: the choices came from an earlier burst of the same frame sizes */
    int32_t form;              /* the recorded graph kept: 0 with the cross-frame launches, 1 with per-frame kernels */
    float form_ms[2];          /* a replay of each form at creation (ms; 0 when not timed) */
} clim_burst_tuning;

clim_status clim_burst_create(clim_ctx* ctx, int32_t n, const clim_image* const* raws,
                              clim_image* const* out_rgba8, clim_image* const* out_f32,
                              const clim_demosaic_params* params, uint32_t flags, clim_burst** burst);
clim_status clim_burst_launch(clim_burst* burst);
clim_status clim_burst_frame_times(clim_burst* burst, float* ms);
clim_status clim_burst_tuning_read(const clim_burst* burst, clim_burst_tuning* tuning);
void clim_burst_destroy(clim_burst* burst);

/* ------------------------------------------------- synthetic scenes
 * Seeded synthetic RAW frames (smooth background, 200 Gaussian blobs, 50
 * hard-edged rectangles; Poisson-Gaussian noise y = a*Poisson(x/a) + N(0,b);
 * DN = clamp(rint(black + y*(white-black)), 0, white)), generated on the
 * device with Philox4x32-10 keyed by the seeds.  clean (nullable, RGBA32F)
 * receives the noise-free linear scene.  Not a reference function: the
 * reference reads DNG files (src/gls_image_tiff.cpp:255); this is the input
 * generator for tests, benchmarks and learned-stage training data. */
clim_status clim_generate_bayer(clim_ctx* ctx, clim_image* raw, clim_image* clean, int32_t bayer,
                                uint32_t scene_seed, uint32_t noise_seed, float noise_a,
                                float noise_b, float black_level, float white_level);

/* --------------------------------------------- learned denoise strength
 * No reference counterpart: a method change at the per-level filter plug
 * point (ImageDenoiser::denoise, cpp/pyramidal_denoise.hpp:24-34).  A
 * 4-layer 3x3 convolutional network (8 -> 32 -> 32 -> 32 -> 20 channels,
 * bfloat16 weights, FP32 accumulation on the matrix cores) runs once per
 * frame on the grid of pyramid level 3, after the per-level noise fits, and
 * predicts for every texel and level a multiplier of the bilateral range
 * sigma (cl/demosaic.cl:735-765) or of the guided filter's epsilon
 * (cl/demosaic.cl:767-851), and the level's error relative to the plain
 * strength.  A level uses the multipliers only when its gate is open; shut,
 * it runs the plain filter and its output is the plain pyramid's bit for
 * bit.  It applies to the Bilateral and GuidedFast denoisers of pyramids with
 * at least 4 levels (otherwise every level is plain).
 *
 * Parameters: CLIM_STRENGTH_PARAMS floats, per layer l the weights
 * W[tap = 3 ky + kx][cin][cout] then the biases b[cout]; weights are rounded
 * to bfloat16 when set. */
#define CLIM_STRENGTH_PARAMS 26612
typedef enum {
    CLIM_STRENGTH_OFF = 0,            /* the plain pyramid (default) */
    CLIM_STRENGTH_GATED = 1,          /* a level opens when its mean predicted log2 error ratio is < 0 */
    CLIM_STRENGTH_FORCE_PLAIN = 2,    /* the network runs, every gate shut: the plain output */
    CLIM_STRENGTH_FORCE_LEARNED = 3   /* every gate open */
} clim_strength_mode;
clim_status clim_strength_set_weights(clim_ctx* ctx, const float* params, int64_t n);
clim_status clim_strength_get_weights(clim_ctx* ctx, float* params, int64_t n);
/* applies to later clim_raw_to_rgb / clim_raw_to_rgb_batch /
 * clim_pyramidal_denoise calls on the context; needs weights unless OFF */
clim_status clim_strength_set_mode(clim_ctx* ctx, int32_t mode);
/* Blocking.  The last frame's gate: open[l], and (nullable) the mean
 * predicted log2 error ratio of level l over the grid. */
clim_status clim_strength_gate_read(clim_ctx* ctx, int32_t open[5], float mean_pred[5]);
/* Blocking.  The last frame's multiplier map: 5 planes of gw x gh floats
 * (the grid of level 3), plane l level l's multipliers.  With map null only
 * gw and gh are written. */
clim_status clim_strength_map_read(clim_ctx* ctx, float* map, int64_t capacity, int32_t* gw, int32_t* gh);
/* Blocking.  The training loss and its gradient on one batch (host arrays):
 * feat n x h x w x 8 network inputs (rounded to bfloat16), curves
 * n x h x w x 70 error curves ([level][filter][k] for log2 strength (k-3)/2),
 * norm n x 10 normalisers ([level][filter], > 0 to count).  out_raw
 * (nullable): n x h x w x 32 raw outputs (channels >= 20 zero); grad
 * (nullable): CLIM_STRENGTH_PARAMS; loss: the strength and prediction terms. */
clim_status clim_strength_loss_grad(clim_ctx* ctx, const float* params, const float* feat, const float* curves,
                                    const float* norm, int32_t n, int32_t h, int32_t w, float* out_raw, float* grad,
                                    double loss[2]);
typedef struct {
    int32_t frames;         /* synthetic frames (seeds seed + i, noise seeds seed + 100000 + i) */
    int32_t width, height;  /* frame size (multiples of 8, at least 256) */
    int32_t crop;           /* crop size on the level-3 grid (32: 256 x 256 frame pixels) */
    int32_t epochs, batch;  /* passes over the crops, crops per step */
    float lr;               /* Adam step (a tenth over the last fifth of the epochs) */
    uint32_t seed;
    int32_t bayer;
    int32_t noise_levels;   /* frame i has Poisson gain a_(i mod noise_levels) of 0.0005, 0.001, 0.002,
                               0.004, 0.008, 0.016, 0.003, 0.006; read noise a / 200 */
    int32_t verbose;        /* progress lines on stderr (frames generated, epochs) */
} clim_strength_train_config;
/* Blocking.  Generates the crops, trains from seeded initial weights with
 * the library's kernels, and sets the result as the context's weights.
 * log (nullable, 2 x epochs floats): each epoch's sampled losses. */
clim_status clim_strength_train(clim_ctx* ctx, const clim_strength_train_config* cfg, float* log, int64_t log_n);

/* ------------------------------------------- allocation-free variant
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  clim_compute_raw_noise_statistics with every work buffer
 * (sample lists, block partials, fit state) in one static device array of
 * the library, CLIM_NOALLOC_SCRATCH_BYTES on each device it runs on, held
 * until the process ends; the call allocates nothing and does not block.
 * nlf_dev (device memory, 8 floats) receives A[4], B[4]; status_dev
 * (nullable, device memory) receives 0 (fitted), 1 (no sample in range:
 * nlf_dev receives zeros) or 2 (no inliers in the refit: the first-pass
 * model).  Mosaics whose quad grid, each axis rounded up to 16, exceeds
 * CLIM_NOALLOC_MAX_QUADS (e.g. larger than 4096 x 3072 pixels) give
 * CLIM_EINVAL.  Every call on one device shares the array: calls that
 * overlap on one GPU (two streams or two contexts) race. */
#define CLIM_NOALLOC_MAX_QUADS 3145728
#define CLIM_NOALLOC_SCRATCH_BYTES (CLIM_NOALLOC_MAX_QUADS * 32 + 1048576)
clim_status clim_compute_raw_noise_statistics_noalloc(clim_ctx* ctx, const clim_image* scaled, int32_t bayer,
                                                      float* nlf_dev, int32_t* status_dev);

/* ------------------------------------------------------------ profiling
 * Per-stage device timing with HIP events on the context's stream (the
 * counterpart of the reference's PRINT_EXECUTION_TIME wall-clock block,
 * cpp/raw_converter.cpp:22,97-99,176-183).  While enabled, every kernel
 * launch of clim_raw_to_rgb is bracketed by two events; clim_profile_read
 * waits for the stream, adds the elapsed times per stage and clears the
 * record.  Stage names: clim_profile_stage_name(i), i < CLIM_PROFILE_STAGES. */
#define CLIM_PROFILE_STAGES 14
typedef struct {
    double ms[CLIM_PROFILE_STAGES];         /* summed kernel time per stage */
    int64_t launches[CLIM_PROFILE_STAGES];  /* launches per stage */
    int64_t texels[CLIM_PROFILE_STAGES];    /* output texels (or fitted samples) per stage */
} clim_profile;
clim_status clim_profile_enable(clim_ctx* ctx, int32_t enable);
/* Restricts event recording to the stages whose bit (1 << stage) is set in
 * stage_mask (default: all stages), so that a timed run can measure one
 * kernel while recording as few events as possible. */
clim_status clim_profile_select(clim_ctx* ctx, uint32_t stage_mask);
clim_status clim_profile_read(clim_ctx* ctx, clim_profile* out);
const char* clim_profile_stage_name(int32_t stage);

#ifdef __cplusplus
}
#endif

#endif /* CLIMAGE_H */
