// kernels.h - launchers of the gfx950 kernels (internal to the library).
#pragma once

#include "device_common.h"

namespace clim {

// ----------------------------------------------------- noise-model regression
constexpr int kFitSums = 21;  // sums of the largest pass (raw first pass: 5 x 4 channels + count)

struct FitState {
    clim_fx128 tot[kFitSums];  // totals of the current pass
    double A[4], B[4], err2[4];
    unsigned long long n0;
    int status;            // 0 fitted, 1 no sample in range (fallback), 2 no inliers (first fit)
    float out[8];          // effective model: A[0..nch), B[0..nch)
    float cal[8];          // out * mult (calibrated per-level model)
    unsigned int reduced;  // blocks of the running combination that are done (0 between launches)
};

constexpr int kFitBlocksPerJob = 2048;
// blocks of a statistics launch: as many 256-thread blocks per CU as the
// kernel's registers allow (raw: five waves per SIMD; levels: four, which
// measured faster than five with fewer registers)
constexpr int kRawStatsGrid = 5 * 256, kLevelStatsGrid = 4 * 256;

// A fit's first pass runs inside the statistics kernel, which reads `src`
// (the scaled mosaic with Bayer table `b`, or an RGBA level) and appends the
// samples that pass the range test to sample lists; the refit streams over
// those.  Raw fits list the rawNoiseStatistics mean and var (float4 R, G, B,
// G2) in m and v, level fits the noiseStatistics values (mean Y, var Y,
// var Cb, var Cr) in m.  The lists are split in runs, one per statistics
// block: block bi of a job owns the slots from t0(bi) * seg (t0: its first
// tile, seg: statistics per tile) and writes its sample count to counts[bi].
struct FitJob {
    Img src;
    Bayer b;
    Img m, v;                 // sample lists (float4 slots from m.p, v.p)
    long long cap;            // slots in each list
    int nblocks, block_base;  // set by the launchers
    int ntiles, seg;          // set by the statistics launchers
    unsigned int* counts;     // [kFitBlocksPerJob] samples listed per block
    FitState* st;
    clim_fx128* partials;     // nblocks x accumulator count
    float nlf_in[8];          // fallback model (written to out when no sample is in range)
    float mult[8];            // calibration multipliers
};

struct FitJobs {
    FitJob j[5];
    int njobs;
};

// Statistics of every job plus the first regression pass (block partials,
// sample lists; sets the jobs' block shares and tiling for launch_fit_rest).
// smul: read the jobs' src as the uint16 mosaic, scaled with (smul, black);
// null: src is the scaled FP32 mosaic.  False (nothing launched) when a
// job's lists have fewer than ntiles * seg slots.
bool launch_raw_stats_fit(FitJobs& jobs, const float* smul, float black, hipStream_t st);
// jobs [j0, j1) of `jobs` in one launch (their block shares are set in
// `jobs`); down (one job, downsample_exact2x(src, *down)): the launch also
// writes downsampleImage of the job's level into *down.  False as above.
bool launch_noise_stats_fit(FitJobs& jobs, int j0, int j1, const Img* down, hipStream_t st);
// list slots a statistics launch needs for a w x h level (texels) or a w x h
// mosaic (raw: quads), as allocation sizes: the width and height to allocate
// an RGBA32F image with
void fit_list_dims(bool raw, int w, int h, int* aw, int* ah);
// downsampleImage from `in` to `out` is an exact 2x reduction (every tap at
// an odd texel offset with weight 1/2); launch_noise_stats_fit can then do it
// on the way
bool downsample_exact2x(const Img& in, const Img& out);
// Gives each job a run of blocks in proportion to units[j] (at most one per
// unit and kFitBlocksPerJob, at least one), about `grid` in all; returns the
// total.
int fit_share_blocks(FitJobs& jobs, const long long* units, int grid);
// the passes after the statistics kernel: combine and solve, refit over the
// sample lists (the statistics launches' block shares), combine and solve
void launch_fit_rest(FitJobs& jobs, bool raw, hipStream_t st);

// ------------------------------------------------------------ demosaic stages
void launch_scale_raw(const Img& raw, const Img& out, const Bayer& b, const float smul[4], float black,
                      hipStream_t st);
void launch_raw_stats(const Img& scaled, const Bayer& b, const Img& mean, const Img& var, hipStream_t st);
// the mosaic a demosaic kernel reads: the scaled FP32 mosaic, or the uint16
// mosaic with the scaleRawData parameters (scaled on the fly)
struct MosaicArg {
    Img im;
    Bayer b;
    float4 smul;
    float black;
    bool u16;
    template <bool U16>
    MosaicSrc<U16> src() const {
        return MosaicSrc<U16>{im, b, smul, black};
    }
};
void launch_green(const MosaicArg& raw, const Img& green, int bayer, const ParamRef& lv, hipStream_t st);
void launch_red_blue(const MosaicArg& raw, const Img& green, const Img& out, int bayer, const Bayer& b,
                     const ParamRef& cv, int rotate_180, bool fuse, float clip, const float c2y[9],
                     hipStream_t st);
// row 0 and column 0 of an RGBA32F image to zero (the texels a rotated
// interpolateRedBlue leaves unwritten)
void launch_zero_row0_col0(const Img& out, hipStream_t st);
void launch_blend_highlights(const Img& in, const Img& out, float clip, hipStream_t st);
void launch_transform(const Img& in, const Img& out, const float mat[9], hipStream_t st);
void launch_despeckle(const Img& in, const Img& out, float var_a, float var_b, hipStream_t st);
// interpolateRedBlue + blendHighlights + transformImage(c2y) + despeckleImage
void launch_rb_despeckle(const MosaicArg& raw, const Img& green, const Img& out, int bayer, const Bayer& b,
                         const ParamRef& cv, float clip, const float c2y[9], float var_a, float var_b,
                         hipStream_t st);

// ----------------------------------------------------- burst (cross-frame)
// One frame of a burst in the cross-frame green and red/blue + despeckle
// launches (a device table, frame-major).
struct BurstFrame {
    MosaicSrc<true> raw;   // the uint16 mosaic, scaled on the fly
    Img green, ycc;        // the frame's green plane and despeckled YCbCr image
    int gxo, gyo;          // raw_green site
    ParamRef lv, cv;       // luma / chroma variance for interpolateGreen / RedBlue
    Bayer b;
    int red_is_c1;
    float clip;
    float c2y[9];          // camera -> YCbCr
    float var_a, var_b;    // despeckleImage (pyramidNlf[0])
    int tiles_x[2];        // tiles per row: [0] green, [1] red/blue + despeckle
    int tile_end[2];       // running totals of the frames' tile counts
};
// tile counters of one cross-frame kernel (zero between launches)
struct BurstQueue {
    unsigned int next[8];
    unsigned int done;
    unsigned int pad[7];
};
// tiles of a w x h frame for kind 0 (green) / 1 (red/blue + despeckle)
void burst_tiles(int w, int h, int kind, int* tiles_x, int* tiles);
// every frame's tiles of one stage, `blocks` persistent blocks (a multiple of
// 8), `grab` tiles per take; waves (green: 4, 6 or 8): the register budget
// the kernel is built for, as waves per SIMD (red/blue + despeckle has one
// build: its registers allow 4)
// half: the YCbCr images are FP16 (the mixed-precision form)
void launch_demosaic_burst(int kind, int waves, bool half, const BurstFrame* frames, int total_tiles, BurstQueue* q,
                           int blocks, int grab, hipStream_t st);

// ------------------------------------------------------------- pyramid stages
void launch_downsample(const Img& in, const Img& out, hipStream_t st);
void launch_noise_stats(const Img& in, const Img& out, hipStream_t st);
// Per-texel denoise strength of the level filters (the learned-strength
// network's map, strength.h): the bilateral range threshold is multiplied by
// m^2 (the range sigma by m), the guided epsilon by m.  Off (the plain
// filters, untouched arithmetic) when map is null and constant <= 0.
struct StrengthRef {
    const float* map;   // [5][gh][gw] on the grid of level 3: plane l holds level l's multipliers
    int gw, gh;
    const int* open;    // [5] device: level l reads the map only when open[l] != 0
    float constant;     // > 0: this multiplier at every texel (the map is not read)
};
// grid texels per texel of `level`: 2^(level - 3)
__device__ __forceinline__ float strength_scale(int level) {
    return level == 0 ? 0.125f : level == 1 ? 0.25f : level == 2 ? 0.5f : level == 3 ? 1.0f : 2.0f;
}
// The multiplier at texel (x, y) of `level`, or 0 when the strength is off
// there.  The map is sampled bilinearly (clamped) at the texel's centre,
// grid coordinate (x + 1/2) 2^(level - 3) - 1/2.
__device__ __forceinline__ float strength_at(const StrengthRef& s, int level, int x, int y) {
    if (s.constant > 0.0f) return s.constant;
    if (!s.map || !s.open[level]) return 0.0f;
    const float sc = strength_scale(level);
    const float gx = ((float)x + 0.5f) * sc - 0.5f, gy = ((float)y + 0.5f) * sc - 0.5f;
    const float fx = floorf(gx), fy = floorf(gy);
    const float ax = gx - fx, ay = gy - fy;
    const int ix = (int)fx, iy = (int)fy;
    const int x0 = clampi(ix, 0, s.gw - 1), x1 = clampi(ix + 1, 0, s.gw - 1);
    const int y0 = clampi(iy, 0, s.gh - 1), y1 = clampi(iy + 1, 0, s.gh - 1);
    const float* m = s.map + (size_t)level * s.gw * s.gh;
    const float v00 = m[(size_t)y0 * s.gw + x0], v01 = m[(size_t)y0 * s.gw + x1];
    const float v10 = m[(size_t)y1 * s.gw + x0], v11 = m[(size_t)y1 * s.gw + x1];
    const float top = v00 + (v01 - v00) * ax, bot = v10 + (v11 - v10) * ax;
    return top + (bot - top) * ay;
}

// The map texels under one tile (x0, y0, TX x TY texels) of a level no
// coarser than the grid (level <= 3), read once per block into LDS (mt,
// kMapStageMax floats, clamped like strength_at's reads); on is false when
// the strength is off for the level, constant, or the level is coarser (then
// strength_at reads the map directly).  Every thread of the block calls it
// (256-thread blocks; a __syncthreads follows before strength_at_staged).
constexpr int kMapStageMax = 34 * 18;
struct MapStage {
    int gx0, gy0, w;
    bool on;
};
template <int TX, int TY>
__device__ __forceinline__ MapStage stage_strength(const StrengthRef& s, int level, int x0, int y0, float* mt,
                                                   int tid) {
    MapStage r{0, 0, 0, false};
    if (s.constant > 0.0f || !s.map || level > 3 || !s.open[level]) return r;
    const float sc = strength_scale(level);
    r.gx0 = (int)floorf(((float)x0 + 0.5f) * sc - 0.5f);
    r.gy0 = (int)floorf(((float)y0 + 0.5f) * sc - 0.5f);
    const int gx1 = (int)floorf(((float)(x0 + TX - 1) + 0.5f) * sc - 0.5f) + 1;
    const int gy1 = (int)floorf(((float)(y0 + TY - 1) + 0.5f) * sc - 0.5f) + 1;
    r.w = gx1 - r.gx0 + 1;
    const int n = r.w * (gy1 - r.gy0 + 1);
    for (int i = tid; i < n; i += 256) {
        const int gx = clampi(r.gx0 + i % r.w, 0, s.gw - 1), gy = clampi(r.gy0 + i / r.w, 0, s.gh - 1);
        mt[i] = s.map[((size_t)level * s.gh + gy) * s.gw + gx];
    }
    r.on = true;
    return r;
}
// strength_at with the four texels from the staged tile (the same values and
// arithmetic); falls back to strength_at when the tile is not staged
__device__ __forceinline__ float strength_at_staged(const StrengthRef& s, const MapStage& m, const float* mt,
                                                    int level, int x, int y) {
    if (!m.on) return strength_at(s, level, x, y);
    const float sc = strength_scale(level);
    const float gx = ((float)x + 0.5f) * sc - 0.5f, gy = ((float)y + 0.5f) * sc - 0.5f;
    const float fx = floorf(gx), fy = floorf(gy);
    const float ax = gx - fx, ay = gy - fy;
    const int i0 = (int)fx - m.gx0, j0 = (int)fy - m.gy0;
    const float v00 = mt[j0 * m.w + i0], v01 = mt[j0 * m.w + i0 + 1];
    const float v10 = mt[(j0 + 1) * m.w + i0], v11 = mt[(j0 + 1) * m.w + i0 + 1];
    const float top = v00 + (v01 - v00) * ax, bot = v10 + (v11 - v10) * ax;
    return top + (bot - top) * ay;
}

// level: the texels' pyramid level (for the strength map)
void launch_bilateral(const Img& in, const Img& out, const ParamRef& nlf, float chromaBoost, float gradientBoost,
                      hipStream_t st, const StrengthRef* sr = nullptr, int level = 0);
// denoiseImage of n pyramid levels in one launch (base: set by the launcher);
// entry l is pyramid level level[l] (the finest first)
struct BilateralLevels {
    Img in[5], out[5];
    ParamRef nlf[5];
    float chromaBoost[5], gradientBoost[5];
    int base[5], level[5];
    int n;
    StrengthRef sr;
};
void launch_bilateral_levels(BilateralLevels& lv, hipStream_t st);
// denoiseImage of one FP32 pyramid level (in) and reassembleImage of the
// result with the next level (n1 noisy, d1 reassembled) into out, in one
// launch: the denoised level stays on chip.  Only for an exact 2x next level
// (bilateral_reassemble_fits).
bool bilateral_reassemble_fits(const Img& in, const Img& n1, const Img& d1, const Img& out);
void launch_bilateral_reassemble(const Img& in, const Img& n1, const Img& d1, const Img& out, const ParamRef& nlf,
                                 float chromaBoost, float gradientBoost, float sharpening, const StrengthRef& sr,
                                 int level, hipStream_t st);
void launch_guided_fast(const Img& in, const Img& out, const ParamRef& nlf, hipStream_t st,
                        const StrengthRef* sr = nullptr, int level = 0);
// convertTosRGB applied to reassembleImage's result in the same kernel
struct ToneLaunch {
    Img out8, outf;  // outf.p may be null
    const float* pre;  // YCbCr -> camera (with exposure), or null
    const float* cam;
    float contrast, saturation, slope, exposureBias, blacks;
    // localToneMapping: the guide (pyramid level 3; p null = off) and
    // guidedFilterEps, shadows, highlights, detail; the mask's YCbCr -> camera
    // transform is `pre`
    Img ltm_guide;
    float ltm[4];
};
void launch_reassemble(const Img& d0, const Img& n1, const Img& d1, const Img& out, float sharpening,
                       const ParamRef& nlf, hipStream_t st, const ToneLaunch* tone = nullptr);
// mask.p null: no local tone mapping
void launch_tone(const Img& in, const Img& mask, const Img& out8, const Img& outf, const float* pre,
                 const float cam[9], float contrast, float saturation, float slope, float exposureBias, float blacks,
                 hipStream_t st);
// localToneMappingMask: in the denoised YCbCr image, out R32F of its size
void launch_ltm_mask(const Img& in, const Img& guide, const float ltm[4], const float T[9], const Img& out,
                     hipStream_t st);

// -------------------------------------------- high-noise raw branch (RGBA raw)
// One (R, G, B, G2) quad per texel, w/2 x h/2 (k_rawrgba.hip).
void launch_bayer_to_raw_rgba(const Img& raw, const Img& rgba, const Bayer& b, hipStream_t st);
void launch_raw_rgba_to_bayer(const Img& rgba, const Img& raw, const Bayer& b, hipStream_t st);
void launch_despeckle_raw_rgba(const Img& in, const Img& out, hipStream_t st);      // out != in
void launch_denoise_raw_rgba(const Img& in, const ParamRef& var, const Img& out, hipStream_t st);  // out != in
// the four in one pass: mosaic -> despeckle -> denoise -> FP32 mosaic (out != m.im)
void launch_raw_predenoise(const MosaicArg& m, const ParamRef& var, const Img& out, hipStream_t st);

// -------------------------------------------------- half-resolution preview
// fasteDebayer (FP32 mosaic -> half-size RGBA) and the whole
// fastDemosaicImage (mosaic -> half-size RGBA8, k_preview.hip)
void launch_fast_debayer(const Img& raw, const Img& rgba, const Bayer& b, hipStream_t st);
void launch_fast_preview(const MosaicArg& m, const Img& out8, const Img& outf, const float cam[9], float contrast,
                         float saturation, float slope, float exposureBias, float blacks, hipStream_t st);

// ----------------------------------------------- multi-pass guided filter
// scratch: 7 images of w*h texels (float4) + 1 of float2, see k_guided.hip
struct GuidedScratch {
    Img mean_I, s1, s2, ipr, ipg, ipb, tmp4, tmp2;
};
void launch_guided_filter(const Img& in, const Img& out, int fs, const ParamRef& eps, const GuidedScratch& g,
                          hipStream_t st);
// the same filter in one launch with no scratch, for the filter sizes where
// guided_filter_fused(fs) holds (3 and 5; the pyramid uses 5)
bool guided_filter_fused(int fs);
void launch_guided_filter_fused(const Img& in, const Img& out, int fs, const ParamRef& eps, hipStream_t st);

// -------------------------------------------------------- synthetic scenes
constexpr int kSceneBlobs = 200, kSceneRects = 50;
struct SceneParams {
    float base[3], gx[3], gy[3], gxy[3], amp[3], fx[3], fy[3], ph[3];
    float bcx[kSceneBlobs], bcy[kSceneBlobs], bw[kSceneBlobs], binv[kSceneBlobs], ba[kSceneBlobs][3];
    float rx0[kSceneRects], ry0[kSceneRects], rx1[kSceneRects], ry1[kSceneRects], rc[kSceneRects][3];
};
void scene_draw_params(int w, int h, uint32_t seed, SceneParams* p);
void launch_scene(const Img& raw, const Img& clean, const SceneParams* dparams, int bayer_site_color_packed,
                  uint32_t noise_seed, float a, float b, float black, float white, hipStream_t st);

}  // namespace clim
