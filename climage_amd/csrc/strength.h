// strength.h - the learned-strength network (internal to the library).
//
// A 4-layer 3x3 convolutional network (8 -> 32 -> 32 -> 32 -> 20 channels,
// ReLU between layers, zero padding) evaluated on the grid of pyramid level
// kSnGrid = 3 (an eighth of the frame's width and height).  Inputs: 5x5
// statistics of that level against the fitted noise models (k_sn_features).  Outputs per
// level l (0..4): channel 4l + 0 the log2 bilateral range multiplier, 4l + 1
// the log2 guided-epsilon multiplier (applied squared), 4l + 2 / 4l + 3 the
// predicted log2 ratio of the level's error with the learned strength to its
// error with the plain strength, for the bilateral and the guided filter.
//
// Storage: activations NHWC bfloat16 ([N][H][W][C], C = 8 or 32); weights
// in bfloat16 for the MFMA operands (FP32 master copy, FP32 biases); every
// product is accumulated in FP32 by v_mfma_f32_32x32x16_bf16.
#pragma once

#include <cstdint>

#include "device_common.h"

namespace clim {

constexpr int kSnLayers = 4;
constexpr int kSnFeat = 8;        // input channels
constexpr int kSnHidden = 32;     // hidden channels
constexpr int kSnOut = 20;        // output channels (4 per level)
constexpr int kSnLevels = 5;
constexpr int kSnGrid = 3;        // the pyramid level whose texels the network runs on
constexpr int kSnCurveK = 7;      // strengths of a training error curve: log2 m = (k - 3) / 2
constexpr int kSnCurves = kSnLevels * 2 * kSnCurveK;  // [level][filter][k] per grid texel
constexpr float kSnSMax = 1.5f;   // |log2 m| bound (the curves' span)

__host__ __device__ constexpr int sn_cin(int l) { return l == 0 ? kSnFeat : kSnHidden; }
__host__ __device__ constexpr int sn_cout(int l) { return l == kSnLayers - 1 ? kSnOut : kSnHidden; }
// K of a layer's forward GEMM, padded to whole 16-deep MFMA steps
__host__ __device__ constexpr int sn_kpad(int l) { return (9 * sn_cin(l) + 15) / 16 * 16; }
// canonical parameter layout: per layer W[tap][cin][cout] (tap = 3 ky + kx) then b[cout]
__host__ __device__ constexpr int sn_layer_params(int l) { return 9 * sn_cin(l) * sn_cout(l) + sn_cout(l); }
__host__ __device__ constexpr int sn_layer_offset(int l) {
    return l == 0 ? 0 : sn_layer_offset(l - 1) + sn_layer_params(l - 1);
}
constexpr int kSnParams = sn_layer_offset(kSnLayers);

// device weights in operand layouts (rebuilt from the master by launch_sn_prep)
struct SnWeights {
    float* master;          // [kSnParams] FP32, canonical
    uint16_t* fwd[kSnLayers];   // [32][sn_kpad(l)] bf16, row = output channel, k = tap * cin + ci
    uint16_t* dgrad[kSnLayers]; // l >= 1: [32][288] bf16, row = input channel, k = (8 - tap) * 32 + co
    float* bias[kSnLayers];     // [32] FP32 (zero past cout)
};

// Conv epilogues
enum SnEpilogue {
    SN_RELU = 0,   // bf16 out = relu(acc + bias)
    SN_MASK = 1,   // bf16 out = acc where mask > 0 (backward through the previous ReLU)
    SN_RAW = 2     // fp32 out [P][32] = acc + bias (training forward of the last layer)
};

struct SnConv {
    const uint16_t* x;      // [N][H][W][cin]
    const uint16_t* wa;     // [32][kpad]
    const float* bias;      // [32] or null
    int n, h, w;
    uint16_t* y;            // SN_RELU / SN_MASK: [N][H][W][32]
    const uint16_t* mask;   // SN_MASK: [N][H][W][32]
    float* y32;             // SN_RAW: [N][H][W][32]
};
// cin = 8 or 32
void launch_sn_conv(const SnConv& a, int cin, SnEpilogue epi, hipStream_t st);

// inference of one frame in one launch: features of the grid level (lvg) and all
// four layers with the activations in LDS, writing the map [5][gh][gw]
// (plane l: level l's multipliers) and adding the predicted log2 error
// ratios' sums (2^-16 units) to gate_sum[5]; filter: bit l set when the
// guided filter denoises level l (its epsilon multiplier, 4^s, is mapped),
// else the bilateral's range multiplier 2^s
void launch_sn_infer(const Img& lvg, const float* nlf0, const float* nlfg, const SnWeights& wt, float* map,
                     int filter, long long* gate_sum, hipStream_t st);

// weight gradient of one layer: grad (canonical layout of the layer, FP32)
// = sum over pixels of x(shifted) * dz, bias gradient = sum dz; partials:
// kSnWgradBlocks x (9 cin 32 + 32) floats of scratch
constexpr int kSnWgradBlocks = 256;
void launch_sn_wgrad(const uint16_t* x, const uint16_t* dz, int n, int h, int w, int cin, int cout, float* partials,
                     float* grad, hipStream_t st);

// master -> operand layouts
void launch_sn_prep(const SnWeights& wt, hipStream_t st);
// Adam step on the master weights (t: step number from 1)
void launch_sn_adam(float* master, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                    float eps, int t, hipStream_t st);

// Training loss on raw outputs out32 [P][32] against error curves [P][kSnCurves]
// and per-image normalisers norm [N][5][2]: writes dz [P][32] (bf16) and per-block
// loss partials (two floats per block: strength loss, prediction loss);
// returns the block count
int launch_sn_loss(const float* out32, const float* curves, const float* norm, int n, int hw, float scale,
                   uint16_t* dz, float* loss_partials, hipStream_t st);

// features of the grid level's image lvg (RGBA32F) with the calibrated noise
// models of level 0 and of the grid level (device: A_Y, A_Cb, A_Cr, B_Y, B_Cb, B_Cr)
void launch_sn_features(const Img& lvg, const float* nlf0, const float* nlfg, uint16_t* feat, hipStream_t st);

// error curve entry (level, filter, k) on the grid gw x gh: mean over the
// level's texels in each grid texel (level <= kSnGrid) or the texel under it
// (coarser levels) of |y2c (den - clean)|^2
void launch_sn_curve(const Img& den, const Img& clean, const float y2c[9], int level, int gw, int gh, float* curves,
                     int slot, hipStream_t st);

// crops: copy crop (cx, cy) of size cs x cs from the frame grid (gw wide) of
// features / curves into training-set slot `dst`
void launch_sn_gather(const uint16_t* feat, const float* curves, int gw, int cs, const int2* origins, int ncrops,
                      uint16_t* feat_out, float* curves_out, float* norm_out, hipStream_t st);
// batch assembly from the training set by crop index
void launch_sn_batch(const uint16_t* feat_set, const float* curves_set, const float* norm_set, const int* idx,
                     int nb, int cs, uint16_t* feat_out, float* curves_out, float* norm_out, hipStream_t st);

// the gate: open[l] = mode 3 ? 1 : mode 2 ? 0 : (sum[l] < 0) (mode 1), l < levels
void launch_sn_gate(const long long* sum, int levels, int mode, int* open, hipStream_t st);

}  // namespace clim
