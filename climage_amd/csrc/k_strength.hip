// k_strength.hip - kernels of the learned-strength network (strength.h).
//
// The convolution is an implicit GEMM on v_mfma_f32_32x32x16_bf16: one
// block computes 32 x 8 output pixels of one image; its input tile (34 x 10
// pixels with the zero-padded halo) is staged in LDS, the weights (the A
// operand, 32 output channels x K) sit in registers for the whole block, and
// each wave multiplies them by two 32-pixel rows of the tile (the B operand,
// K x 32 pixels).  K runs over (tap, input channel): one 16-deep step is half
// a tap's 32 channels (cin 32) or two taps' 8 channels (cin 8).  The result
// of a step, D[channel][pixel], leaves each lane one pixel and 16 channels in
// four runs of four, stored as 8-byte (bf16) or 16-byte (FP32) chunks.
//
// The backward pass uses the same kernel: the input gradient of a layer is
// the convolution of the output gradient with the tap-flipped, transposed
// weights (exact with zero padding), masked by the previous layer's ReLU.
// The weight gradient (k_sn_wgrad) is a GEMM over pixels: D[co][(tap, ci)]
// = sum_p dz[p][co] x[p + tap][ci], accumulated per block over a run of
// tiles, written as per-block partials and summed in a fixed order.
#include <algorithm>

#include "strength.h"

namespace clim {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite values)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

constexpr int kCTX = 32, kCTY = 8;  // output tile of a convolution block

// ------------------------------------------------------------- convolution
template <int CIN, int EPI>
__global__ void __launch_bounds__(256) k_sn_conv(SnConv a) {
    constexpr int LX = kCTX + 2, LY = kCTY + 2;
    constexpr int PS = CIN == 32 ? 40 : 8;  // LDS pixel stride (bf16): 80 B keeps a wave's 16-B reads on distinct banks
    constexpr int CH = CIN / 8;             // 16-byte chunks per pixel
    constexpr int KSTEPS = CIN == 32 ? 18 : 5;
    constexpr int KPAD = KSTEPS * 16;
    __shared__ __align__(16) uint16_t tile[LY * LX * PS];

    const int tiles_x = (a.w + kCTX - 1) / kCTX, tiles_y = (a.h + kCTY - 1) / kCTY;
    const int per_img = tiles_x * tiles_y;
    const int img = blockIdx.x / per_img, r = blockIdx.x % per_img;
    const int x0 = (r % tiles_x) * kCTX, y0 = (r / tiles_x) * kCTY;
    const int tid = threadIdx.x;
    const uint16_t* xin = a.x + (size_t)img * a.h * a.w * CIN;
    for (int i = tid; i < LY * LX * CH; i += 256) {
        const int p = i / CH, q = i % CH;
        const int gx = x0 - 1 + p % LX, gy = y0 - 1 + p / LX;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gx >= 0 && gx < a.w && gy >= 0 && gy < a.h)
            v = *reinterpret_cast<const uint4*>(xin + ((size_t)gy * a.w + gx) * CIN + q * 8);
        *reinterpret_cast<uint4*>(&tile[p * PS + q * 8]) = v;
    }
    const int wave = tid >> 6, lane = tid & 63, half = lane >> 5, col = lane & 31;
    // A operand: lane (col, half) holds weights of output channel col, k = 16 s + 8 half + [0, 8)
    bf16x8 af[KSTEPS];
    const uint16_t* wrow = a.wa + (size_t)col * KPAD + 8 * half;
#pragma unroll
    for (int s = 0; s < KSTEPS; s++) af[s] = as_bf16x8(*reinterpret_cast<const uint4*>(wrow + 16 * s));
    __syncthreads();

    f32x16 acc[2];
#pragma unroll
    for (int g = 0; g < 2; g++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[g][i] = 0.0f;
#pragma unroll
    for (int s = 0; s < KSTEPS; s++) {
        const int tap = CIN == 32 ? (s >> 1) : 2 * s + half;
        const int ci0 = CIN == 32 ? 16 * (s & 1) + 8 * half : 0;
        const int dy = tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int g = 0; g < 2; g++) {
            const int py = 2 * wave + g;  // tile row of this 32-pixel group
            uint4 bv = make_uint4(0u, 0u, 0u, 0u);
            if (CIN == 32 || tap < 9)  // cin 8: the tenth "tap" of the last step is K padding
                bv = *reinterpret_cast<const uint4*>(&tile[((py + 1 + dy) * LX + (col + 1 + dx)) * PS + ci0]);
            acc[g] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[s], as_bf16x8(bv), acc[g], 0, 0, 0);
        }
    }

    // D[channel][pixel]: lane (col, half) holds pixel col, channels 8 j + 4 half + i
#pragma unroll
    for (int g = 0; g < 2; g++) {
        const int px = x0 + col, py = y0 + 2 * wave + g;
        const bool ok = px < a.w && py < a.h;
        const size_t pix = ((size_t)img * a.h + py) * a.w + px;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int c0 = 8 * j + 4 * half;
            float v[4];
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = acc[g][4 * j + i] + (a.bias ? a.bias[c0 + i] : 0.0f);
            if (EPI == SN_RELU || EPI == SN_MASK) {
                if (EPI == SN_MASK && ok) {
                    const uint2 mk = *reinterpret_cast<const uint2*>(a.mask + pix * 32 + c0);
                    const uint32_t mw[4] = {mk.x & 0xFFFFu, mk.x >> 16, mk.y & 0xFFFFu, mk.y >> 16};
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const bool pos = (mw[i] & 0x8000u) == 0u && (mw[i] & 0x7FFFu) != 0u;
                        v[i] = pos ? v[i] : 0.0f;
                    }
                }
                if (EPI == SN_RELU)
#pragma unroll
                    for (int i = 0; i < 4; i++) v[i] = fmaxf(v[i], 0.0f);
                if (ok) {
                    uint2 o;
                    o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                    o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                    *reinterpret_cast<uint2*>(a.y + pix * 32 + c0) = o;
                }
            } else {  // SN_RAW
                if (ok) *reinterpret_cast<float4*>(a.y32 + pix * 32 + c0) = make_float4(v[0], v[1], v[2], v[3]);
            }
        }
    }
}

void launch_sn_conv(const SnConv& a, int cin, SnEpilogue epi, hipStream_t st) {
    const int blocks = a.n * ((a.w + kCTX - 1) / kCTX) * ((a.h + kCTY - 1) / kCTY);
    const dim3 grd(blocks), blk(256);
#define SN_CONV_CASE(C, E) \
    if (cin == C && epi == E) { hipLaunchKernelGGL((k_sn_conv<C, E>), grd, blk, 0, st, a); return; }
    SN_CONV_CASE(8, SN_RELU)
    SN_CONV_CASE(32, SN_RELU)
    SN_CONV_CASE(32, SN_MASK)
    SN_CONV_CASE(32, SN_RAW)
#undef SN_CONV_CASE
}

// ---------------------------------------------------- fused inference
// The features and all four layers of a 16 x 16 tile of the grid in one
// block, the activations in LDS: layer l computes the region of its output
// that the remaining layers need (22, 20, 18, 16 texels square), so the
// activations never reach HBM (each round trip is 49 MB per 12 MP frame).
// Region texels outside the image are zero, which is the zero padding every
// layer's input has in the layer-by-layer form; each output texel sees the
// same bfloat16 inputs and the same MFMA sequence, so the results are that
// form's.  The features are computed from a 28 x 28 tile of the grid level staged in
// LDS (clamped reads, the window sums in k_sn_features's order).
// Two blocks share a CU (80 KB of LDS each); a block reads each layer's
// weights (its A operand) from global memory into registers, the first
// layer's while the features are computed.  The B operand is read from LDS
// a chunk of MFMA steps ahead.
namespace {
constexpr int kFT = 16;              // output tile
constexpr int kR0 = kFT + 8;         // feature region (halo 4)
constexpr int kL2 = kR0 + 4;         // grid-level texels under the feature windows
constexpr int kPS = 40;              // activation pixel stride (bf16)
struct InferSmem {
    uint16_t f[kR0 * kR0 * 8];                      // features
    uint16_t a[(kFT + 6) * (kFT + 6) * kPS];        // layer 1 and layer 3 outputs
    union {
        float4 lv[kL2 * kL2];                       // grid-level texels (before layer 1)
        uint16_t b[(kFT + 4) * (kFT + 4) * kPS];    // layer 2 output
    };
    float gate[4][2][kSnLevels];
};
}  // namespace

__device__ __forceinline__ float sn_pred_f(const float* nlf, int c, float y) {
    return fmaxf(nlf[c] + nlf[3 + c] * y, 1e-10f);
}

// the A operand of a layer (weights of output channel col, k = 16 s + 8 half
// + [0, 8)) into registers
template <int KSTEPS>
__device__ __forceinline__ void sn_load_weights(bf16x8 (&af)[18], const uint16_t* w, int half, int col) {
    const uint16_t* wrow = w + col * (KSTEPS * 16) + 8 * half;
#pragma unroll
    for (int s = 0; s < KSTEPS; s++) af[s] = as_bf16x8(*reinterpret_cast<const uint4*>(wrow + 16 * s));
}

// B operand of step s for output texel q of an S x S region over an
// (S + 2)-wide input region (pixel stride PSI bf16)
template <int CIN, int S, int PSI>
__device__ __forceinline__ uint4 sn_bfrag(const uint16_t* in, int q, int s, int half) {
    constexpr int SI = S + 2;
    const int tap = CIN == 32 ? (s >> 1) : 2 * s + half;
    const int ci0 = CIN == 32 ? 16 * (s & 1) + 8 * half : 0;
    if (CIN == 8 && tap >= 9) return make_uint4(0u, 0u, 0u, 0u);  // K padding of the last step
    const int qy = q / S, qx = q % S;
    return *reinterpret_cast<const uint4*>(in + ((qy + tap / 3) * SI + (qx + tap % 3)) * PSI + ci0);
}

// Two 32-texel groups (q0, q1; the second only when two) through one layer:
// the B fragments of the next chunk of steps are read from LDS while the
// current chunk's MFMAs run.
template <int CIN, int S>
__device__ __forceinline__ void sn_pair(const uint16_t* in, const bf16x8 (&af)[18], int q0, int q1, bool two,
                                        int half, f32x16& acc0, f32x16& acc1) {
    constexpr int KSTEPS = CIN == 32 ? 18 : 5;
    constexpr int PSI = CIN == 32 ? kPS : 8;
    constexpr int CH = CIN == 32 ? 6 : 5, NCH = KSTEPS / CH;
    uint4 bb[2][2][CH];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        acc0[i] = 0.0f;
        acc1[i] = 0.0f;
    }
#pragma unroll
    for (int u = 0; u < CH; u++) {
        bb[0][0][u] = sn_bfrag<CIN, S, PSI>(in, q0, u, half);
        bb[0][1][u] = sn_bfrag<CIN, S, PSI>(in, q1, u, half);
    }
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        if (c + 1 < NCH) {
#pragma unroll
            for (int u = 0; u < CH; u++) {
                bb[(c + 1) & 1][0][u] = sn_bfrag<CIN, S, PSI>(in, q0, (c + 1) * CH + u, half);
                bb[(c + 1) & 1][1][u] = sn_bfrag<CIN, S, PSI>(in, q1, (c + 1) * CH + u, half);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[c * CH + u], as_bf16x8(bb[c & 1][0][u]), acc0, 0, 0, 0);
            if (two)
                acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[c * CH + u], as_bf16x8(bb[c & 1][1][u]), acc1, 0,
                                                               0, 0);
        }
    }
}

// one hidden layer: out region S x S from in region (S + 2) x (S + 2); the
// output texel (qx, qy) is at image position (ox + qx, oy + qy) and is zero
// outside the image.  A wave takes 32-texel groups wave, wave + 4, ... two
// at a time.
template <int CIN, int S>
__device__ __forceinline__ void sn_layer_lds(const uint16_t* in, uint16_t* out, const bf16x8 (&af)[18],
                                             const float* bias, int ox, int oy, int gw, int gh, int wave, int half,
                                             int col) {
    constexpr int NG = (S * S + 31) / 32;
    float bv[16];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) bv[4 * j + i] = bias[8 * j + 4 * half + i];
    for (int g = wave; g < NG; g += 8) {
        const int g2 = g + 4;
        const int q0 = min(32 * g + col, S * S - 1), q1 = min(32 * g2 + col, S * S - 1);
        f32x16 acc0, acc1;
        sn_pair<CIN, S>(in, af, q0, q1, g2 < NG, half, acc0, acc1);
#pragma unroll
        for (int k = 0; k < 2; k++) {
            const int gg = k ? g2 : g;
            const int q = 32 * gg + col;
            if (gg >= NG || q >= S * S) continue;
            const int qy = q / S, qx = q % S;
            const int x = ox + qx, y = oy + qy;
            const bool inside = x >= 0 && x < gw && y >= 0 && y < gh;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const float a = k ? acc1[4 * j + i] : acc0[4 * j + i];
                    v[i] = inside ? fmaxf(a + bv[4 * j + i], 0.0f) : 0.0f;
                }
                uint2 o;
                o.x = (uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                o.y = (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(out + q * kPS + 8 * j + 4 * half) = o;
            }
        }
    }
}

template <bool H>
__global__ void __launch_bounds__(256, 2) k_sn_infer(Img lv, const float* nlf0, const float* nlfg, SnWeights wt,
                                                     float* map, int filter, long long* gate_sum) {
    __shared__ InferSmem sm;
    const int gw = lv.w, gh = lv.h;
    const int tiles_x = (gw + kFT - 1) / kFT;
    const int x0 = (blockIdx.x % tiles_x) * kFT, y0 = (blockIdx.x / tiles_x) * kFT;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, half = lane >> 5, col = lane & 31;
    bf16x8 wa[18];
    sn_load_weights<5>(wa, wt.fwd[0], half, col);
    // grid-level texels (clamped) under the feature windows: origin (x0 - 6, y0 - 6)
    for (int i = tid; i < kL2 * kL2; i += 256) {
        const int x = clampi(x0 - 6 + i % kL2, 0, gw - 1), y = clampi(y0 - 6 + i / kL2, 0, gh - 1);
        sm.lv[i] = ldt4_<H>(lv, x, y);
    }
    __syncthreads();
    // features of the region (x0 - 4, y0 - 4) .. + 24, zero outside the image
    for (int i = tid; i < kR0 * kR0; i += 256) {
        const int rx = i % kR0, ry = i / kR0;
        const int x = x0 - 4 + rx, y = y0 - 4 + ry;
        uint4 o = make_uint4(0u, 0u, 0u, 0u);
        if (x >= 0 && x < gw && y >= 0 && y < gh) {
            float s[3] = {0, 0, 0}, qq[3] = {0, 0, 0};
            for (int j = 0; j < 5; j++)
                for (int k = 0; k < 5; k++) {
                    const float4 v = sm.lv[(ry + j) * kL2 + rx + k];
                    s[0] += v.x; s[1] += v.y; s[2] += v.z;
                    qq[0] += v.x * v.x; qq[1] += v.y * v.y; qq[2] += v.z * v.z;
                }
            const float mY = s[0] / 25.0f;
            float f[8];
            f[0] = mY;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const float m = s[c] / 25.0f, var = qq[c] / 25.0f - m * m;
                f[1 + c] = log2f(fmaxf(var, 1e-12f) / sn_pred_f(nlfg, c, mY)) * 0.25f;
            }
            f[4] = log2f(sn_pred_f(nlfg, 0, mY)) * 0.125f + 1.0f;
            f[5] = log2f(sn_pred_f(nlfg, 1, mY)) * 0.125f + 1.0f;
            f[6] = log2f(sn_pred_f(nlf0, 0, mY)) * 0.125f + 1.0f;
            f[7] = 1.0f;
            o.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
            o.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
            o.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
            o.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
        }
        *reinterpret_cast<uint4*>(&sm.f[i * 8]) = o;
    }
    __syncthreads();
    sn_layer_lds<8, kFT + 6>(sm.f, sm.a, wa, wt.bias[0], x0 - 3, y0 - 3, gw, gh, wave, half, col);
    sn_load_weights<18>(wa, wt.fwd[1], half, col);
    __syncthreads();
    sn_layer_lds<32, kFT + 4>(sm.a, sm.b, wa, wt.bias[1], x0 - 2, y0 - 2, gw, gh, wave, half, col);
    sn_load_weights<18>(wa, wt.fwd[2], half, col);
    __syncthreads();
    sn_layer_lds<32, kFT + 2>(sm.b, sm.a, wa, wt.bias[2], x0 - 1, y0 - 1, gw, gh, wave, half, col);
    sn_load_weights<18>(wa, wt.fwd[3], half, col);
    __syncthreads();
    // last layer: 16 x 16 texels, two 32-texel groups per wave; the map and
    // the predicted error ratios
    {
        constexpr int S = kFT;
        const int g0 = 2 * wave;
        f32x16 acc[2];
        sn_pair<32, S>(sm.a, wa, 32 * g0 + col, 32 * (g0 + 1) + col, true, half, acc[0], acc[1]);
#pragma unroll
        for (int gg = 0; gg < 2; gg++) {
            const int q = 32 * (g0 + gg) + col, qy = q / S, qx = q % S;
            const int x = x0 + qx, y = y0 + qy;
            const bool ok = x < gw && y < gh;
            // lane (col, half) holds levels 2 j + half (channels 4 l .. 4 l + 3)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int l = 2 * j + half;
                float r = 0.0f;
                if (l < kSnLevels) {
                    const int f = (filter >> l) & 1;
                    const float sv = fminf(fmaxf(acc[gg][4 * j + f] + wt.bias[3][4 * l + f], -kSnSMax), kSnSMax);
                    if (ok) map[((size_t)l * gh + y) * gw + x] = exp2f(f ? 2.0f * sv : sv);
                    r = ok ? acc[gg][4 * j + 2 + f] + wt.bias[3][4 * l + 2 + f] : 0.0f;
                }
                // sum over the 32 texels of this half (fixed butterfly order)
#pragma unroll
                for (int off = 16; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
                if (col == 0 && l < kSnLevels) sm.gate[wave][gg][l] = r;
            }
        }
    }
    __syncthreads();
    if (tid < kSnLevels) {
        float s = 0.0f;
        for (int w = 0; w < 4; w++)
            for (int gg = 0; gg < 2; gg++) s += sm.gate[w][gg][tid];
        atomicAdd(reinterpret_cast<unsigned long long*>(gate_sum + tid),
                  (unsigned long long)llrint((double)s * 65536.0));
    }
}

void launch_sn_infer(const Img& lvg, const float* nlf0, const float* nlfg, const SnWeights& wt, float* map,
                     int filter, long long* gate_sum, hipStream_t st) {
    const int tiles = ((lvg.w + kFT - 1) / kFT) * ((lvg.h + kFT - 1) / kFT);
    with_storage(lvg, [&](auto H) {
        hipLaunchKernelGGL((k_sn_infer<H()>), dim3(tiles), dim3(256), 0, st, lvg, nlf0, nlfg, wt, map, filter,
                           gate_sum);
    });
}

// ---------------------------------------------------------- weight gradient
template <int CIN>
__global__ void __launch_bounds__(256) k_sn_wgrad(const uint16_t* x, const uint16_t* dz, int n, int h, int w,
                                                  float* partials) {
    constexpr int LX = kCTX + 2, LY = kCTY + 2;
    constexpr int PS = CIN == 32 ? 33 : 9;          // odd pixel strides spread the 2-byte gathers over banks
    constexpr int NT = CIN == 32 ? 9 : 3;           // 32-column D tiles: one per tap (cin 32), four taps each (cin 8)
    constexpr int PW = (NT + 3) / 4;                // D tiles per wave
    constexpr int STRIDE = 9 * CIN * 32 + 32;       // floats per block partial
    __shared__ uint16_t xt[LY * LX * PS];
    __shared__ uint16_t dt[256 * 33];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, half = lane >> 5, col = lane & 31;
    const int tiles_x = (w + kCTX - 1) / kCTX, tiles_y = (h + kCTY - 1) / kCTY, per_img = tiles_x * tiles_y;
    const int total = n * per_img;
    f32x16 acc[PW];
#pragma unroll
    for (int k = 0; k < PW; k++)
#pragma unroll
        for (int i = 0; i < 16; i++) acc[k][i] = 0.0f;
    float bsum = 0.0f;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int img = t / per_img, r = t % per_img;
        const int x0 = (r % tiles_x) * kCTX, y0 = (r / tiles_x) * kCTY;
        __syncthreads();  // the previous tile is consumed
        const uint16_t* xi = x + (size_t)img * h * w * CIN;
        for (int i = tid; i < LY * LX * CIN; i += 256) {
            const int p = i / CIN, c = i % CIN;
            const int gx = x0 - 1 + p % LX, gy = y0 - 1 + p / LX;
            xt[p * PS + c] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? xi[((size_t)gy * w + gx) * CIN + c] : 0;
        }
        const uint16_t* di = dz + (size_t)img * h * w * 32;
        for (int i = tid; i < 256 * 32; i += 256) {
            const int p = i >> 5, c = i & 31;
            const int gx = x0 + (p & 31), gy = y0 + (p >> 5);
            dt[p * 33 + c] = (gx < w && gy < h) ? di[((size_t)gy * w + gx) * 32 + c] : 0;
        }
        __syncthreads();
#pragma unroll 2
        for (int s = 0; s < 16; s++) {
            // K step s: pixels 16 s + 8 half + j of the tile
            uint4 araw;
            uint16_t* ah = reinterpret_cast<uint16_t*>(&araw);
#pragma unroll
            for (int j = 0; j < 8; j++) ah[j] = dt[(16 * s + 8 * half + j) * 33 + col];
            const bf16x8 av = as_bf16x8(araw);
#pragma unroll
            for (int k = 0; k < PW; k++) {
                const int d = wave + 4 * k;
                if (d >= NT) continue;
                const int tap = CIN == 32 ? d : 4 * d + (col >> 3);
                const int ci = CIN == 32 ? col : (col & 7);
                const int dy = tap / 3 - 1, dx = tap % 3 - 1;
                uint4 braw = make_uint4(0u, 0u, 0u, 0u);
                uint16_t* bh = reinterpret_cast<uint16_t*>(&braw);
                if (tap < 9) {
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const int p = 16 * s + 8 * half + j, py = p >> 5, px = p & 31;
                        bh[j] = xt[((py + 1 + dy) * LX + (px + 1 + dx)) * PS + ci];
                    }
                }
                acc[k] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, as_bf16x8(braw), acc[k], 0, 0, 0);
            }
        }
        if (tid < 32)
            for (int p = 0; p < 256; p++) bsum += bf2f(dt[p * 33 + tid]);
    }
    float* out = partials + (size_t)blockIdx.x * STRIDE;
#pragma unroll
    for (int k = 0; k < PW; k++) {
        const int d = wave + 4 * k;
        if (d >= NT) continue;
        const int tap = CIN == 32 ? d : 4 * d + (col >> 3);
        const int ci = CIN == 32 ? col : (col & 7);
        if (tap >= 9) continue;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int co0 = 8 * j + 4 * half;
            *reinterpret_cast<float4*>(out + ((size_t)tap * CIN + ci) * 32 + co0) =
                make_float4(acc[k][4 * j], acc[k][4 * j + 1], acc[k][4 * j + 2], acc[k][4 * j + 3]);
        }
    }
    if (tid < 32) out[9 * CIN * 32 + tid] = bsum;
}

// the blocks' partials (scratch layout [tap][ci][32] + bias[32]) summed in
// block order into the canonical layout W[tap][ci][cout], b[cout]
__global__ void k_sn_wreduce(const float* partials, int cin, int cout, int nblocks, float* grad) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int nw = 9 * cin * cout;
    if (e >= nw + cout) return;
    const int stride = 9 * cin * 32 + 32;
    const int src = e < nw ? (e / cout) * 32 + e % cout : 9 * cin * 32 + (e - nw);
    float s = 0.0f;
    for (int b = 0; b < nblocks; b++) s += partials[(size_t)b * stride + src];
    grad[e] = s;
}

void launch_sn_wgrad(const uint16_t* x, const uint16_t* dz, int n, int h, int w, int cin, int cout, float* partials,
                     float* grad, hipStream_t st) {
    if (cin == 32)
        hipLaunchKernelGGL(k_sn_wgrad<32>, dim3(kSnWgradBlocks), dim3(256), 0, st, x, dz, n, h, w, partials);
    else
        hipLaunchKernelGGL(k_sn_wgrad<8>, dim3(kSnWgradBlocks), dim3(256), 0, st, x, dz, n, h, w, partials);
    const int ne = 9 * cin * cout + cout;
    hipLaunchKernelGGL(k_sn_wreduce, dim3((ne + 255) / 256), dim3(256), 0, st, partials, cin, cout, kSnWgradBlocks,
                       grad);
}

// ------------------------------------------------------- weights and Adam
__global__ void k_sn_prep(SnWeights wt) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    // forward operands: layer 0 [32][80], layers 1..3 [32][288]
    int base = 0;
    for (int l = 0; l < kSnLayers; l++) {
        const int cin = sn_cin(l), cout = sn_cout(l), kp = sn_kpad(l), n = 32 * kp;
        if (e >= base && e < base + n) {
            const int co = (e - base) / kp, k = (e - base) % kp;
            float v = 0.0f;
            if (co < cout && k < 9 * cin) v = wt.master[sn_layer_offset(l) + k * cout + co];
            wt.fwd[l][co * kp + k] = f2bf(v);
            return;
        }
        base += n;
    }
    // input-gradient operands: layers 1..3 [32][288], row ci, k = (8 - tap) 32 + co
    for (int l = 1; l < kSnLayers; l++) {
        const int cout = sn_cout(l), n = 32 * 288;
        if (e >= base && e < base + n) {
            const int ci = (e - base) / 288, k = (e - base) % 288;
            const int tap = 8 - k / 32, co = k % 32;
            float v = 0.0f;
            if (co < cout) v = wt.master[sn_layer_offset(l) + (tap * 32 + ci) * cout + co];
            wt.dgrad[l][ci * 288 + k] = f2bf(v);
            return;
        }
        base += n;
    }
    for (int l = 0; l < kSnLayers; l++) {
        if (e >= base && e < base + 32) {
            const int co = e - base, cout = sn_cout(l);
            wt.bias[l][co] = co < cout ? wt.master[sn_layer_offset(l) + 9 * sn_cin(l) * cout + co] : 0.0f;
            return;
        }
        base += 32;
    }
}

void launch_sn_prep(const SnWeights& wt, hipStream_t st) {
    int n = 0;
    for (int l = 0; l < kSnLayers; l++) n += 32 * sn_kpad(l) + 32;
    n += 3 * 32 * 288;
    hipLaunchKernelGGL(k_sn_prep, dim3((n + 255) / 256), dim3(256), 0, st, wt);
}

__global__ void k_sn_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                          float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void launch_sn_adam(float* master, const float* grad, float* m, float* v, int n, float lr, float b1, float b2,
                    float eps, int t, hipStream_t st) {
    const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_sn_adam, dim3((n + 255) / 256), dim3(256), 0, st, master, grad, m, v, n, lr, b1, b2, eps, c1,
                       c2);
}

// --------------------------------------------------------------------- loss
// Per pixel and (level, filter) with normaliser Z > 0 (the crop's mean error
// at the plain strength): the strength loss E(s) / Z, E the error curve
// interpolated linearly in s = log2 m between its seven knots s_k = (k - 3)/2
// (s clamped to the knots' span, the excess penalised by (s - clamp)^2 / 2),
// and the prediction loss lambda (r - t)^2 with the target
// t = log2((E(s) + delta) / (E(0) + delta)), delta = Z / 100 (t held fixed).
constexpr float kSnLambda = 0.1f;

__global__ void __launch_bounds__(256) k_sn_loss(const float* out32, const float* curves, const float* norm, int n,
                                                 int hw, float scale, uint16_t* dz, float* partials) {
    __shared__ float red[2][256];
    const int tid = threadIdx.x;
    const long long p = (long long)blockIdx.x * 256 + tid;
    float ls = 0.0f, lr = 0.0f;
    float g[kSnOut];
#pragma unroll
    for (int c = 0; c < kSnOut; c++) g[c] = 0.0f;
    const bool ok = p < (long long)n * hw;
    if (ok) {
        const int img = (int)(p / hw);
        const float* o = out32 + p * 32;
        const float* E = curves + p * kSnCurves;
#pragma unroll
        for (int l = 0; l < kSnLevels; l++) {
#pragma unroll
            for (int f = 0; f < 2; f++) {
                const float Z = norm[(img * kSnLevels + l) * 2 + f];
                if (!(Z > 0.0f)) continue;
                const float* e = E + (l * 2 + f) * kSnCurveK;
                const float s = o[4 * l + f];
                const float sc = fminf(fmaxf(s, -kSnSMax), kSnSMax);
                const float t = (sc + kSnSMax) * 2.0f;
                const int k = min((int)t, kSnCurveK - 2);
                const float fr = t - (float)k;
                const float Es = e[k] + fr * (e[k + 1] - e[k]);
                float gs = (e[k + 1] - e[k]) * 2.0f / Z;
                ls += Es / Z;
                if (s != sc) {
                    gs = s - sc;
                    ls += 0.5f * (s - sc) * (s - sc);
                }
                const float dl = 0.01f * Z;
                const float tgt = log2f((Es + dl) / (e[3] + dl));
                const float rr = o[4 * l + 2 + f];
                lr += kSnLambda * (rr - tgt) * (rr - tgt);
                g[4 * l + f] = gs * scale;
                g[4 * l + 2 + f] = 2.0f * kSnLambda * (rr - tgt) * scale;
            }
        }
        uint32_t wv[16];
#pragma unroll
        for (int c = 0; c < 16; c++) {
            const float a0 = 2 * c < kSnOut ? g[2 * c] : 0.0f, a1 = 2 * c + 1 < kSnOut ? g[2 * c + 1] : 0.0f;
            wv[c] = (uint32_t)f2bf(a0) | ((uint32_t)f2bf(a1) << 16);
        }
        uint4* d = reinterpret_cast<uint4*>(dz + p * 32);
#pragma unroll
        for (int q = 0; q < 4; q++) d[q] = make_uint4(wv[4 * q], wv[4 * q + 1], wv[4 * q + 2], wv[4 * q + 3]);
    }
    red[0][tid] = ls;
    red[1][tid] = lr;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            red[0][tid] += red[0][tid + s];
            red[1][tid] += red[1][tid + s];
        }
        __syncthreads();
    }
    if (tid == 0) {
        partials[2 * blockIdx.x] = red[0][0];
        partials[2 * blockIdx.x + 1] = red[1][0];
    }
}

int launch_sn_loss(const float* out32, const float* curves, const float* norm, int n, int hw, float scale,
                   uint16_t* dz, float* loss_partials, hipStream_t st) {
    const long long P = (long long)n * hw;
    const int blocks = (int)((P + 255) / 256);
    hipLaunchKernelGGL(k_sn_loss, dim3(blocks), dim3(256), 0, st, out32, curves, norm, n, hw, scale, dz,
                       loss_partials);
    return blocks;
}

// ----------------------------------------------------------------- features
// Per texel of the grid level: 5x5 (clamped) mean and variance of Y, Cb,
// Cr, and the noise models' predicted variances at the mean luma:
//   f0 mean Y, f1..f3 log2(var / predicted var) / 4 for Y, Cb, Cr (the grid
//   level's model), f4, f5 log2(predicted var Y, Cb there) / 8 + 1, f6 the same
//   for Y at level 0, f7 = 1.
__device__ __forceinline__ float sn_pred(const float* nlf, int c, float y) {
    return fmaxf(nlf[c] + nlf[3 + c] * y, 1e-10f);
}

template <bool H>
__global__ void __launch_bounds__(256) k_sn_features(Img lv, const float* nlf0, const float* nlfg, uint16_t* feat) {
    const int x = blockIdx.x * 32 + (threadIdx.x & 31), y = blockIdx.y * 8 + (threadIdx.x >> 5);
    if (x >= lv.w || y >= lv.h) return;
    float s[3] = {0, 0, 0}, q[3] = {0, 0, 0};
    for (int j = -2; j <= 2; j++) {
        for (int i = -2; i <= 2; i++) {
            const float4 v = ld4_<H>(lv, x + i, y + j);
            s[0] += v.x; s[1] += v.y; s[2] += v.z;
            q[0] += v.x * v.x; q[1] += v.y * v.y; q[2] += v.z * v.z;
        }
    }
    const float mY = s[0] / 25.0f;
    float f[8];
    f[0] = mY;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float m = s[c] / 25.0f, var = q[c] / 25.0f - m * m;
        f[1 + c] = log2f(fmaxf(var, 1e-12f) / sn_pred(nlfg, c, mY)) * 0.25f;
    }
    f[4] = log2f(sn_pred(nlfg, 0, mY)) * 0.125f + 1.0f;
    f[5] = log2f(sn_pred(nlfg, 1, mY)) * 0.125f + 1.0f;
    f[6] = log2f(sn_pred(nlf0, 0, mY)) * 0.125f + 1.0f;
    f[7] = 1.0f;
    uint4 o;
    o.x = (uint32_t)f2bf(f[0]) | ((uint32_t)f2bf(f[1]) << 16);
    o.y = (uint32_t)f2bf(f[2]) | ((uint32_t)f2bf(f[3]) << 16);
    o.z = (uint32_t)f2bf(f[4]) | ((uint32_t)f2bf(f[5]) << 16);
    o.w = (uint32_t)f2bf(f[6]) | ((uint32_t)f2bf(f[7]) << 16);
    *reinterpret_cast<uint4*>(feat + ((size_t)y * lv.w + x) * 8) = o;
}

void launch_sn_features(const Img& lvg, const float* nlf0, const float* nlfg, uint16_t* feat, hipStream_t st) {
    const dim3 grd((lvg.w + 31) / 32, (lvg.h + 7) / 8);
    with_storage(lvg, [&](auto H) {
        hipLaunchKernelGGL((k_sn_features<H()>), grd, dim3(256), 0, st, lvg, nlf0, nlfg, feat);
    });
}

// ------------------------------------------------------------ error curves
struct Mat3 {
    float m[9];
};

__global__ void __launch_bounds__(256) k_sn_curve(Img den, Img clean, Mat3 y2c, int level, int gw, int gh,
                                                  float* curves, int slot) {
    const int gx = blockIdx.x * 32 + (threadIdx.x & 31), gy = blockIdx.y * 8 + (threadIdx.x >> 5);
    if (gx >= gw || gy >= gh) return;
    auto err = [&](int x, int y) {
        const float4 a = row_f4(den, y)[x], b = row_f4(clean, y)[x];
        const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
        float e = 0.0f;
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const float v = y2c.m[3 * r] * d0 + y2c.m[3 * r + 1] * d1 + y2c.m[3 * r + 2] * d2;
            e += v * v;
        }
        return e;
    };
    float e = 0.0f;
    if (level <= kSnGrid) {
        const int sx = 1 << (kSnGrid - level);
        for (int j = 0; j < sx; j++)
            for (int i = 0; i < sx; i++) e += err(gx * sx + i, gy * sx + j);
        e /= (float)(sx * sx);
    } else {
        const int sh = level - kSnGrid;
        e = err(min(gx >> sh, den.w - 1), min(gy >> sh, den.h - 1));
    }
    curves[((size_t)gy * gw + gx) * kSnCurves + slot] = e;
}

void launch_sn_curve(const Img& den, const Img& clean, const float y2c[9], int level, int gw, int gh, float* curves,
                     int slot, hipStream_t st) {
    Mat3 m;
    for (int i = 0; i < 9; i++) m.m[i] = y2c[i];
    const dim3 grd((gw + 31) / 32, (gh + 7) / 8);
    hipLaunchKernelGGL(k_sn_curve, grd, dim3(256), 0, st, den, clean, m, level, gw, gh, curves, slot);
}

// ---------------------------------------------------------- training sets
// one block per crop: features and curves copied into the set, the crop's
// normalisers (mean plain-strength error per level and filter) reduced in a
// fixed order
__global__ void __launch_bounds__(256) k_sn_gather(const uint16_t* feat, const float* curves, int gw, int cs,
                                                   const int2* origins, uint16_t* feat_out, float* curves_out,
                                                   float* norm_out) {
    __shared__ float red[10][256];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int2 o = origins[c];
    float acc[10];
#pragma unroll
    for (int q = 0; q < 10; q++) acc[q] = 0.0f;
    for (int i = tid; i < cs * cs; i += 256) {
        const int x = o.x + i % cs, y = o.y + i / cs;
        const size_t src = (size_t)y * gw + x, dst = (size_t)c * cs * cs + i;
        *reinterpret_cast<uint4*>(feat_out + dst * 8) = *reinterpret_cast<const uint4*>(feat + src * 8);
        const float* e = curves + src * kSnCurves;
        float* d = curves_out + dst * kSnCurves;
        for (int k = 0; k < kSnCurves; k++) d[k] = e[k];
#pragma unroll
        for (int q = 0; q < 10; q++) acc[q] += e[q * kSnCurveK + 3];
    }
#pragma unroll
    for (int q = 0; q < 10; q++) red[q][tid] = acc[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s)
#pragma unroll
            for (int q = 0; q < 10; q++) red[q][tid] += red[q][tid + s];
        __syncthreads();
    }
    if (tid < 10) norm_out[(size_t)c * 10 + tid] = red[tid][0] / (float)(cs * cs);
}

void launch_sn_gather(const uint16_t* feat, const float* curves, int gw, int cs, const int2* origins, int ncrops,
                      uint16_t* feat_out, float* curves_out, float* norm_out, hipStream_t st) {
    hipLaunchKernelGGL(k_sn_gather, dim3(ncrops), dim3(256), 0, st, feat, curves, gw, cs, origins, feat_out,
                       curves_out, norm_out);
}

__global__ void __launch_bounds__(256) k_sn_batch(const uint16_t* feat_set, const float* curves_set,
                                                  const float* norm_set, const int* idx, int cs, uint16_t* feat_out,
                                                  float* curves_out, float* norm_out) {
    const int b = blockIdx.x, tid = threadIdx.x;
    const size_t src = (size_t)idx[b] * cs * cs, dst = (size_t)b * cs * cs;
    for (int i = tid; i < cs * cs; i += 256)
        *reinterpret_cast<uint4*>(feat_out + (dst + i) * 8) = *reinterpret_cast<const uint4*>(feat_set + (src + i) * 8);
    const size_t nc = (size_t)cs * cs * kSnCurves;
    for (size_t i = tid; i < nc; i += 256) curves_out[dst * kSnCurves + i] = curves_set[src * kSnCurves + i];
    if (tid < 10) norm_out[b * 10 + tid] = norm_set[(size_t)idx[b] * 10 + tid];
}

void launch_sn_batch(const uint16_t* feat_set, const float* curves_set, const float* norm_set, const int* idx,
                     int nb, int cs, uint16_t* feat_out, float* curves_out, float* norm_out, hipStream_t st) {
    hipLaunchKernelGGL(k_sn_batch, dim3(nb), dim3(256), 0, st, feat_set, curves_set, norm_set, idx, cs, feat_out,
                       curves_out, norm_out);
}

// --------------------------------------------------------------------- gate
// Reads the predicted-error sums of the frame, decides per level and clears
// the sums for the next frame (vector stores).
__global__ void k_sn_gate(long long* sum, int levels, int mode, int* open) {
    const int l = threadIdx.x;
    if (l >= kSnLevels) return;
    const long long s = sum[l];
    int o = 0;
    if (l < levels) o = mode == 3 ? 1 : mode == 2 ? 0 : (s < 0 ? 1 : 0);
    open[l] = o;
    open[kSnLevels + l] = (int)(s >> 32);          // the sum, for the host's read-back
    open[2 * kSnLevels + l] = (int)(s & 0xFFFFFFFFll);
    sum[l] = 0;
}

void launch_sn_gate(const long long* sum, int levels, int mode, int* open, hipStream_t st) {
    hipLaunchKernelGGL(k_sn_gate, dim3(1), dim3(64), 0, st, const_cast<long long*>(sum), levels, mode, open);
}

}  // namespace clim
