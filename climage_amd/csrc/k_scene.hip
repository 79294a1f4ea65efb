// k_scene.hip - seeded synthetic RAW scenes on the device.
//
// Scene (all FP32, linear RGB):
//   background per channel, u = (x+0.5)/W, v = (y+0.5)/H:
//     base + gx(u-1/2) + gy(v-1/2) + gxy(u-1/2)(v-1/2) + amp cos(2pi(fx u + fy v + ph)),
//     clamped to [0.05, 0.6]
//   200 Gaussian blobs (sigma 5..60 px, amplitude -0.3..0.6 per channel,
//     evaluated within +-4 sigma), then 50 hard-edged rectangles (sides
//     2..20 % of the frame, colour U[0,1]^3) painted in order, clipped to [0,1].
// Noise y = a*Poisson(x/a) + N(0, b): Poisson by sequential inversion for
// lambda < 64, else rint(lambda + sqrt(lambda) z); z by Box-Muller.
// DN = clamp(rint(black + y (white - black)), 0, white).
// Random numbers: Philox4x32-10; parameter k = word k%4 of counter (k/4,0,0,0)
// under key (scene_seed, 0x53434E45); pixel i = y*W + x uses counter
// (i,0,0,0) under key (noise_seed, 0x4E4F4953).
#include "kernels.h"

namespace clim {

static float param_u(uint32_t seed, uint32_t k) {
    clim_u32x4 ctr = {{k / 4u, 0u, 0u, 0u}};
    clim_u32x4 r = clim_philox4x32_10(ctr, seed, 0x53434E45u);
    return clim_u01(r.v[k % 4u]);
}

void scene_draw_params(int w, int h, uint32_t seed, SceneParams* p) {
    uint32_t k = 0;
    for (int c = 0; c < 3; c++) {
        p->base[c] = 0.15f + 0.25f * param_u(seed, k++);
        p->gx[c] = 0.2f * (param_u(seed, k++) - 0.5f);
        p->gy[c] = 0.2f * (param_u(seed, k++) - 0.5f);
        p->gxy[c] = 0.2f * (param_u(seed, k++) - 0.5f);
        p->amp[c] = 0.08f * param_u(seed, k++);
        p->fx[c] = 0.5f + 2.5f * param_u(seed, k++);
        p->fy[c] = 0.5f + 2.5f * param_u(seed, k++);
        p->ph[c] = param_u(seed, k++);
    }
    for (int b = 0; b < kSceneBlobs; b++) {
        p->bcx[b] = (float)w * param_u(seed, k++);
        p->bcy[b] = (float)h * param_u(seed, k++);
        const float sigma = 5.0f + 55.0f * param_u(seed, k++);
        p->bw[b] = 4.0f * sigma;
        p->binv[b] = 1.0f / (2.0f * sigma * sigma);
        for (int c = 0; c < 3; c++) p->ba[b][c] = -0.3f + 0.9f * param_u(seed, k++);
    }
    for (int r = 0; r < kSceneRects; r++) {
        const float rw = (float)w * (0.02f + 0.18f * param_u(seed, k++));
        const float rh = (float)h * (0.02f + 0.18f * param_u(seed, k++));
        p->rx0[r] = ((float)w - rw) * param_u(seed, k++);
        p->ry0[r] = ((float)h - rh) * param_u(seed, k++);
        p->rx1[r] = p->rx0[r] + rw;
        p->ry1[r] = p->ry0[r] + rh;
        for (int c = 0; c < 3; c++) p->rc[r][c] = param_u(seed, k++);
    }
}

__device__ __forceinline__ float gauss(uint32_t r0, uint32_t r1) {
    return sqrtf(-2.0f * clim_logf(clim_u01_open0(r0))) * clim_cos2pif(clim_u01(r1));
}

__global__ void __launch_bounds__(256) k_scene(Img raw, Img clean, const SceneParams* __restrict__ p, int colors,
                                               uint32_t noise_seed, float a, float b, float black, float white) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    const int w = raw.p ? raw.w : clean.w, h = raw.p ? raw.h : clean.h;
    if (x >= w || y >= h) return;
    const float fxp = (float)x + 0.5f, fyp = (float)y + 0.5f;
    const float u = fxp / (float)w, v = fyp / (float)h;
    const float du = u - 0.5f, dv = v - 0.5f;
    float o[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float bg = p->base[c] + p->gx[c] * du + p->gy[c] * dv + p->gxy[c] * (du * dv) +
                         p->amp[c] * clim_cos2pif(p->fx[c] * u + p->fy[c] * v + p->ph[c]);
        o[c] = clim_clampf(bg, 0.05f, 0.6f);
    }
    for (int k = 0; k < kSceneBlobs; k++) {
        const float dx = fxp - p->bcx[k], dy = fyp - p->bcy[k];
        if (fabsf(dx) > p->bw[k] || fabsf(dy) > p->bw[k]) continue;
        const float g = clim_expf(-(dx * dx + dy * dy) * p->binv[k]);
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = o[c] + p->ba[k][c] * g;
    }
    for (int r = 0; r < kSceneRects; r++) {
        if (fxp >= p->rx0[r] && fxp < p->rx1[r] && fyp >= p->ry0[r] && fyp < p->ry1[r]) {
#pragma unroll
            for (int c = 0; c < 3; c++) o[c] = p->rc[r][c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) o[c] = clim_clampf(o[c], 0.0f, 1.0f);
    if (clean.p) wrow_f4(clean, y)[x] = make_float4(o[0], o[1], o[2], 0.0f);
    if (!raw.p) return;

    const int col = (colors >> ((((y & 1) << 1) | (x & 1)) << 1)) & 3;
    const float xs = col == 0 ? o[0] : col == 1 ? o[1] : o[2];
    clim_u32x4 ctr = {{(uint32_t)((size_t)y * w + x), 0u, 0u, 0u}};
    const clim_u32x4 rr = clim_philox4x32_10(ctr, noise_seed, 0x4E4F4953u);
    const float lambda = xs / a;
    float kf;
    if (lambda < 64.0f) {
        const float uu = clim_u01(rr.v[0]);
        float pk = clim_expf(-lambda);
        float F = pk;
        int kk = 0;
        while (uu > F && kk < 200) {
            kk++;
            pk = pk * lambda / (float)kk;
            F = F + pk;
        }
        kf = (float)kk;
    } else {
        kf = fmaxf(rintf(lambda + sqrtf(lambda) * gauss(rr.v[0], rr.v[1])), 0.0f);
    }
    const float yv = a * kf + sqrtf(b) * gauss(rr.v[2], rr.v[3]);
    const float dn = rintf(black + yv * (white - black));
    reinterpret_cast<uint16_t*>(raw.p + (size_t)y * raw.pitch)[x] = (uint16_t)clim_clampf(dn, 0.0f, white);
}

void launch_scene(const Img& raw, const Img& clean, const SceneParams* dparams, int colors, uint32_t noise_seed,
                  float a, float b, float black, float white, hipStream_t st) {
    const int w = raw.p ? raw.w : clean.w, h = raw.p ? raw.h : clean.h;
    dim3 blk(64, 4), grd((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(k_scene, grd, blk, 0, st, raw, clean, dparams, colors, noise_seed, a, b, black, white);
}

}  // namespace clim
