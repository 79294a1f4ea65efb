/*
 * scene_oracle.c - CPU restatement of the seeded synthetic RAW scenes.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the device-side generator).
 *
 * The reference has no synthetic-scene generator; its only synthetic-input
 * pattern mosaics a PNG into a Bayer DNG (cpp/raw_pipeline.cpp:148-214).
 * The scene below is the one the benchmark configurations specify:
 *   background  per channel c, u = (x+0.5)/W, v = (y+0.5)/H:
 *               base + gx(u-1/2) + gy(v-1/2) + gxy(u-1/2)(v-1/2)
 *                    + amp*cos(2pi(fx u + fy v + phase)), clamped to [0.05, 0.6]
 *   200 blobs   centre uniform in the frame, sigma ~ U[5, 60] px, per-channel
 *               amplitude ~ U[-0.3, 0.6], evaluated inside +-4 sigma
 *   50 rects    sides U[2%, 20%] of the frame, uniform position, colour U[0,1]^3,
 *               painted last, in order
 *   clip        [0, 1]
 *   noise       y = a*Poisson(x/a) + N(0, b): Poisson by sequential inversion
 *               for lambda < 64, else rint(lambda + sqrt(lambda) z) (normal
 *               approximation); N(0,1) by Box-Muller
 *   quantise    DN = clamp(rint(black + y (white - black)), 0, white)
 * Parameters: uniform k of the parameter stream is word k%4 of
 * Philox4x32-10(counter (k/4, 0, 0, 0), key (scene_seed, 0x53434E45)).
 * Noise at pixel i = y*W + x: Philox4x32-10(counter (i, 0, 0, 0),
 * key (noise_seed, 0x4E4F4953)).
 */
#include <stdint.h>
#include <stdlib.h>
#include <math.h>

#include "clim_math.h"

#define N_BLOBS 200
#define N_RECTS 50

typedef struct {
    float base[3], gx[3], gy[3], gxy[3], amp[3], fx[3], fy[3], ph[3];
    float bcx[N_BLOBS], bcy[N_BLOBS], bw[N_BLOBS], binv[N_BLOBS], ba[N_BLOBS][3];
    float rx0[N_RECTS], ry0[N_RECTS], rx1[N_RECTS], ry1[N_RECTS], rc[N_RECTS][3];
} scene_params;

static float param_u(uint32_t seed, uint32_t k) {
    clim_u32x4 ctr = {{k / 4u, 0u, 0u, 0u}};
    clim_u32x4 r = clim_philox4x32_10(ctr, seed, 0x53434E45u);
    return clim_u01(r.v[k % 4u]);
}

static void draw_params(int w, int h, uint32_t seed, scene_params* p) {
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
    for (int b = 0; b < N_BLOBS; b++) {
        p->bcx[b] = (float)w * param_u(seed, k++);
        p->bcy[b] = (float)h * param_u(seed, k++);
        float sigma = 5.0f + 55.0f * param_u(seed, k++);
        p->bw[b] = 4.0f * sigma;
        p->binv[b] = 1.0f / (2.0f * sigma * sigma);
        for (int c = 0; c < 3; c++) p->ba[b][c] = -0.3f + 0.9f * param_u(seed, k++);
    }
    for (int r = 0; r < N_RECTS; r++) {
        float rw = (float)w * (0.02f + 0.18f * param_u(seed, k++));
        float rh = (float)h * (0.02f + 0.18f * param_u(seed, k++));
        p->rx0[r] = ((float)w - rw) * param_u(seed, k++);
        p->ry0[r] = ((float)h - rh) * param_u(seed, k++);
        p->rx1[r] = p->rx0[r] + rw;
        p->ry1[r] = p->ry0[r] + rh;
        for (int c = 0; c < 3; c++) p->rc[r][c] = param_u(seed, k++);
    }
}

static void scene_pixel(const scene_params* p, int w, int h, int x, int y, float out[3]) {
    float fxp = (float)x + 0.5f, fyp = (float)y + 0.5f;
    float u = fxp / (float)w, v = fyp / (float)h;
    float du = u - 0.5f, dv = v - 0.5f;
    for (int c = 0; c < 3; c++) {
        float bg = p->base[c] + p->gx[c] * du + p->gy[c] * dv + p->gxy[c] * (du * dv) +
                   p->amp[c] * clim_cos2pif(p->fx[c] * u + p->fy[c] * v + p->ph[c]);
        out[c] = clim_clampf(bg, 0.05f, 0.6f);
    }
    for (int b = 0; b < N_BLOBS; b++) {
        float dx = fxp - p->bcx[b], dy = fyp - p->bcy[b];
        if (fabsf(dx) > p->bw[b] || fabsf(dy) > p->bw[b]) continue;
        float g = clim_expf(-(dx * dx + dy * dy) * p->binv[b]);
        for (int c = 0; c < 3; c++) out[c] = out[c] + p->ba[b][c] * g;
    }
    for (int r = 0; r < N_RECTS; r++) {
        if (fxp >= p->rx0[r] && fxp < p->rx1[r] && fyp >= p->ry0[r] && fyp < p->ry1[r]) {
            for (int c = 0; c < 3; c++) out[c] = p->rc[r][c];
        }
    }
    for (int c = 0; c < 3; c++) out[c] = clim_clampf(out[c], 0.0f, 1.0f);
}

static float gauss(uint32_t r0, uint32_t r1) {
    return sqrtf(-2.0f * clim_logf(clim_u01_open0(r0))) * clim_cos2pif(clim_u01(r1));
}

/* CFA colour of site (x, y): 0 = R, 1 = G, 2 = B, from the OpenCL table
 * (cl/demosaic.cl:29-34); both greens are G. */
static int cfa_color(int bayer, int x, int y) {
    static const int off[4][4][2] = {
        {{1, 0}, {0, 0}, {0, 1}, {1, 1}},
        {{0, 1}, {0, 0}, {1, 0}, {1, 1}},
        {{0, 0}, {0, 1}, {1, 1}, {1, 0}},
        {{1, 1}, {0, 1}, {0, 0}, {1, 0}},
    };
    static const int color_of[4] = {0, 1, 2, 1};
    for (int c = 0; c < 4; c++)
        if ((x & 1) == off[bayer][c][0] && (y & 1) == off[bayer][c][1]) return color_of[c];
    return 1;
}

void oscene_generate(int w, int h, int bayer, uint32_t scene_seed, uint32_t noise_seed, float a, float b,
                     float black, float white, uint16_t* raw, float* clean_rgba) {
    scene_params* p = (scene_params*)malloc(sizeof(scene_params));
    draw_params(w, h, scene_seed, p);
    const float sb = sqrtf(b);
#pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < h; y++) {
        for (int x = 0; x < w; x++) {
            float rgb[3];
            scene_pixel(p, w, h, x, y, rgb);
            size_t i = (size_t)y * w + x;
            if (clean_rgba) {
                clean_rgba[i * 4 + 0] = rgb[0];
                clean_rgba[i * 4 + 1] = rgb[1];
                clean_rgba[i * 4 + 2] = rgb[2];
                clean_rgba[i * 4 + 3] = 0.0f;
            }
            if (!raw) continue;
            float xs = rgb[cfa_color(bayer, x, y)];
            clim_u32x4 ctr = {{(uint32_t)i, 0u, 0u, 0u}};
            clim_u32x4 r = clim_philox4x32_10(ctr, noise_seed, 0x4E4F4953u);
            float lambda = xs / a;
            float k;
            if (lambda < 64.0f) {
                float u = clim_u01(r.v[0]);
                float pk = clim_expf(-lambda);
                float F = pk;
                int kk = 0;
                while (u > F && kk < 200) {
                    kk++;
                    pk = pk * lambda / (float)kk;
                    F = F + pk;
                }
                k = (float)kk;
            } else {
                k = fmaxf(rintf(lambda + sqrtf(lambda) * gauss(r.v[0], r.v[1])), 0.0f);
            }
            float yv = a * k + sb * gauss(r.v[2], r.v[3]);
            float dn = rintf(black + yv * (white - black));
            raw[i] = (uint16_t)clim_clampf(dn, 0.0f, white);
        }
    }
    free(p);
}
