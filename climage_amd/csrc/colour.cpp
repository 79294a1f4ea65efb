// colour.cpp - host-side colour math of the C ABI (no HIP): cam_ycbcr
// (cpp/demosaic_utils.cpp:809-822) and the 3x3 inverse (src/gls_linalg.hpp:544-557),
// both in FP64 and rounded to float.  Kept apart from the HIP translation
// units so that host-only builds (the DNG reader's sanitizer build) link it.
#include "climage.h"

extern "C" void clim_cam_ycbcr(const float rgb_cam[9], float out[9]) {
    // xyz_rgb row Y of cpp/demosaic.hpp:91-95 times rgb_cam gives K_R, K_G, K_B
    const double xyz_y[3] = {0.2126729, 0.7151522, 0.0721750};
    double k[3];
    for (int j = 0; j < 3; j++)
        k[j] = xyz_y[0] * (double)rgb_cam[0 * 3 + j] + xyz_y[1] * (double)rgb_cam[1 * 3 + j] +
               xyz_y[2] * (double)rgb_cam[2 * 3 + j];
    const double KR = k[0], KG = k[1], KB = k[2];
    const double m[9] = {KR, KG, KB,
                         -0.5 * KR / (1.0 - KB), -0.5 * KG / (1.0 - KB), 0.5,
                         0.5, -0.5 * KG / (1.0 - KR), -0.5 * KB / (1.0 - KR)};
    for (int i = 0; i < 9; i++) out[i] = (float)m[i];
}

// cofactor inverse in FP64, scaled, rounded to float; 1 when singular
int inverse3x3_scaled(const float mf[9], double scale, float out[9]) {
    double m[9];
    for (int i = 0; i < 9; i++) m[i] = mf[i];
    const double c00 = m[4] * m[8] - m[5] * m[7];
    const double c01 = m[5] * m[6] - m[3] * m[8];
    const double c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    if (det == 0.0) return 1;
    const double inv[9] = {c00, m[2] * m[7] - m[1] * m[8], m[1] * m[5] - m[2] * m[4],
                           c01, m[0] * m[8] - m[2] * m[6], m[2] * m[3] - m[0] * m[5],
                           c02, m[1] * m[6] - m[0] * m[7], m[0] * m[4] - m[1] * m[3]};
    for (int i = 0; i < 9; i++) out[i] = (float)(inv[i] / det * scale);
    return 0;
}

extern "C" clim_status clim_inverse3x3(const float m[9], float out[9]) {
    if (!m || !out) return CLIM_EINVAL;
    return inverse3x3_scaled(m, 1.0, out) ? CLIM_ENUMERIC : CLIM_OK;
}
