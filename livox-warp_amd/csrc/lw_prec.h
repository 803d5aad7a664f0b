// Per-voxel mean and residual precision, computed once per voxel when its moments change instead of
// once per point and iteration. The reference evaluates the same quantities inside every k_icp
// thread (python/livox_warp/odom.py:325, :336-368); they depend only on the voxel's moments and the
// fixed parameters (voxel, noise^2, tight^2, alpha), so caching them leaves every residual unchanged.
#pragma once

#include "lw_gn.h"

#ifndef LW_HD
#define LW_HD
#endif

namespace lw {

// Mean (odom.py:325, float32: centre + sum / count) of a voxel with axis indices (ix, iy, iz).
LW_HD inline void voxel_mean(int32_t ix, int32_t iy, int32_t iz, const float *s, int32_t c, float voxel, float *mu) {
    float cf = (float)c;
    mu[0] = ((float)ix + 0.5f) * voxel + s[0] / cf;
    mu[1] = ((float)iy + 0.5f) * voxel + s[1] / cf;
    mu[2] = ((float)iz + 0.5f) * voxel + s[2] / cf;
}

// odom.py:336-368 for one voxel: M, Mg (xx xy xz yy yz zz), the tightest direction q; returns 1 for the
// shape branch (rn = r.q), 0 for the weak isotropic branch (rn = |r|). The covariance is formed in
// float32 exactly as the reference forms it; its eigen-decomposition and the weights are float64 and
// rounded to float32 once.
LW_HD inline int voxel_precision(const float *s, const float *o6, int32_t c, float voxel, float noise2, float tight2,
                                 float alpha, float *M6, float *Mg6, float *q3) {
    float cf = (float)c;
    float m[3] = {s[0] / cf, s[1] / cf, s[2] / cf};
    // o6: xx xy xz yy yz zz
    double cov[9];
    cov[0] = (double)(o6[0] / cf - m[0] * m[0]);
    cov[1] = cov[3] = (double)(o6[1] / cf - m[0] * m[1]);
    cov[2] = cov[6] = (double)(o6[2] / cf - m[0] * m[2]);
    cov[4] = (double)(o6[3] / cf - m[1] * m[1]);
    cov[5] = cov[7] = (double)(o6[4] / cf - m[1] * m[2]);
    cov[8] = (double)(o6[5] / cf - m[2] * m[2]);
    double lam[3], V[9];
    gn::sym_eig<3>(cov, lam, V);
    int nt = 0;
    LW_UNROLL for (int a = 0; a < 3; ++a) nt += lam[a] < (double)tight2 ? 1 : 0;
    if (c >= 5 && nt <= 1) {
        // the tightest direction: first index of the smallest eigenvalue (odom.py:348-352)
        double q0 = V[0], q1 = V[3], q2 = V[6], lmin = lam[0];
        if (lam[1] < lmin) { lmin = lam[1]; q0 = V[1]; q1 = V[4]; q2 = V[7]; }
        if (lam[2] < lmin) { lmin = lam[2]; q0 = V[2]; q1 = V[5]; q2 = V[8]; }
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, G[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        LW_UNROLL for (int a = 0; a < 3; ++a) {
            double e = (lam[a] > 0.0 ? lam[a] : 0.0) + (double)noise2;
            bool tight = !(lam[a] >= (double)tight2);
            double wa = tight ? 1.0 / e : (double)alpha / e;
            LW_UNROLL for (int i = 0; i < 3; ++i)
                LW_UNROLL for (int j = 0; j < 3; ++j) {
                    double t = wa * V[3 * i + a] * V[3 * j + a];
                    M[3 * i + j] += t;
                    G[3 * i + j] += tight ? t : 0.0;
                }
        }
        M6[0] = (float)M[0]; M6[1] = (float)M[1]; M6[2] = (float)M[2];
        M6[3] = (float)M[4]; M6[4] = (float)M[5]; M6[5] = (float)M[8];
        Mg6[0] = (float)G[0]; Mg6[1] = (float)G[1]; Mg6[2] = (float)G[2];
        Mg6[3] = (float)G[4]; Mg6[4] = (float)G[5]; Mg6[5] = (float)G[8];
        q3[0] = (float)q0; q3[1] = (float)q1; q3[2] = (float)q2;
        return 1;
    }
    float wa = alpha / (0.25f * voxel * voxel + noise2);
    M6[0] = wa; M6[1] = 0.0f; M6[2] = 0.0f; M6[3] = wa; M6[4] = 0.0f; M6[5] = wa;
    for (int k = 0; k < 6; ++k) Mg6[k] = 0.0f;
    q3[0] = q3[1] = q3[2] = 0.0f;
    return 0;
}

}  // namespace lw
