// The float64 Gauss-Newton step (lw_gn.h's gn::step) spread over the 64 lanes of one wave, for the common
// case: every direction constrained (J^T Hg J - thr I positive definite) and the damped system positive
// definite. Matrices live in LDS, one entry a lane, and the Cholesky factorisations go column by column with
// the column's entries in parallel; the substitutions keep one row of the factor a lane in registers and pass
// the unknowns by readlane. Each lane forms its values with the scalar routine's exact operation order (the
// same products summed in the same order, the same divisions and square roots), so the result equals
// gn::step's bit for bit. Anything else (a degenerate direction to hold, or an indefinite system needing the LU
// fallback) reports "not handled" and the caller runs gn::step in one lane.
//
// The scalar routine compiles to tens of kilobytes of straight-line code executed by one lane; this one is
// a few hundred instructions per lane.
#pragma once

#include <hip/hip_runtime.h>

#include "lw_common.h"
#include "lw_gn.h"

namespace lw {
namespace gnw {

// LDS layout (doubles) of the step's scratch
constexpr int W_H = 0, W_HG = 36, W_A = 72, W_J = 108, W_JI = 144, W_M1 = 180, W_M2 = 216, W_M3 = 252, W_L = 288,
              W_L2 = 324, W_RHS = 360, W_T = 366, W_DOUBLES = 382;

// Index of (p, q), p <= q, in a row-major upper triangle of a 6x6 matrix (odom.py:387, np.triu_indices(6)).
__device__ inline int triu(int p, int q) { return p * 6 - (p * (p - 1)) / 2 + (q - p); }

// Entry l (< 36) of C = A B, summed over k in order (gn::mat6_mul).
__device__ inline double mul_entry(const double *A, const double *B, int l) {
    const int i = l / 6, j = l % 6;
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += A[6 * i + k] * B[6 * k + j];
    return s;
}

// Entry l (< 36) of C = A^T B (gn::mat6_tmul).
__device__ inline double tmul_entry(const double *A, const double *B, int l) {
    const int i = l / 6, j = l % 6;
    double s = 0.0;
    for (int k = 0; k < 6; ++k) s += A[6 * k + i] * B[6 * k + j];
    return s;
}

// A double held by lane src (the same lane for every lane) in every lane.
__device__ inline double bcast(double x, int src) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | (uint64_t)lo));
}

// Two Cholesky factorisations side by side, column by column: lanes 0-31 factor A0 - shift0 I into L0, lanes
// 32-63 A1 - shift1 I into L1 (lower, LDS). In each half, every lane forms column j's pivot and lane i > j the
// entry below it (gn::pd_above / gn::solve6 order). pd0 / pd1: whether every pivot of each was positive.
__device__ inline void wave_cholesky2(const double *A0, double shift0, double *L0, const double *A1, double shift1,
                                      double *L1, bool &pd0, bool &pd1) {
    const int lane = lane_id();
    const int li = lane & 31;
    const double *A = lane < 32 ? A0 : A1;
    double *L = lane < 32 ? L0 : L1;
    const double shift = lane < 32 ? shift0 : shift1;
    bool pd = true;
    for (int j = 0; j < 6; ++j) {
        double s = A[6 * j + j] - shift;
        for (int k = 0; k < j; ++k) s -= L[6 * j + k] * L[6 * j + k];
        if (!(s > 0.0)) pd = false;
        const double d = sqrt(s > 0.0 ? s : 1.0);
        if (li > j && li < 6) {
            double t = A[6 * li + j];
            for (int k = 0; k < j; ++k) t -= L[6 * li + k] * L[6 * j + k];
            L[6 * li + j] = t / d;
        }
        if (li == j) L[6 * j + j] = d;
        wave_sync();
    }
    pd0 = __shfl(pd ? 1 : 0, 0, 64) != 0;
    pd1 = __shfl(pd ? 1 : 0, 32, 64) != 0;
}

// Solve L L^T x = rhs with lane r < 6 holding row r of L in a[0..r] and rhs[r]; every lane returns x. Forward: y_k = t_k / L_kk, then each row below subtracts L_rk y_k (so row r subtracts k = 0, 1, ...
// in order); back: x_i = (y_i - sum_{k > i} L_ki x_k) / L_ii with k ascending, on every lane. The operations
// and their order are gn::solve6's.
__device__ inline void wave_chol_solve(const double (&a)[6], double rhs_r, double (&x)[6]) {
    const int lane = lane_id();
    double t = rhs_r, y[6];
    LW_UNROLL for (int k = 0; k < 6; ++k) {
        y[k] = bcast(t, k) / bcast(a[k], k);
        if (lane < 6 && lane > k) t = t - a[k] * y[k];
    }
    LW_UNROLL for (int i = 5; i >= 0; --i) {
        double u = y[i];
        LW_UNROLL for (int k = i + 1; k < 6; ++k) u = u - bcast(a[i], k) * x[k];
        x[i] = u / bcast(a[i], i);
    }
}

// Entry (r, c) of K = [t]x for the pose T (row-major 4x4 in LDS), as gn::sensor_basis lays it out:
// {0, -t2, t1; t2, 0, -t0; -t1, t0, 0}. Selects, not a private array: a lane-dependent index into a private
// array would put it in scratch memory.
__device__ inline double skew_entry(const double *Ts, int r, int c) {
    const int e = 3 * r + c;
    double v = 0.0;
    v = e == 1 ? -Ts[11] : v;
    v = e == 2 ? Ts[7] : v;
    v = e == 3 ? Ts[11] : v;
    v = e == 5 ? -Ts[3] : v;
    v = e == 6 ? -Ts[7] : v;
    v = e == 7 ? Ts[3] : v;
    return v;
}

// gn::sensor_basis entry by entry (same products in the same order): J = [[R, [c]x R / L], [0, R / L]],
// J^-1 = [[R^T, -R^T [c]x], [0, L R^T]], with R = T[0:3, 0:3] and c = T[0:3, 3] (Ts: T in LDS).
__device__ inline void basis_entry(const double *Ts, double L, int l, double &j_out, double &ji_out) {
    const int i = l / 6, j = l % 6;
    j_out = 0.0;
    ji_out = 0.0;
    if (i < 3 && j < 3) {
        j_out = Ts[4 * i + j];
        ji_out = Ts[4 * j + i];
    } else if (i < 3) {
        const int jj = j - 3;
        const double kr = skew_entry(Ts, i, 0) * Ts[jj] + skew_entry(Ts, i, 1) * Ts[4 + jj] +
                          skew_entry(Ts, i, 2) * Ts[8 + jj];  // (K R)[i][jj], mat3_mul order
        j_out = kr / L;
        ji_out = -(Ts[i] * skew_entry(Ts, 0, jj) + Ts[4 + i] * skew_entry(Ts, 1, jj) +
                   Ts[8 + i] * skew_entry(Ts, 2, jj));  // -(R^T K)[i][jj]
    } else if (j >= 3) {
        const int ii = i - 3, jj = j - 3;
        j_out = Ts[4 * ii + jj] / L;
        ji_out = L * Ts[4 * jj + ii];
    }
}

struct WaveOut {
    double delta[6];  // in every lane
    double rms;
    int32_t corr;
    int32_t n_degen;
};

// The step on all 64 lanes of the calling wave (the other waves of its workgroup, if any, take no part). sh: W_DOUBLES doubles of LDS. Returns
// 1 when handled (lane 0's o holds delta, and H_out, the registration's copy of H, has been written), 0 when
// the caller must run gn::step instead, or -1 for gn::STEP_FEW (corr < min_corr). T: the pose. xi: the twist
// from T to the prediction (gn::twist_between(T, T_pred), read when P.prior_info > 0), computed by the caller
// (the persistent loop does it while the pass's search runs). With stamp set (measurement), lane 0 records
// the device wall clock in t_setup, t_factor and t_solve after the set-up, the two factorisations (the
// degeneracy test's and the damped system's, side by side) and the substitution.
__device__ inline int wave_step(const double *v, const double *T, const double *xi, const gn::StepParams &P,
                                double *sh, WaveOut &o, double *H_out, bool stamp, unsigned long long &t_setup,
                                unsigned long long &t_factor, unsigned long long &t_solve) {
    const int lane = lane_id();
    o.corr = (int32_t)v[27];
    o.n_degen = 0;
    if (o.corr < P.min_corr) return -1;
    double *H = sh + W_H, *Hg = sh + W_HG, *A = sh + W_A, *J = sh + W_J, *Ji = sh + W_JI, *M1 = sh + W_M1,
           *M2 = sh + W_M2, *M3 = sh + W_M3, *L = sh + W_L, *L2 = sh + W_L2, *rhs = sh + W_RHS, *Ts = sh + W_T;
    const bool prior = P.prior_info > 0.0, degen = P.degen_thr > 0.0;
    if (lane < 16) Ts[lane] = T[lane];
    wave_sync();
    double Lr = v[31] / (double)o.corr;
    if (!(Lr > 1.0)) Lr = 1.0;
    const double rr = v[29] > 0.0 ? v[29] : 0.0;
    const double ww = v[30] > 1e-9 ? v[30] : 1e-9;
    o.rms = sqrt(rr / ww);
    if (lane < 36) {
        const int i = lane / 6, j = lane % 6;
        const int u = i <= j ? triu(i, j) : triu(j, i);
        H[lane] = v[u];
        Hg[lane] = v[32 + u];
        double bj, bji;
        basis_entry(Ts, Lr, lane, bj, bji);
        J[lane] = bj;
        Ji[lane] = bji;
    }
    wave_sync();
    double mean = 0.0;
    for (int i = 0; i < 6; ++i) mean += H[7 * i];
    mean /= 6.0;
    if (lane < 36) {
        double a = H[lane];
        if (lane % 7 == 0) a += 1e-9 * mean + P.damping * H[lane];
        if (prior) M1[lane] = tmul_entry(Ji, Ji, lane) * P.prior_info;  // P = prior_info J^-T J^-1
        if (degen) M2[lane] = mul_entry(Hg, J, lane);
        A[lane] = a;
    }
    if (lane < 6) rhs[lane] = -v[21 + lane];
    wave_sync();
    if (prior || degen) {
        if (lane < 36) {
            if (prior) A[lane] += M1[lane];
            if (degen) M3[lane] = tmul_entry(J, M2, lane);  // J^T Hg J
        }
        if (prior && lane < 6) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += M1[6 * lane + k] * xi[k];
            rhs[lane] += s;
        }
        wave_sync();
    }
    if (stamp) t_setup = wall_clock64();
    // the damped system and the degeneracy test (J^T Hg J - thr I positive definite), factored together
    bool pd_a, pd_g;
    wave_cholesky2(A, 0.0, L, degen ? M3 : A, degen ? P.degen_thr : 0.0, L2, pd_a, pd_g);
    if (stamp) t_factor = wall_clock64();
    if (degen && !pd_g) return 0;  // a direction to hold: the scalar routine does it
    if (!pd_a) return 0;           // indefinite: the scalar routine's LU fallback
    double lrow[6];
    const int r6 = lane < 6 ? lane : 0;
    LW_UNROLL for (int m = 0; m < 6; ++m) lrow[m] = m <= r6 ? L[6 * r6 + m] : 0.0;
    wave_chol_solve(lrow, rhs[r6], o.delta);
    if (stamp) t_solve = wall_clock64();
    if (lane < 36) H_out[lane] = H[lane];
    return 1;
}

}  // namespace gnw
}  // namespace lw
