// Float64 Gauss-Newton step of the scan-to-map registration, as one sequential routine that runs in one
// lane on the device (at the end of an accumulation pass) and, compiled for the host, in tests.
//
// It restates the host step of python/livox_warp/odom.py:387-455 and the loop body of :768-796:
// Levenberg-Marquardt damping, a prior toward the motion prediction in sensor coordinates, the
// degenerate-direction hold (eigen-decomposition of the surface-normal information in sensor
// coordinates) and the left-multiplied twist update. The reference calls numpy/scipy for the dense
// algebra; here the symmetric eigen problems are cyclic Jacobi sweeps, the SPD systems are Cholesky
// solves with an LU (partial pivoting) fallback so that, as with numpy.linalg.solve, only an exactly
// singular system stops the iteration, and the rotation maps are closed forms.
//
// Every array has a fixed size and every loop a fixed trip count (fully unrolled), so on the device all
// of it lives in registers: no private (scratch) memory. The reduced solve of the constrained step (k of
// 6 eigen-directions) is the 6x6 system with identity rows for the held directions, which eliminates to
// the same numbers as the k x k system.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define LW_HD __host__ __device__
#define LW_UNROLL _Pragma("unroll")
#else
#define LW_HD
#define LW_UNROLL
#endif

namespace lw {
namespace gn {

LW_HD inline void mat3_mul(const double *A, const double *B, double *C) {
    LW_UNROLL for (int i = 0; i < 3; ++i)
        LW_UNROLL for (int j = 0; j < 3; ++j)
            C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

LW_HD inline void mat6_mul(const double *A, const double *B, double *C) {
    LW_UNROLL for (int i = 0; i < 6; ++i)
        LW_UNROLL for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            LW_UNROLL for (int k = 0; k < 6; ++k) s += A[6 * i + k] * B[6 * k + j];
            C[6 * i + j] = s;
        }
}

// C = A^T B
LW_HD inline void mat6_tmul(const double *A, const double *B, double *C) {
    LW_UNROLL for (int i = 0; i < 6; ++i)
        LW_UNROLL for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            LW_UNROLL for (int k = 0; k < 6; ++k) s += A[6 * k + i] * B[6 * k + j];
            C[6 * i + j] = s;
        }
}

LW_HD inline void mat6_vec(const double *A, const double *x, double *y) {
    LW_UNROLL for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        LW_UNROLL for (int k = 0; k < 6; ++k) s += A[6 * i + k] * x[k];
        y[i] = s;
    }
}

LW_HD inline void mat6_tvec(const double *A, const double *x, double *y) {
    LW_UNROLL for (int i = 0; i < 6; ++i) {
        double s = 0.0;
        LW_UNROLL for (int k = 0; k < 6; ++k) s += A[6 * k + i] * x[k];
        y[i] = s;
    }
}

// Rotation matrix of a rotation vector (Rodrigues; series near zero).
LW_HD inline void rot_from_rotvec(const double *v, double *R) {
    double th2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    double th = sqrt(th2);
    double a, b;  // R = I + a [v]x + b [v]x^2
    if (th < 1e-4) {
        a = 1.0 - th2 / 6.0;
        b = 0.5 - th2 / 24.0;
    } else {
        a = sin(th) / th;
        b = (1.0 - cos(th)) / th2;
    }
    double K[9] = {0.0, -v[2], v[1], v[2], 0.0, -v[0], -v[1], v[0], 0.0};
    double K2[9];
    mat3_mul(K, K, K2);
    LW_UNROLL for (int i = 0; i < 9; ++i) R[i] = a * K[i] + b * K2[i];
    R[0] += 1.0;
    R[4] += 1.0;
    R[8] += 1.0;
}

// Rotation vector of a rotation matrix: unit quaternion by the largest-pivot method, then 2 atan2(|v|, w).
LW_HD inline void rotvec_from_rot(const double *R, double *out) {
    const double tr = R[0] + R[4] + R[8];
    double qx, qy, qz, qw;
    if (tr >= R[0] && tr >= R[4] && tr >= R[8]) {
        qx = R[7] - R[5]; qy = R[2] - R[6]; qz = R[3] - R[1]; qw = 1.0 + tr;
    } else if (R[0] >= R[4] && R[0] >= R[8]) {
        qx = 1.0 - tr + 2.0 * R[0]; qy = R[3] + R[1]; qz = R[6] + R[2]; qw = R[7] - R[5];
    } else if (R[4] >= R[8]) {
        qx = R[1] + R[3]; qy = 1.0 - tr + 2.0 * R[4]; qz = R[7] + R[5]; qw = R[2] - R[6];
    } else {
        qx = R[2] + R[6]; qy = R[5] + R[7]; qz = 1.0 - tr + 2.0 * R[8]; qw = R[3] - R[1];
    }
    double n = sqrt(qx * qx + qy * qy + qz * qz + qw * qw);
    double sg = qw < 0.0 ? -1.0 / n : 1.0 / n;
    qx *= sg; qy *= sg; qz *= sg; qw *= sg;
    double s = sqrt(qx * qx + qy * qy + qz * qz);
    double ang = 2.0 * atan2(s, qw);
    double scale;
    if (ang <= 1e-3) {
        double a2 = ang * ang;
        scale = 2.0 + a2 / 12.0 + 7.0 * a2 * a2 / 2880.0;
    } else {
        scale = ang / sin(ang * 0.5);
    }
    out[0] = scale * qx;
    out[1] = scale * qy;
    out[2] = scale * qz;
}

// One Jacobi rotation zeroing A[p][q] (N x N, row-major), accumulated into V.
template <int N, int P, int Q>
LW_HD inline void jacobi_rot(double *A, double *V) {
    const double apq = A[N * P + Q];
    if (apq == 0.0) return;
    const double theta = (A[N * Q + Q] - A[N * P + P]) / (2.0 * apq);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
    LW_UNROLL for (int k = 0; k < N; ++k) {
        double akp = A[N * k + P], akq = A[N * k + Q];
        A[N * k + P] = c * akp - s * akq;
        A[N * k + Q] = s * akp + c * akq;
    }
    LW_UNROLL for (int k = 0; k < N; ++k) {
        double apk = A[N * P + k], aqk = A[N * Q + k];
        A[N * P + k] = c * apk - s * aqk;
        A[N * Q + k] = s * apk + c * aqk;
    }
    A[N * P + Q] = 0.0;
    A[N * Q + P] = 0.0;
    LW_UNROLL for (int k = 0; k < N; ++k) {
        double vkp = V[N * k + P], vkq = V[N * k + Q];
        V[N * k + P] = c * vkp - s * vkq;
        V[N * k + Q] = s * vkp + c * vkq;
    }
}

template <int N>
LW_HD inline void jacobi_sweep(double *A, double *V);

template <>
LW_HD inline void jacobi_sweep<3>(double *A, double *V) {
    jacobi_rot<3, 0, 1>(A, V); jacobi_rot<3, 0, 2>(A, V); jacobi_rot<3, 1, 2>(A, V);
}

template <>
LW_HD inline void jacobi_sweep<6>(double *A, double *V) {
    jacobi_rot<6, 0, 1>(A, V); jacobi_rot<6, 0, 2>(A, V); jacobi_rot<6, 0, 3>(A, V); jacobi_rot<6, 0, 4>(A, V);
    jacobi_rot<6, 0, 5>(A, V); jacobi_rot<6, 1, 2>(A, V); jacobi_rot<6, 1, 3>(A, V); jacobi_rot<6, 1, 4>(A, V);
    jacobi_rot<6, 1, 5>(A, V); jacobi_rot<6, 2, 3>(A, V); jacobi_rot<6, 2, 4>(A, V); jacobi_rot<6, 2, 5>(A, V);
    jacobi_rot<6, 3, 4>(A, V); jacobi_rot<6, 3, 5>(A, V); jacobi_rot<6, 4, 5>(A, V);
}

// Symmetric eigen-decomposition by cyclic Jacobi: lam[i] with eigenvector column i of V (unsorted).
template <int N>
LW_HD inline void sym_eig(const double *A_in, double *lam, double *V) {
    double A[N * N];
    LW_UNROLL for (int i = 0; i < N * N; ++i) A[i] = A_in[i];
    LW_UNROLL for (int i = 0; i < N; ++i)
        LW_UNROLL for (int j = 0; j < N; ++j) V[N * i + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, dia = 0.0;
        LW_UNROLL for (int i = 0; i < N; ++i) {
            dia += A[N * i + i] * A[N * i + i];
            LW_UNROLL for (int j = i + 1; j < N; ++j) off += A[N * i + j] * A[N * i + j];
        }
        if (off == 0.0 || off <= 1e-36 * dia) break;
        jacobi_sweep<N>(A, V);
    }
    LW_UNROLL for (int i = 0; i < N; ++i) lam[i] = A[N * i + i];
}

// Solve A x = b (6x6) where only the rows/columns with keep[i] take part; the others are identity rows
// with a zero right-hand side (x[i] = 0). Cholesky when positive definite, else LU with partial pivoting.
// Returns false only for an exactly singular system (numpy.linalg.solve's LinAlgError).
LW_HD inline bool solve6(const double *A_in, const double *b_in, const bool *keep, double *x) {
    double A[36], b[6];
    LW_UNROLL for (int i = 0; i < 6; ++i) {
        LW_UNROLL for (int j = 0; j < 6; ++j)
            A[6 * i + j] = (keep[i] && keep[j]) ? A_in[6 * i + j] : (i == j ? 1.0 : 0.0);
        b[i] = keep[i] ? b_in[i] : 0.0;
    }
    double L[36];
    bool pd = true;
    LW_UNROLL for (int j = 0; j < 6; ++j) {
        double s = A[6 * j + j];
        LW_UNROLL for (int k = 0; k < j; ++k) s -= L[6 * j + k] * L[6 * j + k];
        if (!(s > 0.0)) pd = false;
        double d = sqrt(s > 0.0 ? s : 1.0);
        L[6 * j + j] = d;
        LW_UNROLL for (int i = j + 1; i < 6; ++i) {
            double t = A[6 * i + j];
            LW_UNROLL for (int k = 0; k < j; ++k) t -= L[6 * i + k] * L[6 * j + k];
            L[6 * i + j] = t / d;
        }
    }
    if (pd) {
        double y[6];
        LW_UNROLL for (int i = 0; i < 6; ++i) {
            double t = b[i];
            LW_UNROLL for (int k = 0; k < i; ++k) t -= L[6 * i + k] * y[k];
            y[i] = t / L[6 * i + i];
        }
        LW_UNROLL for (int i = 5; i >= 0; --i) {
            double t = y[i];
            LW_UNROLL for (int k = i + 1; k < 6; ++k) t -= L[6 * k + i] * x[k];
            x[i] = t / L[6 * i + i];
        }
        return true;
    }
    // LU with partial pivoting; row swaps as selects so every index stays a compile-time constant
    LW_UNROLL for (int c = 0; c < 6; ++c) {
        double best = fabs(A[6 * c + c]);
        int piv = c;
        LW_UNROLL for (int i = c + 1; i < 6; ++i)
            if (fabs(A[6 * i + c]) > best) {
                best = fabs(A[6 * i + c]);
                piv = i;
            }
        if (best == 0.0) return false;
        LW_UNROLL for (int i = c + 1; i < 6; ++i) {
            const bool sw = (i == piv);
            LW_UNROLL for (int k = 0; k < 6; ++k) {
                double u = A[6 * c + k], w = A[6 * i + k];
                A[6 * c + k] = sw ? w : u;
                A[6 * i + k] = sw ? u : w;
            }
            double u = b[c], w = b[i];
            b[c] = sw ? w : u;
            b[i] = sw ? u : w;
        }
        LW_UNROLL for (int i = c + 1; i < 6; ++i) {
            double f = A[6 * i + c] / A[6 * c + c];
            LW_UNROLL for (int k = c; k < 6; ++k) A[6 * i + k] -= f * A[6 * c + k];
            b[i] -= f * b[c];
        }
    }
    LW_UNROLL for (int i = 5; i >= 0; --i) {
        double t = b[i];
        LW_UNROLL for (int k = i + 1; k < 6; ++k) t -= A[6 * i + k] * x[k];
        x[i] = t / A[6 * i + i];
    }
    return true;
}

// odom.py:390-394: the symmetric 6x6 matrix of a row-major upper triangle.
LW_HD inline void sym6(const double *upper, double *H) {
    int idx = 0;
    LW_UNROLL for (int p = 0; p < 6; ++p)
        LW_UNROLL for (int q = p; q < 6; ++q) {
            H[6 * p + q] = upper[idx];
            H[6 * q + p] = upper[idx];
            ++idx;
        }
}

// odom.py:425-433 and its inverse: J = [[R, [c]x R / L], [0, R / L]], J^-1 = [[R^T, -R^T [c]x], [0, L R^T]].
LW_HD inline void sensor_basis(const double *T, double L, double *J, double *Ji) {
    const double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
    const double c[3] = {T[3], T[7], T[11]};
    const double K[9] = {0.0, -c[2], c[1], c[2], 0.0, -c[0], -c[1], c[0], 0.0};
    double KR[9], RtK[9];
    mat3_mul(K, R, KR);
    LW_UNROLL for (int i = 0; i < 3; ++i)
        LW_UNROLL for (int j = 0; j < 3; ++j)
            RtK[3 * i + j] = R[i] * K[j] + R[3 + i] * K[3 + j] + R[6 + i] * K[6 + j];
    LW_UNROLL for (int i = 0; i < 36; ++i) J[i] = Ji[i] = 0.0;
    LW_UNROLL for (int i = 0; i < 3; ++i)
        LW_UNROLL for (int j = 0; j < 3; ++j) {
            J[6 * i + j] = R[3 * i + j];
            J[6 * i + 3 + j] = KR[3 * i + j] / L;
            J[6 * (3 + i) + 3 + j] = R[3 * i + j] / L;
            Ji[6 * i + j] = R[3 * j + i];
            Ji[6 * i + 3 + j] = -RtK[3 * i + j];
            Ji[6 * (3 + i) + 3 + j] = L * R[3 * j + i];
        }
}

// odom.py:397-401: left twist (rho, theta) with exp(twist) T_from ~= T_to.
LW_HD inline void twist_between(const double *Tf, const double *Tt, double *xi) {
    const double Rf[9] = {Tf[0], Tf[1], Tf[2], Tf[4], Tf[5], Tf[6], Tf[8], Tf[9], Tf[10]};
    const double Rt[9] = {Tt[0], Tt[1], Tt[2], Tt[4], Tt[5], Tt[6], Tt[8], Tt[9], Tt[10]};
    double R[9];
    LW_UNROLL for (int i = 0; i < 3; ++i)
        LW_UNROLL for (int j = 0; j < 3; ++j)
            R[3 * i + j] = Rt[3 * i] * Rf[3 * j] + Rt[3 * i + 1] * Rf[3 * j + 1] + Rt[3 * i + 2] * Rf[3 * j + 2];
    const double tf[3] = {Tf[3], Tf[7], Tf[11]};
    LW_UNROLL for (int i = 0; i < 3; ++i)
        xi[i] = Tt[4 * i + 3] - (R[3 * i] * tf[0] + R[3 * i + 1] * tf[1] + R[3 * i + 2] * tf[2]);
    rotvec_from_rot(R, xi + 3);
}

// odom.py:404-409: R' = Exp(d_theta) R, t' = Exp(d_theta) t + d_rho.
LW_HD inline void apply_twist(double *T, const double *d) {
    double Rz[9];
    rot_from_rotvec(d + 3, Rz);
    const double R[9] = {T[0], T[1], T[2], T[4], T[5], T[6], T[8], T[9], T[10]};
    const double t[3] = {T[3], T[7], T[11]};
    double Rn[9];
    mat3_mul(Rz, R, Rn);
    LW_UNROLL for (int i = 0; i < 3; ++i) {
        LW_UNROLL for (int j = 0; j < 3; ++j) T[4 * i + j] = Rn[3 * i + j];
        T[4 * i + 3] = Rz[3 * i] * t[0] + Rz[3 * i + 1] * t[1] + Rz[3 * i + 2] * t[2] + d[i];
    }
}

struct StepParams {
    double damping;
    double prior_info;
    double degen_thr;
    int32_t min_corr;
    int32_t pad;
};

struct StepOut {
    double delta[6];
    double rms;
    double H[36];
    int32_t corr;
    int32_t n_degen;
};

enum { STEP_OK = 0, STEP_FEW = 1, STEP_SINGULAR = 2 };

// True when A - shift I is positive definite (a Cholesky factorisation succeeds): every eigenvalue of the
// symmetric A exceeds shift.
LW_HD inline bool pd_above(const double *A, double shift) {
    double L[36];
    bool pd = true;
    LW_UNROLL for (int j = 0; j < 6; ++j) {
        double s = A[6 * j + j] - shift;
        LW_UNROLL for (int k = 0; k < j; ++k) s -= L[6 * j + k] * L[6 * j + k];
        if (!(s > 0.0)) pd = false;
        double d = sqrt(s > 0.0 ? s : 1.0);
        L[6 * j + j] = d;
        LW_UNROLL for (int i = j + 1; i < 6; ++i) {
            double t = A[6 * i + j];
            LW_UNROLL for (int k = 0; k < j; ++k) t -= L[6 * i + k] * L[6 * j + k];
            L[6 * i + j] = t / d;
        }
    }
    return pd;
}

// odom.py:436-455
LW_HD inline bool constrained_step(const double *A, const double *rhs, const double *Hg, const double *T, double L,
                                   double thr, const double *to_pred, double *delta, int32_t *n_degen) {
    double J[36], Ji[36], tmp[36], JHJ[36];
    sensor_basis(T, L, J, Ji);
    mat6_mul(Hg, J, tmp);
    mat6_tmul(J, tmp, JHJ);
    if (pd_above(JHJ, thr)) {  // every direction constrained: the plain solve (odom.py:447-448), no eigen sweeps
        *n_degen = 0;
        const bool all[6] = {true, true, true, true, true, true};
        return solve6(A, rhs, all, delta);
    }
    double lam[6], V[36];
    sym_eig<6>(JHJ, lam, V);
    bool good[6];
    int ng = 0;
    LW_UNROLL for (int i = 0; i < 6; ++i) {
        good[i] = lam[i] >= thr;
        ng += good[i] ? 1 : 0;
    }
    *n_degen = 6 - ng;
    if (ng == 6) {
        const bool all[6] = {true, true, true, true, true, true};
        return solve6(A, rhs, all, delta);
    }
    // d0 = Vd Vd^T (J^-1 to_pred)
    double s[6], d0[6] = {0, 0, 0, 0, 0, 0};
    mat6_vec(Ji, to_pred, s);
    LW_UNROLL for (int c = 0; c < 6; ++c) {
        double proj = 0.0;
        LW_UNROLL for (int k = 0; k < 6; ++k) proj += V[6 * k + c] * s[k];
        if (good[c]) proj = 0.0;
        LW_UNROLL for (int k = 0; k < 6; ++k) d0[k] += V[6 * k + c] * proj;
    }
    if (ng > 0) {
        double As[36];
        mat6_mul(A, J, tmp);
        mat6_tmul(J, tmp, As);
        double Jr[6], Ad[6], g[6];
        mat6_tvec(J, rhs, Jr);
        mat6_vec(As, d0, Ad);
        LW_UNROLL for (int i = 0; i < 6; ++i) g[i] = Jr[i] - Ad[i];
        // S = V^T As V, y = V^T g over the kept directions
        double AV[36], S[36], y[6], z[6];
        mat6_mul(As, V, AV);
        mat6_tmul(V, AV, S);
        mat6_tvec(V, g, y);
        if (!solve6(S, y, good, z)) return false;
        LW_UNROLL for (int a = 0; a < 6; ++a) {
            double za = good[a] ? z[a] : 0.0;
            LW_UNROLL for (int i = 0; i < 6; ++i) d0[i] += V[6 * i + a] * za;
        }
    }
    mat6_vec(J, d0, delta);
    return true;
}

// One iteration's host step (odom.py:769-793) from the reduced accumulator v (ACC layout).
LW_HD inline int step(const double *v, const double *T, const double *T_pred, const StepParams &P, StepOut &o) {
    o.corr = (int32_t)v[27];
    o.n_degen = 0;
    if (o.corr < P.min_corr) return STEP_FEW;
    double H[36], Hg[36], b[6];
    sym6(v + 0, H);
    sym6(v + 32, Hg);
    LW_UNROLL for (int i = 0; i < 6; ++i) b[i] = v[21 + i];
    double L = v[31] / (double)o.corr;
    if (!(L > 1.0)) L = 1.0;
    const double rr = v[29] > 0.0 ? v[29] : 0.0;
    const double ww = v[30] > 1e-9 ? v[30] : 1e-9;
    o.rms = sqrt(rr / ww);
    double mean = 0.0;
    LW_UNROLL for (int i = 0; i < 6; ++i) mean += H[7 * i];
    mean /= 6.0;
    double A[36], rhs[6];
    LW_UNROLL for (int i = 0; i < 36; ++i) A[i] = H[i];
    LW_UNROLL for (int i = 0; i < 6; ++i) A[7 * i] += 1e-9 * mean + P.damping * H[7 * i];
    LW_UNROLL for (int i = 0; i < 6; ++i) rhs[i] = -b[i];
    double xi[6];
    twist_between(T, T_pred, xi);
    if (P.prior_info > 0.0) {
        double J[36], Ji[36], JtJ[36], Pxi[6];
        sensor_basis(T, L, J, Ji);
        mat6_tmul(Ji, Ji, JtJ);
        LW_UNROLL for (int i = 0; i < 36; ++i) JtJ[i] *= P.prior_info;
        LW_UNROLL for (int i = 0; i < 36; ++i) A[i] += JtJ[i];
        mat6_vec(JtJ, xi, Pxi);
        LW_UNROLL for (int i = 0; i < 6; ++i) rhs[i] += Pxi[i];
    }
    LW_UNROLL for (int i = 0; i < 36; ++i) o.H[i] = H[i];
    bool ok;
    if (P.degen_thr > 0.0) {
        ok = constrained_step(A, rhs, Hg, T, L, P.degen_thr, xi, o.delta, &o.n_degen);
    } else {
        const bool all[6] = {true, true, true, true, true, true};
        ok = solve6(A, rhs, all, o.delta);
    }
    return ok ? STEP_OK : STEP_SINGULAR;
}

// odom.py:797-800: min/max eigenvalue of H normalised by its diagonal.
LW_HD inline double cond_of(const double *H) {
    double dg[6], Hn[36], lam[6], V[36];
    LW_UNROLL for (int i = 0; i < 6; ++i) dg[i] = sqrt(H[7 * i] > 1e-12 ? H[7 * i] : 1e-12);
    LW_UNROLL for (int i = 0; i < 6; ++i)
        LW_UNROLL for (int j = 0; j < 6; ++j) Hn[6 * i + j] = H[6 * i + j] / (dg[i] * dg[j]);
    sym_eig<6>(Hn, lam, V);
    double lo = lam[0], hi = lam[0];
    LW_UNROLL for (int i = 1; i < 6; ++i) {
        lo = lam[i] < lo ? lam[i] : lo;
        hi = lam[i] > hi ? lam[i] : hi;
    }
    lo = lo > 0.0 ? lo : 0.0;
    hi = hi > 1e-12 ? hi : 1e-12;
    return lo / hi;
}

}  // namespace gn
}  // namespace lw
