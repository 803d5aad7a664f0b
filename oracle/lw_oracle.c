/*
 * CPU restatement of livox-warp's scan-to-map odometry kernels -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: tests, the smoke check and the CPU baseline leg of the benchmark call
 * it as the checker. The product (the HIP library) never links or calls it.
 *
 * Semantics follow Warp's CPU backend, which runs a launch as a serial loop over thread ids on one
 * host thread: every "atomic" below is a plain read-modify-write in index order. Device math that
 * the reference does in float32 is done in float32 here (compile with -ffp-contract=off so no
 * multiply-add is fused); the accumulators are float64 as in the reference.
 *
 * Reference files (paths relative to the reference repository root):
 *   python/livox_warp/gpu.py      :59-64 key constants, :128-139 attr fields, :231-265 keys, hash,
 *                                  probe, :268-284 deskew
 *   python/livox_warp/odom.py     :72-94 constants, :97-123 k_frame_append, :126-160 k_ds_insert,
 *                                  :163-178 k_deskew, :181-230 k_reg_insert, :233-272 accumulate,
 *                                  :275-375 k_icp, :378-384 k_reduce
 *   python/livox_warp/gpu.py      :55-73 map constants, :156-167 keep_raw, :373-413 k_map_insert,
 *                                  :415-467 map_centroid and local_plane, :470-595 k_carve
 *
 * Where the reference's bits come from Warp internals that are not in the reference repository,
 * this file makes a stated choice and the result is "parity unpinned" at the bit level (pinned by
 * tolerance tests instead):
 *   - wp.transform_point: row r is ((m[r][0]*p0 + m[r][1]*p1) + m[r][2]*p2) + m[r][3], no FMA;
 *   - wp.eig3 (odom.py:342): replaced by a float64 cyclic Jacobi on the float32 covariance; the
 *     per-direction weights M, Mg and the tightest direction are formed in float64 and rounded to
 *     float32 (M and Mg do not depend on the eigenbasis chosen, only on the eigenvalues);
 *   - wp.quat_from_axis_angle / wp.quat_rotate (gpu.py:275): the textbook unit-quaternion rotation
 *     v' = v (2w^2 - 1) + 2 (q.v) q + 2 w (q x v) with sinf/cosf;
 *   - 1 << ret for a return index >= 31 (undefined for a 32-bit int): such a point is rejected;
 *   - wp.eig3 in local_plane (gpu.py:453): a float64 cyclic Jacobi on the float32 covariance
 *     (lwo_sym_eig3, the same sweeps and stopping test as the device's), so the surface test and the
 *     normal are float64 where the reference's are float32.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

/* Device-side arithmetic type: float32 (the reference's kernels), or float64 with -DLWO_FP64 (the
 * all-float64 oracle of the precision study, built as liblw_oracle64.so from this same file). */
#ifdef LWO_FP64
typedef double real;
#define LWO_FLOOR floor
#define LWO_SQRT sqrt
#define LWO_SIN sin
#define LWO_COS cos
#else
typedef float real;
#define LWO_FLOOR floorf
#define LWO_SQRT sqrtf
#define LWO_SIN sinf
#define LWO_COS cosf
#endif

#define LWO_KEY_BITS 21
#define LWO_KEY_BIAS (1ULL << 20)
#define LWO_KEY_MASK ((1ULL << 21) - 1ULL)
#define LWO_HASH_MUL 0x9E3779B97F4A7C15ULL /* gpu.py:64 stores it as int64 -7046029254386353131 */
#define LWO_PROBES 64                      /* gpu.py:257, odom.py:153, odom.py:219 */
#define LWO_REG_CAP 4096                   /* odom.py:83 */
#define LWO_ACC_SLOTS 64                   /* odom.py:72 */
#define LWO_ACC_STRIDE 64                  /* odom.py:82 */

enum { ACC_H = 0, ACC_B = 21, ACC_COUNT = 27, ACC_ERR = 28, ACC_RMS = 29, ACC_WEIGHT = 30, ACC_RANGE = 31,
       ACC_HG = 32 }; /* odom.py:73-81 */

/* ---- keys, hash, probe (gpu.py:231-265) ------------------------------------------------------ */

int64_t lwo_voxel_key_ijk(int32_t ix, int32_t iy, int32_t iz)
{
    uint64_t x = ((uint64_t)(int64_t)ix + LWO_KEY_BIAS) & LWO_KEY_MASK;
    uint64_t y = ((uint64_t)(int64_t)iy + LWO_KEY_BIAS) & LWO_KEY_MASK;
    uint64_t z = ((uint64_t)(int64_t)iz + LWO_KEY_BIAS) & LWO_KEY_MASK;
    return (int64_t)(x | (y << LWO_KEY_BITS) | (z << (2 * LWO_KEY_BITS)));
}

int32_t lwo_slot_hash(int64_t k)
{
    uint64_t h = (uint64_t)k * LWO_HASH_MUL; /* wrapping, as int64 multiply wraps in Warp */
    h ^= h >> 29;                            /* arithmetic vs logical shift: same low 31 bits */
    return (int32_t)(h & 0x7FFFFFFFULL);
}

int32_t lwo_find_slot(int64_t key, const int64_t *keys, int32_t mask)
{
    int32_t h = lwo_slot_hash(key) & mask;
    for (int probe = 0; probe < LWO_PROBES; ++probe) {
        int32_t s = (h + probe) & mask;
        int64_t k = keys[s];
        if (k == key) return s;
        if (k == -1) return -1;
    }
    return -1;
}

static inline int32_t ifloor(real x) { return (int32_t)LWO_FLOOR(x); }

int64_t lwo_voxel_key(const real *w, real inv)
{
    return lwo_voxel_key_ijk(ifloor(w[0] * inv), ifloor(w[1] * inv), ifloor(w[2] * inv));
}

static inline int tag_of(uint32_t a) { return (int)((a >> 8) & 255u); }
static inline int ret_of(uint32_t a) { return (int)((a >> 16) & 255u); }
static inline int ret_ok(uint32_t a, int ret_mask)
{
    int r = ret_of(a);
    return r < 31 && ((1 << r) & ret_mask) != 0;
}

static inline real length3(const real *p) { return LWO_SQRT(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

void lwo_transform_point(const real *T, const real *p, real *w)
{
    for (int r = 0; r < 3; ++r) {
        real a = T[r * 4 + 0] * p[0];
        a = a + T[r * 4 + 1] * p[1];
        a = a + T[r * 4 + 2] * p[2];
        w[r] = a + T[r * 4 + 3];
    }
}

/* ---- frame cutting (odom.py:97-123) ---------------------------------------------------------- */

void lwo_frame_append(const real *sx, const uint32_t *sa, const real *st, int n, real inv_dt, int k_min,
                      int k_open, int cap, int32_t *count, real *f_xyz, uint32_t *f_attr, real *f_t,
                      int32_t *f_frame)
{
    for (int i = 0; i < n; ++i) {
        real tt = st[i] > ((real)0.0) ? st[i] : ((real)0.0);
        int k = ifloor(tt * inv_dt);
        if (k > k_open || k < k_min) continue;
        int j = (*count)++;
        if (j >= cap) continue;
        memcpy(f_xyz + 3 * (size_t)j, sx + 3 * (size_t)i, 3 * sizeof(real));
        f_attr[j] = sa[i];
        f_t[j] = st[i];
        f_frame[j] = k_open;
    }
}

/* ---- deskew (odom.py:163-178, gpu.py:268-284) ------------------------------------------------ */

static void rotvec_rotate(const real *r, const real *p, real *out)
{
    real a = length3(r);
    if (a < ((real)1.0e-9)) {
        out[0] = p[0]; out[1] = p[1]; out[2] = p[2];
        return;
    }
    real ax[3] = {r[0] / a, r[1] / a, r[2] / a};
    real sh = LWO_SIN(a * ((real)0.5)), ch = LWO_COS(a * ((real)0.5));
    real q[3] = {ax[0] * sh, ax[1] * sh, ax[2] * sh};
    real w = ch;
    real c = ((real)2.0) * w * w - ((real)1.0);
    real d = ((real)2.0) * (q[0] * p[0] + q[1] * p[1] + q[2] * p[2]);
    real cx = q[1] * p[2] - q[2] * p[1];
    real cy = q[2] * p[0] - q[0] * p[2];
    real cz = q[0] * p[1] - q[1] * p[0];
    real w2 = ((real)2.0) * w;
    out[0] = p[0] * c + q[0] * d + cx * w2;
    out[1] = p[1] * c + q[1] * d + cy * w2;
    out[2] = p[2] * c + q[2] * d + cz * w2;
}

void lwo_deskew(const real *raw, const real *t, int m, real t_mid, real inv_dt, const real *rho,
                const real *theta, real *out)
{
    for (int i = 0; i < m; ++i) {
        real s = (t[i] - t_mid) * inv_dt;
        s = s < -((real)1.0) ? -((real)1.0) : (s > ((real)1.0) ? ((real)1.0) : s);
        real r[3] = {theta[0] * s, theta[1] * s, theta[2] * s};
        real q[3];
        rotvec_rotate(r, raw + 3 * (size_t)i, q);
        out[3 * (size_t)i + 0] = q[0] + rho[0] * s;
        out[3 * (size_t)i + 1] = q[1] + rho[1] * s;
        out[3 * (size_t)i + 2] = q[2] + rho[2] * s;
    }
}

/* ---- thinning (odom.py:126-160) -------------------------------------------------------------- */

void lwo_ds_insert(const real *xyz, const uint32_t *attr, int n, real inv_v, real min_range, real max_range,
                   int ret_mask, int mask, int64_t *keys, int32_t *count, real *out)
{
    for (int i = 0; i < n; ++i) {
        const real *p = xyz + 3 * (size_t)i;
        real r = length3(p);
        if (r < min_range || r > max_range) continue;
        if (tag_of(attr[i]) != 0 || !ret_ok(attr[i], ret_mask)) continue;
        int64_t key = lwo_voxel_key(p, inv_v);
        int32_t h = lwo_slot_hash(key) & mask;
        for (int probe = 0; probe < LWO_PROBES; ++probe) {
            int32_t s = (h + probe) & mask;
            if (keys[s] == -1) {
                keys[s] = key;
                memcpy(out + 3 * (size_t)(*count), p, 3 * sizeof(real));
                (*count)++;
                break;
            }
            if (keys[s] == key) break;
        }
    }
}

/* ---- map build (odom.py:181-230) ------------------------------------------------------------- */

static inline real center_of(int32_t i, real voxel) { return ((real)i + ((real)0.5)) * voxel; }

void lwo_reg_insert(const real *xyz, const uint32_t *attr, int n, const real *T, real inv_v, real voxel,
                    real max_range, int ret_mask, int mask, int64_t *keys, real *rsum, real *router,
                    int32_t *rcnt, int32_t *stats)
{
    for (int i = 0; i < n; ++i) {
        const real *p = xyz + 3 * (size_t)i;
        if (length3(p) > max_range || tag_of(attr[i]) != 0 || !ret_ok(attr[i], ret_mask)) continue;
        real w[3];
        lwo_transform_point(T, p, w);
        int32_t ix = ifloor(w[0] * inv_v), iy = ifloor(w[1] * inv_v), iz = ifloor(w[2] * inv_v);
        int64_t key = lwo_voxel_key_ijk(ix, iy, iz);
        real r[3] = {w[0] - center_of(ix, voxel), w[1] - center_of(iy, voxel), w[2] - center_of(iz, voxel)};
        int32_t h = lwo_slot_hash(key) & mask;
        int done = 0;
        for (int probe = 0; probe < LWO_PROBES && !done; ++probe) {
            int32_t s = (h + probe) & mask;
            int64_t prev = keys[s];
            if (prev == -1 || prev == key) {
                if (prev == -1) {
                    keys[s] = key;
                    stats[0]++;
                }
                if (rcnt[s] < LWO_REG_CAP) {
                    for (int a = 0; a < 3; ++a) rsum[3 * (size_t)s + a] += r[a];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) router[9 * (size_t)s + 3 * a + b] += r[a] * r[b];
                    rcnt[s] += 1;
                }
                done = 1;
            }
        }
        if (!done) stats[1]++;
    }
}

/* ---- symmetric 3x3 eigen: float64 cyclic Jacobi (stands in for wp.eig3) -------------------- */

void lwo_eig3(const double *A_in, double *ev, double *Q)
{
    double A[3][3], V[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[i][j] = A_in[3 * i + j];
            V[i][j] = (i == j) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1e-40 * dia || off == 0.0) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) { /* A <- A J */
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; ++k) { /* A <- J^T A */
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                A[p][q] = A[q][p] = 0.0;
                for (int k = 0; k < 3; ++k) { /* V <- V J */
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    for (int a = 0; a < 3; ++a) {
        ev[a] = A[a][a];
        for (int k = 0; k < 3; ++k) Q[3 * k + a] = V[k][a]; /* column a = eigenvector a */
    }
}

/*
 * The per-voxel precision of k_icp (odom.py:336-368) from the voxel's moments: M (6 unique entries,
 * xx xy xz yy yz zz), its tight part Mg, the tightest direction q and whether the voxel is a shape
 * (1: rn = r.q) or the weak isotropic branch (0: rn = |r|). Returns the shape flag.
 */
int lwo_voxel_precision(const real *rsum, const real *router, int32_t c, real voxel, real noise2, real tight2,
                        real alpha, real *M6, real *Mg6, real *q3)
{
    real cf = (real)c;
    real mrel[3] = {rsum[0] / cf, rsum[1] / cf, rsum[2] / cf};
    double cov[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) cov[3 * i + j] = (double)(router[3 * i + j] / cf - mrel[i] * mrel[j]);
    double ev[3], Q[9];
    lwo_eig3(cov, ev, Q);
    int nt = 0;
    for (int a = 0; a < 3; ++a) nt += ev[a] < (double)tight2;
    if (c >= 5 && nt <= 1) {
        int kmin = 0;
        if (ev[1] < ev[kmin]) kmin = 1;
        if (ev[2] < ev[kmin]) kmin = 2;
        double M[9] = {0}, Mg[9] = {0};
        for (int a = 0; a < 3; ++a) {
            double q[3] = {Q[0 * 3 + a], Q[1 * 3 + a], Q[2 * 3 + a]};
            double e_a = (ev[a] > 0.0 ? ev[a] : 0.0) + (double)noise2;
            double wa = 1.0 / e_a;
            int tight = !(ev[a] >= (double)tight2);
            if (!tight) wa = (double)alpha / e_a;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    M[3 * i + j] += wa * q[i] * q[j];
                    if (tight) Mg[3 * i + j] += wa * q[i] * q[j];
                }
            if (a == kmin)
                for (int i = 0; i < 3; ++i) q3[i] = (real)q[i];
        }
        const int iu[6] = {0, 1, 2, 4, 5, 8};
        for (int k = 0; k < 6; ++k) {
            M6[k] = (real)M[iu[k]];
            Mg6[k] = (real)Mg[iu[k]];
        }
        return 1;
    }
    real wa = alpha / (((real)0.25) * voxel * voxel + noise2);
    M6[0] = wa; M6[1] = ((real)0.0); M6[2] = ((real)0.0); M6[3] = wa; M6[4] = ((real)0.0); M6[5] = wa;
    for (int k = 0; k < 6; ++k) Mg6[k] = ((real)0.0);
    q3[0] = q3[1] = q3[2] = ((real)0.0);
    return 0;
}

/* ---- accumulation (odom.py:233-272) ---------------------------------------------------------- */

static void accumulate_h(double *acc, const real J[3][6], const real M[3][3], real wgt)
{
    int idx = 0;
    for (int p = 0; p < 6; ++p)
        for (int q = p; q < 6; ++q) {
            /* odom.py:250-252: v = row 0; v += row 1; v += row 2, each row summed left to right */
            real v = ((real)0.0);
            for (int i = 0; i < 3; ++i) {
                real row = M[i][0] * J[i][p] * J[0][q];
                row = row + M[i][1] * J[i][p] * J[1][q];
                row = row + M[i][2] * J[i][p] * J[2][q];
                v = (i == 0) ? row : v + row;
            }
            acc[idx++] += (double)(wgt * v);
        }
}

static void accumulate(double *acc, const real J[3][6], const real M[3][3], const real *r, real wgt)
{
    real Mr[3];
    for (int i = 0; i < 3; ++i) Mr[i] = M[i][0] * r[0] + M[i][1] * r[1] + M[i][2] * r[2];
    accumulate_h(acc + ACC_H, J, M, wgt);
    for (int p = 0; p < 6; ++p) acc[ACC_B + p] += (double)(wgt * (Mr[0] * J[0][p] + Mr[1] * J[1][p] + Mr[2] * J[2][p]));
    acc[ACC_COUNT] += 1.0;
    acc[ACC_ERR] += (double)(wgt * (r[0] * Mr[0] + r[1] * Mr[1] + r[2] * Mr[2]));
    acc[ACC_WEIGHT] += (double)wgt;
}

/* One point of k_icp (odom.py:299-375). Returns the slot of the chosen voxel, or -1. */
static int32_t icp_point(const real *p, int i, const real *T, real inv_v, real voxel, int mask,
                         const int64_t *keys, const real *rsum, const real *router, const int32_t *rcnt, int reach,
                         real max_d2, real noise2, real tight2, real alpha, real kappa, double *acc)
{
    real w[3];
    lwo_transform_point(T, p, w);
    double *slot = acc + (size_t)(i & (LWO_ACC_SLOTS - 1)) * LWO_ACC_STRIDE;
    real J[3][6] = {{((real)1.0), ((real)0.0), ((real)0.0), ((real)0.0), w[2], -w[1]},
                     {((real)0.0), ((real)1.0), ((real)0.0), -w[2], ((real)0.0), w[0]},
                     {((real)0.0), ((real)0.0), ((real)1.0), w[1], -w[0], ((real)0.0)}};
    int32_t bx = ifloor(w[0] * inv_v), by = ifloor(w[1] * inv_v), bz = ifloor(w[2] * inv_v);
    real best_d2 = max_d2;
    int32_t best_s = -1, best_c = 0;
    real best_mu[3] = {0, 0, 0};
    for (int dx = -reach; dx <= reach; ++dx)
        for (int dy = -reach; dy <= reach; ++dy)
            for (int dz = -reach; dz <= reach; ++dz) {
                int32_t jx = bx + dx, jy = by + dy, jz = bz + dz;
                int32_t s = lwo_find_slot(lwo_voxel_key_ijk(jx, jy, jz), keys, mask);
                if (s < 0) continue;
                int32_t c = rcnt[s];
                if (c <= 0) continue;
                real cf = (real)c;
                real mu[3] = {center_of(jx, voxel) + rsum[3 * (size_t)s + 0] / cf,
                               center_of(jy, voxel) + rsum[3 * (size_t)s + 1] / cf,
                               center_of(jz, voxel) + rsum[3 * (size_t)s + 2] / cf};
                real dd[3] = {w[0] - mu[0], w[1] - mu[1], w[2] - mu[2]};
                real d2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
                if (d2 < best_d2) {
                    best_d2 = d2;
                    best_s = s;
                    best_c = c;
                    memcpy(best_mu, mu, 3 * sizeof(real));
                }
            }
    if (best_s < 0) return -1;
    real r[3] = {w[0] - best_mu[0], w[1] - best_mu[1], w[2] - best_mu[2]};
    real M6[6], Mg6[6], q[3];
    int shape = lwo_voxel_precision(rsum + 3 * (size_t)best_s, router + 9 * (size_t)best_s, best_c, voxel, noise2,
                                    tight2, alpha, M6, Mg6, q);
    real M[3][3] = {{M6[0], M6[1], M6[2]}, {M6[1], M6[3], M6[4]}, {M6[2], M6[4], M6[5]}};
    real Mg[3][3] = {{Mg6[0], Mg6[1], Mg6[2]}, {Mg6[1], Mg6[3], Mg6[4]}, {Mg6[2], Mg6[4], Mg6[5]}};
    real rn = shape ? (r[0] * q[0] + r[1] * q[1] + r[2] * q[2]) : length3(r);
    real Mr[3];
    for (int a = 0; a < 3; ++a) Mr[a] = M[a][0] * r[0] + M[a][1] * r[1] + M[a][2] * r[2];
    real d2 = r[0] * Mr[0] + r[1] * Mr[1] + r[2] * Mr[2];
    real wgt = ((real)1.0) / (((real)1.0) + d2 / (kappa * kappa));
    wgt = wgt * wgt;
    slot[ACC_RMS] += (double)(wgt * rn * rn);
    accumulate(slot, J, M, r, wgt);
    accumulate_h(slot + ACC_HG, J, Mg, wgt);
    slot[ACC_RANGE] += (double)length3(p);
    return best_s;
}

/* k_icp over m points (odom.py:275-375) into acc[ACC_SLOTS * ACC_STRIDE]; best_slot (optional) gets
 * each point's chosen slot or -1. */
void lwo_icp(const real *ds, int m, const real *T, real inv_v, real voxel, int mask, const int64_t *keys,
             const real *rsum, const real *router, const int32_t *rcnt, int reach, real max_d2, real noise2,
             real tight2, real alpha, real kappa, double *acc, int32_t *best_slot)
{
    for (int i = 0; i < m; ++i) {
        int32_t s = icp_point(ds + 3 * (size_t)i, i, T, inv_v, voxel, mask, keys, rsum, router, rcnt, reach, max_d2,
                              noise2, tight2, alpha, kappa, acc);
        if (best_slot) best_slot[i] = s;
    }
}

/* k_reduce (odom.py:378-384): fixed order over the 64 partial accumulators. */
void lwo_reduce(const double *acc, double *out)
{
    for (int j = 0; j < LWO_ACC_STRIDE; ++j) {
        double s = 0.0;
        for (int k = 0; k < LWO_ACC_SLOTS; ++k) s += acc[(size_t)k * LWO_ACC_STRIDE + j];
        out[j] = s;
    }
}

/* ---- the integration map (gpu.py:373-595): insert and free-space carving ---------------------- */
/*
 * Record arrays as the reference keeps them (gpu.py:1213-1219): keys (-1 empty), acc (4 floats a slot: x, y,
 * z, reflectivity sums), cnt, last, first, miss; stats[0] voxels claimed, stats[1] points without a slot.
 * poses: 65536 row-major 4x4 float32 matrices indexed by frame id & 65535.
 */
#define LWO_POSE_MASK 65535
#define LWO_MAP_COUNT_CAP 4096 /* gpu.py:55 */

static int keep_raw(const float *p, uint32_t a, float min_range, float max_range, int min_refl, int noise_mask,
                    int ret_mask)
{
    float r = sqrtf(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    if (r < min_range || r > max_range) return 0;
    if ((int)(a & 255u) < min_refl) return 0;
    if (((noise_mask >> (tag_of(a) & 3)) & 1) == 0) return 0;
    if (((ret_mask >> (ret_of(a) & 3)) & 1) == 0) return 0;
    return 1;
}

static void xform_f(const float *T, const float *p, float *w)
{
    for (int r = 0; r < 3; ++r) {
        float a = T[r * 4 + 0] * p[0];
        a = a + T[r * 4 + 1] * p[1];
        a = a + T[r * 4 + 2] * p[2];
        w[r] = a + T[r * 4 + 3];
    }
}

void lwo_map_insert(const float *xyz, const uint32_t *attr, const float *t, const int32_t *frame, int n,
                    const float *poses, float min_range, float max_range, int min_refl, int noise_mask, int ret_mask,
                    float inv_voxel, int32_t mask, int64_t *keys, float *acc, int32_t *cnt, float *last,
                    float *first, int32_t *stats)
{
    for (int i = 0; i < n; ++i) {
        const float *p = xyz + 3 * (int64_t)i;
        if (!keep_raw(p, attr[i], min_range, max_range, min_refl, noise_mask, ret_mask)) continue;
        float w[3];
        xform_f(poses + (int64_t)(frame[i] & LWO_POSE_MASK) * 16, p, w);
        int64_t key = lwo_voxel_key_ijk((int32_t)floorf(w[0] * inv_voxel), (int32_t)floorf(w[1] * inv_voxel),
                                        (int32_t)floorf(w[2] * inv_voxel));
        int32_t h = lwo_slot_hash(key) & mask;
        int done = 0;
        for (int probe = 0; probe < LWO_PROBES && !done; ++probe) {
            int32_t s = (h + probe) & mask;
            int64_t prev = keys[s];
            if (prev == -1) keys[s] = key;
            if (prev == -1 || prev == key) {
                if (prev == -1) stats[0]++;
                int32_t old = cnt[s]++;
                if (old < LWO_MAP_COUNT_CAP) {
                    acc[4 * s] += w[0];
                    acc[4 * s + 1] += w[1];
                    acc[4 * s + 2] += w[2];
                    acc[4 * s + 3] += (float)(attr[i] & 255u);
                }
                if (t[i] > last[s]) last[s] = t[i];
                if (t[i] < first[s]) first[s] = t[i];
                done = 1;
            }
        }
        if (!done) stats[1]++;
    }
}

/* Cyclic Jacobi for a symmetric 3x3 (float64): the device's sweep order, rotation and stopping test. */
static void jrot3(double *A, double *V, int P, int Q)
{
    double apq = A[3 * P + Q];
    if (apq == 0.0) return;
    double theta = (A[3 * Q + Q] - A[3 * P + P]) / (2.0 * apq);
    double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
    double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
    for (int k = 0; k < 3; ++k) {
        double akp = A[3 * k + P], akq = A[3 * k + Q];
        A[3 * k + P] = c * akp - sn * akq;
        A[3 * k + Q] = sn * akp + c * akq;
    }
    for (int k = 0; k < 3; ++k) {
        double apk = A[3 * P + k], aqk = A[3 * Q + k];
        A[3 * P + k] = c * apk - sn * aqk;
        A[3 * Q + k] = sn * apk + c * aqk;
    }
    A[3 * P + Q] = 0.0;
    A[3 * Q + P] = 0.0;
    for (int k = 0; k < 3; ++k) {
        double vkp = V[3 * k + P], vkq = V[3 * k + Q];
        V[3 * k + P] = c * vkp - sn * vkq;
        V[3 * k + Q] = sn * vkp + c * vkq;
    }
}

void lwo_sym_eig3(const double *A_in, double *ev, double *V)
{
    double A[9];
    for (int i = 0; i < 9; ++i) A[i] = A_in[i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[3 * i + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int i = 0; i < 3; ++i) {
            dia += A[3 * i + i] * A[3 * i + i];
            for (int j = i + 1; j < 3; ++j) off += A[3 * i + j] * A[3 * i + j];
        }
        if (off == 0.0 || off <= 1e-36 * dia) break;
        jrot3(A, V, 0, 1);
        jrot3(A, V, 0, 2);
        jrot3(A, V, 1, 2);
    }
    for (int i = 0; i < 3; ++i) ev[i] = A[3 * i + i];
}

static void centroid(int32_t s, const float *acc, const int32_t *cnt, float *c)
{
    float m = 1.0f / (float)(cnt[s] < LWO_MAP_COUNT_CAP ? cnt[s] : LWO_MAP_COUNT_CAP);
    c[0] = acc[4 * s] * m;
    c[1] = acc[4 * s + 1] * m;
    c[2] = acc[4 * s + 2] * m;
}

/* gpu.py:420-467: (normal x, y, z, w) of the surface through voxel (ix, iy, iz), centroids relative to o. */
void lwo_local_plane(int32_t ix, int32_t iy, int32_t iz, const float *o, const int64_t *keys, const float *acc,
                     const int32_t *cnt, int32_t mask, float *out)
{
    float cs[3] = {0, 0, 0}, cc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cn = 0.0f;
    for (int gx = -1; gx < 2; ++gx)
        for (int gy = -1; gy < 2; ++gy)
            for (int gz = -1; gz < 2; ++gz) {
                int32_t sn = lwo_find_slot(lwo_voxel_key_ijk(ix + gx, iy + gy, iz + gz), keys, mask);
                if (sn < 0 || cnt[sn] <= 0) continue;
                float c[3], q[3];
                centroid(sn, acc, cnt, c);
                for (int a = 0; a < 3; ++a) q[a] = c[a] - o[a];
                for (int a = 0; a < 3; ++a) cs[a] = cs[a] + q[a];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) cc[3 * a + b] = cc[3 * a + b] + q[a] * q[b];
                cn = cn + 1.0f;
            }
    if (cn < 4.0f) {
        out[0] = 0.0f; out[1] = 0.0f; out[2] = 1.0f; out[3] = 0.0f;
        return;
    }
    float mu[3] = {cs[0] / cn, cs[1] / cn, cs[2] / cn};
    double cov[9], ev[3], Q[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cov[3 * a + b] = (double)(cc[3 * a + b] / cn - mu[a] * mu[b]);
    lwo_sym_eig3(cov, ev, Q);
    int k = 0;
    if (ev[1] < ev[k]) k = 1;
    if (ev[2] < ev[k]) k = 2;
    int j = (k + 1) % 3, j2 = (k + 2) % 3;
    double mid = ev[j] < ev[j2] ? ev[j] : ev[j2];
    out[0] = (float)Q[k];
    out[1] = (float)Q[3 + k];
    out[2] = (float)Q[6 + k];
    out[3] = ev[k] < 0.25 * mid ? 2.0f : 1.0f;
}

void lwo_carve(const float *xyz, const uint32_t *attr, const int32_t *frames, const float *poses, int n, int every,
               float inv_voxel, int32_t mask, float max_range, int max_steps, const int64_t *keys, const float *acc,
               const int32_t *cnt, int32_t *miss)
{
    for (int64_t i = 0; i < n; i += every) {
        if (tag_of(attr[i]) != 0) continue;
        const float *T = poses + (int64_t)(frames[i] & LWO_POSE_MASK) * 16;
        float o[3] = {T[3], T[7], T[11]}, w[3], d[3];
        xform_f(T, xyz + 3 * i, w);
        for (int a = 0; a < 3; ++a) d[a] = w[a] - o[a];
        float L = sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (L < 0.05f || L > max_range) continue;
        for (int a = 0; a < 3; ++a) d[a] = d[a] / L;
        float voxel = 1.0f / inv_voxel;
        float margin = fmaxf(3.0f * voxel, 0.03f * L);
        float pe[4];
        lwo_local_plane((int32_t)floorf(w[0] * inv_voxel), (int32_t)floorf(w[1] * inv_voxel),
                        (int32_t)floorf(w[2] * inv_voxel), w, keys, acc, cnt, mask, pe);
        if (pe[3] > 0.5f) {
            float cosine = fabsf(pe[0] * d[0] + pe[1] * d[1] + pe[2] * d[2]);
            margin = fmaxf(margin, 1.5f * voxel / fmaxf(cosine, 0.05f));
        }
        float stop = (L - margin) * inv_voxel;
        if (stop <= 1.0f) continue;
        float pos[3] = {o[0] * inv_voxel, o[1] * inv_voxel, o[2] * inv_voxel};
        int32_t ix[3], st[3] = {1, 1, 1};
        float tm[3] = {1.0e30f, 1.0e30f, 1.0e30f}, td[3] = {1.0e30f, 1.0e30f, 1.0e30f};
        for (int a = 0; a < 3; ++a) {
            ix[a] = (int32_t)floorf(pos[a]);
            if (d[a] > 1.0e-9f) {
                tm[a] = ((float)(ix[a] + 1) - pos[a]) / d[a];
                td[a] = 1.0f / d[a];
            } else if (d[a] < -1.0e-9f) {
                st[a] = -1;
                tm[a] = (pos[a] - (float)ix[a]) / (-d[a]);
                td[a] = -1.0f / d[a];
            }
        }
        for (int step = 0; step < max_steps; ++step) {
            int a = (tm[0] < tm[1] && tm[0] < tm[2]) ? 0 : (tm[1] < tm[2] ? 1 : 2);
            float tt = tm[a];
            ix[a] += st[a];
            tm[a] = tm[a] + td[a];
            if (tt >= stop) break;
            int32_t s = lwo_find_slot(lwo_voxel_key_ijk(ix[0], ix[1], ix[2]), keys, mask);
            if (s < 0 || cnt[s] <= 0) continue;
            float cen[3], rel[3], pl[4];
            centroid(s, acc, cnt, cen);
            for (int b = 0; b < 3; ++b) rel[b] = cen[b] - o[b];
            int pierced = 0;
            lwo_local_plane(ix[0], ix[1], ix[2], cen, keys, acc, cnt, mask, pl);
            if (pl[3] > 1.5f) {
                float dn = d[0] * pl[0] + d[1] * pl[1] + d[2] * pl[2];
                if (fabsf(dn) > 1.0e-4f) {
                    float f = (rel[0] * pl[0] + rel[1] * pl[1] + rel[2] * pl[2]) / dn, x[3];
                    for (int b = 0; b < 3; ++b) x[b] = d[b] * f - rel[b];
                    if (sqrtf(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]) < 0.6f * voxel) pierced = 1;
                }
            } else {
                float rd = rel[0] * d[0] + rel[1] * d[1] + rel[2] * d[2], q[3];
                for (int b = 0; b < 3; ++b) q[b] = rel[b] - rd * d[b];
                if (sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]) < 0.5f * voxel) pierced = 1;
            }
            if (pierced) miss[s] += every;
        }
    }
}
