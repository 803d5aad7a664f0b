// The decisions Odometry._finalize / _estimate take between device steps (odom.py:644-692), taken on the
// device so that a frame's whole chain -- deskew, registration, the re-deskew test and second registration,
// the keyframe test and insert -- is queued at once and read back once (lw_odom_frame in lw_abi.hip). Each
// decision is one lane of a tiny launch that reads the registration state and writes a FrameCtl flag the
// next kernels are gated on.
#include "lw_gn.h"
#include "lw_internal.h"

namespace lw {

namespace {

// 4x4 float64 product, row-major, each entry summed over k in order.
__device__ inline void mat4_mul(const double *A, const double *B, double *C) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += A[4 * i + k] * B[4 * k + j];
            C[4 * i + j] = s;
        }
}

// Rotation vector of the rotation block of a 4x4, by scipy's Rotation.from_matrix / as_rotvec method (the
// unit quaternion from the largest of R00, R11, R22 and the trace, normalised by division, w >= 0, then
// 2 atan2(|v|, w) with the series below 1e-3), as odom.py's host helper computes it.
__device__ inline void rotvec_of(const double *T, double *out) {
    const double r00 = T[0], r01 = T[1], r02 = T[2], r10 = T[4], r11 = T[5], r12 = T[6], r20 = T[8], r21 = T[9],
                 r22 = T[10];
    const double tr = r00 + r11 + r22;
    const double top = fmax(fmax(r00, r11), fmax(r22, tr));
    double x, y, z, w;
    if (r00 == top) {
        x = 1.0 - tr + 2.0 * r00; y = r10 + r01; z = r20 + r02; w = r21 - r12;
    } else if (r11 == top) {
        x = r01 + r10; y = 1.0 - tr + 2.0 * r11; z = r21 + r12; w = r02 - r20;
    } else if (r22 == top) {
        x = r02 + r20; y = r12 + r21; z = 1.0 - tr + 2.0 * r22; w = r10 - r01;
    } else {
        x = r21 - r12; y = r02 - r20; z = r10 - r01; w = 1.0 + tr;
    }
    double n = sqrt(x * x + y * y + z * z + w * w);
    if (w < 0.0) n = -n;
    x = x / n; y = y / n; z = z / n; w = w / n;
    const double ang = 2.0 * atan2(sqrt(x * x + y * y + z * z), w);
    double scale;
    if (ang <= 1e-3) {
        const double a2 = ang * ang;
        scale = 2.0 + a2 / 12.0 + 7.0 * a2 * a2 / 2880.0;
    } else {
        scale = ang / sin(ang / 2.0);
    }
    out[0] = scale * x;
    out[1] = scale * y;
    out[2] = scale * z;
}

// The registration's pose and trust (run_registration / Odometry._register): a thinned set below min_corr
// keeps the prediction and is not trusted (odom.py:756-757); otherwise ok when corr >= min_corr.
__device__ inline bool reg_outcome(const RegState *st, const double *T_pred, int32_t min_corr, double *T) {
    const bool has = st->n >= min_corr;
    for (int k = 0; k < 16; ++k) T[k] = has ? st->T[k] : T_pred[k];
    return has && st->corr >= min_corr;
}

// After the first registration: keep a copy of its state, then (one lane) the frame's count and skip test
// (odom.py:676-678) and the re-deskew test (odom.py:683-690) on xi_new = body twist of T_prev^-1 T.
__global__ void k_frame_check(const RegState *st, RegState *st1, FrameCtl *fc, const int32_t *ctr) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(st);
    uint32_t *dst = reinterpret_cast<uint32_t *>(st1);
    for (size_t k = threadIdx.x; k < sizeof(RegState) / 4; k += blockDim.x) dst[k] = src[k];
    if (threadIdx.x != 0) return;
    const int32_t cnt = ctr[6];
    const int32_t m = cnt < fc->frame_cap ? cnt : fc->frame_cap;
    fc->m = m;
    fc->skipped = m < fc->min_points ? 1 : 0;
    double T1[16];
    const bool ok1 = !fc->skipped && reg_outcome(st, fc->T_pred, fc->min_corr, T1);
    for (int k = 0; k < 16; ++k) fc->T1[k] = T1[k];
    fc->ok1 = ok1 ? 1 : 0;
    bool redo = false;
    float xn[6];
    if (ok1 && fc->redeskew && fc->has_prev) {
        double d[16], rv[3];
        mat4_mul(fc->T_prev_inv, T1, d);
        rotvec_of(d, rv);
        xn[0] = (float)d[3];
        xn[1] = (float)d[7];
        xn[2] = (float)d[11];
        for (int k = 0; k < 3; ++k) xn[3 + k] = (float)rv[k];
        float e[6];
        for (int k = 0; k < 6; ++k) e[k] = xn[k] - fc->xi0[k];
        const float nr = sqrtf(e[3] * e[3] + e[4] * e[4] + e[5] * e[5]);
        const float nt = sqrtf(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        redo = (double)nr > fc->redeskew_rot || (double)nt > fc->redeskew_trans;
    }
    fc->redo = redo ? 1 : 0;
    for (int k = 0; k < 6; ++k) fc->xi1[k] = redo ? xn[k] : fc->xi0[k];
}

// After the (possible) second registration: the frame's pose and trust, and the keyframe test
// (odom.py:650-651, 712-717) against the last keyframe, with the float32 pose an insert uses.
__global__ void k_keyframe_check(const RegState *st, FrameCtl *fc) {
    double T[16];
    bool ok;
    if (fc->redo) {
        ok = reg_outcome(st, fc->T1, fc->min_corr, T);
    } else {
        for (int k = 0; k < 16; ++k) T[k] = fc->T1[k];
        ok = fc->ok1 != 0;
    }
    for (int k = 0; k < 16; ++k) {
        fc->T[k] = T[k];
        fc->T32[k] = (float)T[k];
    }
    fc->ok = ok ? 1 : 0;
    bool key = false;
    if (ok) {
        if (!fc->has_key) {
            key = true;
        } else {
            double d[16], rv[3];
            mat4_mul(fc->T_key_inv, T, d);
            rotvec_of(d, rv);
            const double ang = sqrt(rv[0] * rv[0] + rv[1] * rv[1] + rv[2] * rv[2]) * (180.0 / M_PI);
            const double dist = sqrt(d[3] * d[3] + d[7] * d[7] + d[11] * d[11]);
            key = dist >= fc->kf_dist || ang >= fc->kf_angle;
        }
    }
    fc->keyframe = key ? 1 : 0;
}

}  // namespace

hipError_t launch_frame_check(lw_ctx *c) {
    hipLaunchKernelGGL(k_frame_check, dim3(1), dim3(64), 0, c->stream, c->reg, c->reg1, c->fctl, c->ctr);
    return hipGetLastError();
}

hipError_t launch_keyframe_check(lw_ctx *c) {
    hipLaunchKernelGGL(k_keyframe_check, dim3(1), dim3(1), 0, c->stream, c->reg, c->fctl);
    return hipGetLastError();
}

}  // namespace lw
