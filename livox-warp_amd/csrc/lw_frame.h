// Per-point steps of the frame kernels, shared by the grid kernels of lw_frame.hip and the single-workgroup
// frame preparation of lw_odom.hip, so that both compute every point the same way.
#pragma once

#include "lw_common.h"

namespace lw {

// gpu.py:268-284: rotate by the rotation vector theta s (unit quaternion), then add rho s, with
// s = clamp((t - t_mid) inv_dt, -1, 1) and xi = (rho, theta).
__device__ inline void deskew_point(float px, float py, float pz, float tp, float t_mid, float inv_dt,
                                    const float (&xi)[6], float &qx_out, float &qy_out, float &qz_out) {
    float s = (tp - t_mid) * inv_dt;
    s = fminf(fmaxf(s, -1.0f), 1.0f);
    float vx = xi[3] * s, vy = xi[4] * s, vz = xi[5] * s;
    float a = length3(vx, vy, vz);
    float ox = px, oy = py, oz = pz;
    if (!(a < 1.0e-9f)) {
        float ax = vx / a, ay = vy / a, az = vz / a;
        float sh, ch;
        sincosf(a * 0.5f, &sh, &ch);
        float qx = ax * sh, qy = ay * sh, qz = az * sh, qw = ch;
        float c = 2.0f * qw * qw - 1.0f;
        float d = 2.0f * (qx * px + qy * py + qz * pz);
        float cx = qy * pz - qz * py, cy = qz * px - qx * pz, cz = qx * py - qy * px;
        float w2 = 2.0f * qw;
        ox = px * c + qx * d + cx * w2;
        oy = py * c + qy * d + cy * w2;
        oz = pz * c + qz * d + cz * w2;
    }
    qx_out = ox + xi[0] * s;
    qy_out = oy + xi[1] * s;
    qz_out = oz + xi[2] * s;
}

// Thinning filters (odom.py:126-160): a point inside [min_range, max_range], untagged and of an accepted
// return takes part, keyed at its thinning voxel (sensor frame).
__device__ inline bool thin_key(float px, float py, float pz, uint32_t a, float inv_v, float min_range,
                                float max_range, int32_t ret_mask, int64_t &key) {
    const float r = length3(px, py, pz);
    if (r < min_range || r > max_range) return false;
    if (tag_of(a) != 0 || !ret_ok(a, ret_mask)) return false;
    key = voxel_key_ijk(ifloor(px * inv_v), ifloor(py * inv_v), ifloor(pz * inv_v));
    return true;
}

// The single-workgroup kernels (frame preparation, batch append) count their kept points per (round, wave):
// cnt[r * 16 + w] for round r < rounds and wave w of a 1024-thread workgroup. Wave 0 (every lane) turns the
// counts into exclusive offsets in (round, wave) order, in place, and returns the total in every lane.
constexpr int ONE_THREADS = 1024;
constexpr int ONE_WAVES = ONE_THREADS / 64;
constexpr int ONE_ROUNDS = 32;  // points a thread takes: at most ONE_THREADS * ONE_ROUNDS points a launch

__device__ inline int32_t scan_round_counts(int32_t *cnt, int rounds) {
    constexpr int E = ONE_ROUNDS * ONE_WAVES / 64;
    const int lane = threadIdx.x & 63;
    const int n = rounds * ONE_WAVES;
    int32_t v[E];
    int32_t sum = 0;
    for (int e = 0; e < E; ++e) {
        const int idx = lane * E + e;
        v[e] = idx < n ? cnt[idx] : 0;
        sum += v[e];
    }
    int32_t incl = sum;
    for (int d = 1; d < 64; d <<= 1) {
        const int32_t y = __shfl_up(incl, d);
        if (lane >= d) incl += y;
    }
    int32_t run = incl - sum;
    for (int e = 0; e < E; ++e) {
        const int idx = lane * E + e;
        if (idx < n) cnt[idx] = run;
        run += v[e];
    }
    return __shfl(incl, 63);
}

}  // namespace lw
