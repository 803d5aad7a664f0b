// Shared device definitions for the scan-to-map odometry kernels (gfx950).
//
// Key layout, hash and probe follow python/livox_warp/gpu.py:59-64 and :231-265 of the reference;
// constants follow python/livox_warp/odom.py:72-94. Every kernel translation unit is compiled with
// -ffp-contract=off: the float32 steps that decide integer results (voxel indices, closest mean)
// must round each multiply and add on its own, as the serial float32 restatement does.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lw {

constexpr int KEY_BITS = 21;
constexpr uint64_t KEY_BIAS = 1ull << 20;
constexpr uint64_t KEY_MASK = (1ull << 21) - 1ull;
constexpr uint64_t HASH_MUL = 0x9E3779B97F4A7C15ull;  // gpu.py:64 (int64 -7046029254386353131)
constexpr int PROBES = 64;                            // gpu.py:257, odom.py:153/219
constexpr int REG_CAP = 4096;                         // odom.py:83
constexpr int ACC_STRIDE = 64;                        // odom.py:82
constexpr int ACC_H = 0, ACC_B = 21, ACC_COUNT = 27, ACC_ERR = 28, ACC_RMS = 29, ACC_WEIGHT = 30,
              ACC_RANGE = 31, ACC_HG = 32;            // odom.py:73-81
constexpr int ACC_USED = 53;                          // 21 + 6 + 5 + 21

// The calling thread's lane in its wave (64 lanes).
__device__ inline int lane_id() { return (int)(threadIdx.x & 63u); }

// A device pointer p read from a descriptor in device memory (the batched kernels' per-sequence tables),
// re-derived from a kernel argument (base): the compiler then knows it points into global memory and accesses
// it with global instructions (a scalar base, 32-bit lane offsets) instead of flat ones, which also count
// against the LDS counter. The value is p itself (null stays null).
template <class T>
__device__ inline T *via_arg(const void *base, T *p) {
    const int64_t off = (int64_t)(reinterpret_cast<const char *>(p) - reinterpret_cast<const char *>(base));
    return p ? reinterpret_cast<T *>(const_cast<char *>(reinterpret_cast<const char *>(base)) + off) : nullptr;
}

// A barrier among the lanes of one wave, with LDS written before it visible to every lane after it (the
// workgroup barrier would also wait for the workgroup's other waves, which may be at another barrier).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One hash-table slot (32 B): the reference's int64 key, the dense voxel id of its records, and a copy of
// the voxel's mean and count (float), so the closest-mean search reads one slot per neighbour and nothing
// else. Mean and count are written by the precision refresh (count 0 until then: the voxel is skipped).
// disp is the largest probe distance of any key whose home (slot_hash & mask) is this slot, 0 when none
// was displaced: a lookup whose home slot holds another key needs to look no further than that.
struct __align__(32) Slot {
    int64_t key;
    int32_t vid;
    float cnt;
    float mu[3];
    int32_t disp;
};

// Moments of one voxel, relative to the voxel centre (odom.py:198-200): sum r, count, and the six
// distinct entries of sum r r^T (xx xy xz yy yz zz).
struct __align__(8) VoxMom {
    float s[3];
    int32_t cnt;
    float o[6];
};

// Precision of one voxel for the residual (odom.py:336-368): M and its tight part Mg (xx xy xz yy yz
// zz), the tightest direction q and the branch (1: shape, rn = r.q; 0: weak isotropic, rn = |r|).
struct __align__(16) VoxPrec {
    float M[6];
    float Mg[6];
    float q[3];
    float shape;
};

__host__ __device__ inline int64_t voxel_key_ijk(int32_t ix, int32_t iy, int32_t iz) {
    uint64_t x = ((uint64_t)(int64_t)ix + KEY_BIAS) & KEY_MASK;
    uint64_t y = ((uint64_t)(int64_t)iy + KEY_BIAS) & KEY_MASK;
    uint64_t z = ((uint64_t)(int64_t)iz + KEY_BIAS) & KEY_MASK;
    return (int64_t)(x | (y << KEY_BITS) | (z << (2 * KEY_BITS)));
}

// prior_map.py:41-49: the axis indices a key packs.
__host__ __device__ inline void voxel_key_unpack(int64_t key, int32_t &ix, int32_t &iy, int32_t &iz) {
    uint64_t k = (uint64_t)key;
    ix = (int32_t)((int64_t)(k & KEY_MASK) - (int64_t)KEY_BIAS);
    iy = (int32_t)((int64_t)((k >> KEY_BITS) & KEY_MASK) - (int64_t)KEY_BIAS);
    iz = (int32_t)((int64_t)((k >> (2 * KEY_BITS)) & KEY_MASK) - (int64_t)KEY_BIAS);
}

__host__ __device__ inline int32_t slot_hash(int64_t k) {
    uint64_t h = (uint64_t)k * HASH_MUL;
    h ^= h >> 29;
    return (int32_t)(h & 0x7FFFFFFFull);
}

__device__ inline int32_t ifloor(float x) { return (int32_t)floorf(x); }

// A second hash of a voxel key, independent of slot_hash (the high bits of another odd multiplier): the
// neighbourhood bitmap sets and tests two bits a key.
__host__ __device__ inline int32_t nb_hash2(int64_t k) {
    return (int32_t)(((uint64_t)k * 0xD6E8FEB86659FD93ull) >> 33);
}

// A launch whose work a device flag decides: it runs unless gate is given and holds 0.
__device__ inline bool gated_off(const int32_t *gate) { return gate && *gate == 0; }

// A launch's point count: n, or, when n_dev is given, the count on the device capped at n (the grid covers n).
__device__ inline int32_t count_of(int32_t n, const int32_t *n_dev) {
    if (!n_dev) return n;
    const int32_t d = *n_dev;
    return d < n ? d : n;
}

__device__ inline int tag_of(uint32_t a) { return (int)((a >> 8) & 255u); }
__device__ inline bool ret_ok(uint32_t a, int ret_mask) {
    int r = (int)((a >> 16) & 255u);
    return r < 31 && ((1 << r) & ret_mask) != 0;
}

__device__ inline float length3(float x, float y, float z) {
    float d = x * x;
    d = d + y * y;
    d = d + z * z;
    return sqrtf(d);
}

__device__ inline float center_of(int32_t i, float voxel) { return ((float)i + 0.5f) * voxel; }

// wp.transform_point, row r = ((m[r][0] p0 + m[r][1] p1) + m[r][2] p2) + m[r][3].
__device__ inline void transform_point(const float *T, float px, float py, float pz, float &wx, float &wy,
                                       float &wz) {
    float a;
    a = T[0] * px; a = a + T[1] * py; a = a + T[2] * pz; wx = a + T[3];
    a = T[4] * px; a = a + T[5] * py; a = a + T[6] * pz; wy = a + T[7];
    a = T[8] * px; a = a + T[9] * py; a = a + T[10] * pz; wz = a + T[11];
}

// ---- float32 accumulation terms in the reference's order, structural zeros of J skipped ----------
//
// J = [I3 | -[w]x] (odom.py:305-307): rows J0 = (1,0,0,0,w2,-w1), J1 = (0,1,0,-w2,0,w0),
// J2 = (0,0,1,w1,-w0,0). Adding an exact zero leaves a float sum unchanged (up to the sign of a zero
// result), so dropping the terms whose J factor is a structural zero reproduces the reference's
// float32 value of every entry.

// kind of J[i][p]: 0 zero, 1 one, 2 value
__host__ __device__ constexpr int jkind(int i, int p) {
    return p < 3 ? (i == p ? 1 : 0) : ((i == 0 && p == 3) || (i == 1 && p == 4) || (i == 2 && p == 5) ? 0 : 2);
}

template <int I, int P>
__device__ inline float jval(float w0, float w1, float w2) {
    // only called for jkind == 2
    if constexpr (I == 0 && P == 4) return w2;
    else if constexpr (I == 0 && P == 5) return -w1;
    else if constexpr (I == 1 && P == 3) return -w2;
    else if constexpr (I == 1 && P == 5) return w0;
    else if constexpr (I == 2 && P == 3) return w1;
    else return -w0;  // I == 2 && P == 4
}

template <int I, int P>
__device__ inline float jget(float w0, float w1, float w2) {
    if constexpr (jkind(I, P) == 1) return 1.0f;
    else if constexpr (jkind(I, P) == 0) return 0.0f;
    else return jval<I, P>(w0, w1, w2);
}

// term (M[i][j] * J[i][p]) * J[j][q], or "absent" when a J factor is a structural zero
template <int I, int J, int P, int Q>
__device__ inline bool term(const float (&M)[3][3], float w0, float w1, float w2, float &t) {
    if constexpr (jkind(I, P) == 0 || jkind(J, Q) == 0) {
        return false;
    } else {
        float a = M[I][J];
        if constexpr (jkind(I, P) == 2) a = a * jval<I, P>(w0, w1, w2);
        if constexpr (jkind(J, Q) == 2) a = a * jval<J, Q>(w0, w1, w2);
        t = a;
        return true;
    }
}

template <int I, int P, int Q>
__device__ inline bool row_sum(const float (&M)[3][3], float w0, float w1, float w2, float &r) {
    float t;
    bool have = false;
    if (term<I, 0, P, Q>(M, w0, w1, w2, t)) { r = t; have = true; }
    if (term<I, 1, P, Q>(M, w0, w1, w2, t)) { r = have ? r + t : t; have = true; }
    if (term<I, 2, P, Q>(M, w0, w1, w2, t)) { r = have ? r + t : t; have = true; }
    return have;
}

// v(p, q) of odom.py:250-252 (and :268-270)
template <int P, int Q>
__device__ inline float hval(const float (&M)[3][3], float w0, float w1, float w2) {
    float v = 0.0f, r;
    bool have = false;
    if (row_sum<0, P, Q>(M, w0, w1, w2, r)) { v = r; have = true; }
    if (row_sum<1, P, Q>(M, w0, w1, w2, r)) { v = have ? v + r : r; have = true; }
    if (row_sum<2, P, Q>(M, w0, w1, w2, r)) { v = have ? v + r : r; have = true; }
    return v;
}

// (Mr[0] * J0[p] + Mr[1] * J1[p]) + Mr[2] * J2[p] (odom.py:256)
template <int P>
__device__ inline float bval(const float (&Mr)[3], float w0, float w1, float w2) {
    float v = 0.0f;
    bool have = false;
    if constexpr (jkind(0, P) != 0) { float t = Mr[0]; if constexpr (jkind(0, P) == 2) t = t * jval<0, P>(w0, w1, w2); v = t; have = true; }
    if constexpr (jkind(1, P) != 0) { float t = Mr[1]; if constexpr (jkind(1, P) == 2) t = t * jval<1, P>(w0, w1, w2); v = have ? v + t : t; have = true; }
    if constexpr (jkind(2, P) != 0) { float t = Mr[2]; if constexpr (jkind(2, P) == 2) t = t * jval<2, P>(w0, w1, w2); v = have ? v + t : t; have = true; }
    (void)have;
    return v;
}

}  // namespace lw
