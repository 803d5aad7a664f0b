// LVXR recordings decoded on the device (the reference's firing-level reader, python/livox_warp/lvxr.py:
// 76-210, on the record layout of crates/livox/src/record.rs:1-5 and point.rs): the whole file goes up once,
// the host keeps the framing and the per-packet time base (a sequential pass over the datagrams), and one
// thread per firing decodes its record(s) -- cartesian or spherical, one to three returns, or the Mid-40 dual
// profile's pairs of plain records -- into the firing arrays.
#include <cmath>

#include "../../include/livox_warp_amd.h"
#include "lw_common.h"

namespace lw {

namespace {

constexpr int LVXR_HEADER = 18;  // a data packet's header (point.rs DATA_HEADER_LEN)

__device__ inline int32_t rd_i32(const uint8_t *p) {
    return (int32_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
}
__device__ inline uint32_t rd_u32(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ inline uint32_t rd_u16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }

// lvxr.py:146-150: the unit direction of (theta, phi) in 0.01 degree steps and the point at depth mm, float32
// as numpy forms them (degrees * 0.01, then radians by the float32 pi / 180).
__device__ inline void spherical(uint32_t depth_mm, uint32_t theta, uint32_t phi, float *unit, float *xyz) {
    const float d2r = 3.14159265358979323846f / 180.0f;
    const float th = ((float)theta * 0.01f) * d2r, ph = ((float)phi * 0.01f) * d2r;
    const float st = sinf(th), ct = cosf(th), sp = sinf(ph), cp = cosf(ph);
    unit[0] = st * cp;
    unit[1] = st * sp;
    unit[2] = ct;
    const float r = (float)depth_mm * 1e-3f;
    for (int k = 0; k < 3; ++k) xyz[k] = unit[k] * r;
}

// One record of data type dtype at p: up to three returns (xyz, reflectivity, tag) and, for the spherical
// types, the direction (lvxr.py:130-172).
__device__ inline int decode_record(const uint8_t *p, int dtype, float (*xyz)[3], uint8_t *refl, uint8_t *tag,
                                    float *dir, bool &has_dir) {
    has_dir = false;
    switch (dtype) {
    case 0:
    case 2:
        for (int k = 0; k < 3; ++k) xyz[0][k] = (float)rd_i32(p + 4 * k) * 1e-3f;
        refl[0] = p[12];
        tag[0] = dtype == 2 ? p[13] : 0;
        return 1;
    case 1:
    case 3:
        spherical(rd_u32(p), rd_u16(p + 4), rd_u16(p + 6), dir, xyz[0]);
        has_dir = true;
        refl[0] = p[8];
        tag[0] = dtype == 3 ? p[9] : 0;
        return 1;
    case 4:
    case 7: {
        const int R = dtype == 4 ? 2 : 3;
        for (int r = 0; r < R; ++r) {
            for (int k = 0; k < 3; ++k) xyz[r][k] = (float)rd_i32(p + 14 * r + 4 * k) * 1e-3f;
            refl[r] = p[14 * r + 12];
            tag[r] = p[14 * r + 13];
        }
        return R;
    }
    default: {  // 5, 8: one direction, then (depth, reflectivity, tag) per return
        const int R = dtype == 5 ? 2 : 3;
        const uint32_t theta = rd_u16(p), phi = rd_u16(p + 2);
        for (int r = 0; r < R; ++r) {
            spherical(rd_u32(p + 4 + 6 * r), theta, phi, dir, xyz[r]);
            refl[r] = p[4 + 6 * r + 4];
            tag[r] = p[4 + 6 * r + 5];
        }
        has_dir = true;
        return R;
    }
    }
}

__device__ inline int record_size(int dtype) {
    switch (dtype) {
    case 0: return 13;
    case 1: return 9;
    case 2: return 14;
    case 3: return 10;
    case 4: return 28;
    case 5: return 16;
    case 7: return 42;
    default: return 22;  // 8
    }
}

// Block p decodes packet p's firings into rows first[p] .. first[p] + nf[p] of the firing arrays (R return
// slots a row); t = (t_ns[p] + j interval) 1e-9 s for its j-th firing.
__global__ __launch_bounds__(128) void k_lvxr_decode(const uint8_t *bytes, const int64_t *off, const int32_t *dtype,
                                                    const int32_t *nf, const int64_t *t_ns, const int64_t *first,
                                                    int32_t dual, int64_t interval, int32_t R, double *t,
                                                    float *xyz, uint8_t *refl, uint8_t *tag, uint8_t *valid,
                                                    float *dir) {
    const int p = blockIdx.x;
    const int dt = dtype[p];
    const int size = record_size(dt);
    const uint8_t *body = bytes + off[p] + LVXR_HEADER;
    for (int j = threadIdx.x; j < nf[p]; j += blockDim.x) {
        const int64_t row = first[p] + j;
        float x[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        uint8_t rf[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
        float d[3];
        bool has_dir;
        int nr;
        if (dual) {  // two plain records, the returns of one firing; the direction is the first record's
            float x1[3][3];
            uint8_t rf1[3], tg1[3];
            float d1[3];
            bool h1;
            decode_record(body + (int64_t)(2 * j) * size, dt, x, rf, tg, d, has_dir);
            decode_record(body + (int64_t)(2 * j + 1) * size, dt, x1, rf1, tg1, d1, h1);
            for (int k = 0; k < 3; ++k) x[1][k] = x1[0][k];
            rf[1] = rf1[0];
            tg[1] = tg1[0];
            nr = 2;
        } else {
            nr = decode_record(body + (int64_t)j * size, dt, x, rf, tg, d, has_dir);
        }
        t[row] = (double)(t_ns[p] + (int64_t)j * interval) * 1e-9;
        int first_valid = -1;
        for (int r = 0; r < R; ++r) {
            const bool in = r < nr;
            const float a = in ? x[r][0] : 0.0f, b = in ? x[r][1] : 0.0f, c = in ? x[r][2] : 0.0f;
            xyz[(row * R + r) * 3] = a;
            xyz[(row * R + r) * 3 + 1] = b;
            xyz[(row * R + r) * 3 + 2] = c;
            refl[row * R + r] = in ? rf[r] : 0;
            tag[row * R + r] = in ? tg[r] : 0;
            const bool v = a != 0.0f || b != 0.0f || c != 0.0f;
            valid[row * R + r] = v ? 1 : 0;
            if (v && first_valid < 0) first_valid = r;
        }
        // lvxr.py:241-246: a cartesian firing's direction is its first return's, NaN with none
        if (!has_dir) {
            if (first_valid >= 0) {
                const float a = x[first_valid][0], b = x[first_valid][1], c = x[first_valid][2];
                const float n = length3(a, b, c);
                d[0] = a / n;
                d[1] = b / n;
                d[2] = c / n;
            } else {
                d[0] = d[1] = d[2] = __int_as_float(0x7FC00000);
            }
        }
        for (int k = 0; k < 3; ++k) dir[row * 3 + k] = d[k];
    }
}

}  // namespace

}  // namespace lw

// Host side: the two sequential walks of a recording (record.rs:1-5 framing; net.rs Sink's time base).

extern "C" int lw_lvxr_index(const uint8_t *data, int64_t n, int64_t start, int64_t *offs, int32_t *lens, int64_t cap,
                             int64_t *count) {
    if (!data || n < 0 || start < 0 || cap < 0 || !count || (cap > 0 && (!offs || !lens))) return LW_EINVAL;
    int64_t off = start, k = 0;
    while (off + 10 <= n) {
        const int64_t len = (int64_t)data[off + 8] | ((int64_t)data[off + 9] << 8);
        if (off + 10 + len > n) break;  // a truncated last record is dropped
        if (k < cap) {
            offs[k] = off + 10;
            lens[k] = (int32_t)len;
        }
        ++k;
        off += 10 + len;
    }
    *count = k;
    return (cap > 0 && k > cap) ? LW_EINVAL : 0;  // cap 0: count only
}

extern "C" int lw_lvxr_clock(const uint64_t *ts, const int64_t *spans, int64_t n, int64_t jump_ns, int64_t *out) {
    if (n < 0 || (n > 0 && (!ts || !spans || !out))) return LW_EINVAL;
    uint64_t base = 0, expected = 0;
    bool have_base = false, have_expected = false;
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t t = ts[i];
        const int64_t span = spans[i];
        if (!have_base || t < base) {  // the clock went back: a new base
            base = t;
            have_base = true;
            have_expected = false;
        }
        // late by a second or more (and by more than half a packet): close the gap
        if (have_expected && span > 0 && t > expected + (uint64_t)(span / 2) && t - expected >= (uint64_t)jump_ns)
            base += t - expected;
        expected = t + (uint64_t)span;
        have_expected = true;
        out[i] = (int64_t)(t - base);
    }
    return 0;
}

extern "C" int lw_lvxr_decode(const uint8_t *bytes, const int64_t *pk_off, const int32_t *pk_type, const int32_t *pk_nf,
                              const int64_t *pk_t_ns, const int64_t *pk_first, int32_t n_packets, int32_t dual,
                              int64_t interval_ns, int32_t returns, double *t, float *xyz, uint8_t *refl,
                              uint8_t *tag, uint8_t *valid, float *dir, void *stream) {
    if (n_packets < 0 || returns < 1 || returns > 3 || interval_ns <= 0) return LW_EINVAL;
    if (n_packets == 0) return 0;
    if (!bytes || !pk_off || !pk_type || !pk_nf || !pk_t_ns || !pk_first || !t || !xyz || !refl || !tag || !valid ||
        !dir)
        return LW_EINVAL;
    hipLaunchKernelGGL(lw::k_lvxr_decode, dim3((unsigned)n_packets), dim3(128), 0, (hipStream_t)stream, bytes, pk_off,
                       pk_type, pk_nf, pk_t_ns, pk_first, dual, interval_ns, returns, t, xyz, refl, tag, valid, dir);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : -(int)e;
}
