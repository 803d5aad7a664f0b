// Frame cutting, deskew and thinning: k_frame_append (python/livox_warp/odom.py:97-123), k_deskew
// (odom.py:163-178 with gpu.py:268-284) and k_ds_insert (odom.py:126-160).
//
// The reference appends with an atomic counter, so on a GPU its frame order and its thinning winners
// depend on scheduling; Warp's CPU backend runs them in index order. These kernels reproduce the index
// order: appends and thinned points are stream-compacted (per-block count, one-block scan, write), and a
// thinning voxel keeps the point of lowest index (CAS-claim the key, atomicMin the index). The thinning
// table is left clean by the write pass (each claimed slot is reset by its single winner), so no per-call
// clear of the whole table is needed.
#include <climits>

#include "lw_frame.h"
#include "lw_internal.h"

namespace lw {

namespace {

__global__ void k_set_i32_one(int32_t *p, int32_t v) { *p = v; }

// Exclusive prefix of pred across a block of BLOCK threads; *total gets the block's count.
__device__ inline int32_t block_prefix(bool pred, int32_t *wsum, int32_t *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    unsigned long long b = __ballot(pred);
    int32_t lp = __popcll(b & ((1ull << lane) - 1ull));
    if (lane == 0) wsum[wave] = __popcll(b);
    __syncthreads();
    int32_t off = 0, tot = 0;
    for (int w = 0; w < BLOCK / 64; ++w) {
        off += (w < wave) ? wsum[w] : 0;
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return off + lp;
}

// One block: off[b] = base + sum(cnt[0..b)); counter = base + total, base = counter (accumulate) or 0.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_blocks(const int32_t *cnt, int32_t nb, int32_t *off,
                                                              int32_t *counter, int32_t accumulate,
                                                              const int32_t *gate) {
    __shared__ int32_t part[SCAN_THREADS];
    if (gated_off(gate)) return;
    const int t = threadIdx.x;
    const int32_t chunk = (nb + SCAN_THREADS - 1) / SCAN_THREADS;
    const int32_t a = t * chunk, e = min(nb, a + chunk);
    int32_t s = 0;
    for (int32_t b = a; b < e; ++b) s += cnt[b];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < SCAN_THREADS; d <<= 1) {  // inclusive Hillis-Steele
        int32_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const int32_t base = accumulate ? *counter : 0;
    int32_t run = base + (t > 0 ? part[t - 1] : 0);
    for (int32_t b = a; b < e; ++b) {
        off[b] = run;
        run += cnt[b];
    }
    __syncthreads();
    if (t == 0) *counter = base + part[SCAN_THREADS - 1];
}

// A point of time t belongs to the open frame: its frame index floor(max(t, 0) / dt) lies in [k_min, k_open].
__device__ inline bool frame_pred(float t, float inv_dt, int32_t k_min, int32_t k_open) {
    float tt = t > 0.0f ? t : 0.0f;  // wp.max(t, 0.0)
    int32_t k = ifloor(tt * inv_dt);
    return !(k > k_open || k < k_min);
}

__device__ inline bool append_pred(const float *st, int32_t i, float inv_dt, int32_t k_min, int32_t k_open) {
    return frame_pred(st[i], inv_dt, k_min, k_open);
}

// One workgroup appends a batch of at most ONE_THREADS * ONE_ROUNDS points in one launch (the grid path's
// count, scan and write): point i is taken by thread i mod ONE_THREADS in round i / ONE_THREADS, the kept
// points are counted per (round, wave), and written in input order after the frame's count, which the launch
// then advances (reset: the count starts from 0); points past cap are counted and not written, as in the grid
// path.
constexpr int APPEND_BATCH = 8;  // rounds whose loads are issued together

__global__ __launch_bounds__(ONE_THREADS) void k_append_one(const float *sx, const uint32_t *sa, const float *st,
                                                            int32_t n, float inv_dt, int32_t k_min, int32_t k_open,
                                                            int32_t *count, int32_t reset, int32_t cap,
                                                            float *f_raw, uint32_t *f_attr, float *f_t,
                                                            int32_t *f_frame) {
    __shared__ int32_t cnt[ONE_ROUNDS * ONE_WAVES];
    __shared__ int32_t s_base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rounds = (n + ONE_THREADS - 1) / ONE_THREADS;  // the caller keeps n within ONE_ROUNDS rounds
    uint32_t keep = 0u;
    LW_UNROLL for (int b = 0; b < ONE_ROUNDS; b += APPEND_BATCH) {
        if (b >= rounds) break;
        float tv[APPEND_BATCH];
        LW_UNROLL for (int u = 0; u < APPEND_BATCH; ++u) {
            const int32_t i = (b + u) * ONE_THREADS + tid;
            tv[u] = st[i < n ? i : n - 1];
        }
        LW_UNROLL for (int u = 0; u < APPEND_BATCH; ++u) {
            const int32_t i = (b + u) * ONE_THREADS + tid;
            const bool p = b + u < rounds && i < n && frame_pred(tv[u], inv_dt, k_min, k_open);
            keep |= (p ? 1u : 0u) << (b + u);
            const unsigned long long bw = __ballot(p);
            if (lane == 0 && b + u < rounds) cnt[(b + u) * ONE_WAVES + wave] = __popcll(bw);
        }
    }
    __syncthreads();
    if (wave == 0) {
        const int32_t total = scan_round_counts(cnt, rounds);
        if (lane == 0) {
            const int32_t base = reset ? 0 : *count;
            s_base = base;
            *count = base + total;
        }
    }
    __syncthreads();
    const int32_t base = s_base;
    LW_UNROLL for (int b = 0; b < ONE_ROUNDS; b += APPEND_BATCH) {
        if (b >= rounds) break;
        float x[APPEND_BATCH], y[APPEND_BATCH], z[APPEND_BATCH], tv[APPEND_BATCH];
        uint32_t a[APPEND_BATCH];
        LW_UNROLL for (int u = 0; u < APPEND_BATCH; ++u) {
            const int32_t i = (b + u) * ONE_THREADS + tid;
            if ((keep >> (b + u)) & 1u) {
                x[u] = sx[3 * i];
                y[u] = sx[3 * i + 1];
                z[u] = sx[3 * i + 2];
                a[u] = sa[i];
                tv[u] = st[i];
            }
        }
        LW_UNROLL for (int u = 0; u < APPEND_BATCH; ++u) {
            const bool w = ((keep >> (b + u)) & 1u) != 0u;
            const unsigned long long bw = __ballot(w);
            if (w) {
                const int32_t j = base + cnt[(b + u) * ONE_WAVES + wave] + __popcll(bw & ((1ull << lane) - 1ull));
                if (j < cap) {
                    f_raw[3 * j] = x[u];
                    f_raw[3 * j + 1] = y[u];
                    f_raw[3 * j + 2] = z[u];
                    f_attr[j] = a[u];
                    f_t[j] = tv[u];
                    f_frame[j] = k_open;
                }
            }
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_append_count(const float *st, int32_t n, float inv_dt, int32_t k_min,
                                                        int32_t k_open, int32_t *blk_cnt) {
    int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    bool p = i < n && append_pred(st, i, inv_dt, k_min, k_open);
    int32_t c = __syncthreads_count(p);
    if (threadIdx.x == 0) blk_cnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(BLOCK) void k_append_write(const float *sx, const uint32_t *sa, const float *st,
                                                        int32_t n, float inv_dt, int32_t k_min, int32_t k_open,
                                                        const int32_t *blk_off, int32_t cap, float *f_raw,
                                                        uint32_t *f_attr, float *f_t, int32_t *f_frame) {
    __shared__ int32_t wsum[BLOCK / 64];
    int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    bool p = i < n && append_pred(st, i, inv_dt, k_min, k_open);
    int32_t tot;
    int32_t j = blk_off[blockIdx.x] + block_prefix(p, wsum, &tot);
    if (!p || j >= cap) return;
    f_raw[3 * (int64_t)j] = sx[3 * (int64_t)i];
    f_raw[3 * (int64_t)j + 1] = sx[3 * (int64_t)i + 1];
    f_raw[3 * (int64_t)j + 2] = sx[3 * (int64_t)i + 2];
    f_attr[j] = sa[i];
    f_t[j] = st[i];
    f_frame[j] = k_open;
}

// gpu.py:268-284 (deskew_point). xi_dev, when given, holds the twist (rho, theta) in place of the values passed.
__global__ __launch_bounds__(BLOCK) void k_deskew(const float *raw, const float *t, int32_t m, const int32_t *m_dev,
                                                  const int32_t *gate, const float *xi_dev, float t_mid, float inv_dt,
                                                  float r0, float r1, float r2, float h0, float h1, float h2,
                                                  float *out) {
    if (gated_off(gate)) return;
    int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count_of(m, m_dev)) return;
    if (xi_dev) {
        r0 = xi_dev[0];
        r1 = xi_dev[1];
        r2 = xi_dev[2];
        h0 = xi_dev[3];
        h1 = xi_dev[4];
        h2 = xi_dev[5];
    }
    const float xi[6] = {r0, r1, r2, h0, h1, h2};
    float ox, oy, oz;
    deskew_point(raw[3 * (int64_t)i], raw[3 * (int64_t)i + 1], raw[3 * (int64_t)i + 2], t[i], t_mid, inv_dt, xi, ox,
                 oy, oz);
    out[3 * (int64_t)i] = ox;
    out[3 * (int64_t)i + 1] = oy;
    out[3 * (int64_t)i + 2] = oz;
}

// Thinning 1: filters, key at the thinning voxel (sensor frame), claim, lowest index wins.
__global__ __launch_bounds__(BLOCK) void k_thin_claim(const float *xyz, const uint32_t *attr, int32_t m,
                                                      const int32_t *m_dev, float inv_v, float min_range,
                                                      float max_range, int32_t ret_mask, int32_t ds_mask,
                                                      int64_t *ds_keys, int32_t *ds_idx, int32_t *pt_slot,
                                                      const int32_t *gate) {
    if (gated_off(gate)) return;
    int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count_of(m, m_dev)) return;
    pt_slot[i] = -1;
    int64_t key;
    if (!thin_key(xyz[3 * (int64_t)i], xyz[3 * (int64_t)i + 1], xyz[3 * (int64_t)i + 2], attr[i], inv_v, min_range,
                  max_range, ret_mask, key))
        return;
    int32_t h = slot_hash(key) & ds_mask;
    for (int probe = 0; probe < PROBES; ++probe) {
        int32_t s = (h + probe) & ds_mask;
        unsigned long long prev = atomicCAS((unsigned long long *)&ds_keys[s], 0xFFFFFFFFFFFFFFFFull,
                                            (unsigned long long)key);
        if ((int64_t)prev == -1 || (int64_t)prev == key) {
            atomicMin(&ds_idx[s], i);
            pt_slot[i] = s;
            return;
        }
    }
}

__global__ __launch_bounds__(BLOCK) void k_thin_count(const int32_t *pt_slot, const int32_t *ds_idx, int32_t m,
                                                      const int32_t *m_dev, int32_t *blk_cnt, const int32_t *gate) {
    if (gated_off(gate)) return;
    int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    int32_t s = i < count_of(m, m_dev) ? pt_slot[i] : -1;
    bool p = s >= 0 && ds_idx[s] == i;
    int32_t c = __syncthreads_count(p);
    if (threadIdx.x == 0) blk_cnt[blockIdx.x] = c;
}

__global__ __launch_bounds__(BLOCK) void k_thin_write(const float *xyz, const int32_t *pt_slot, int32_t *ds_idx,
                                                      int64_t *ds_keys, int32_t m, const int32_t *m_dev,
                                                      const int32_t *blk_off, float *ds_xyz, const int32_t *gate) {
    __shared__ int32_t wsum[BLOCK / 64];
    if (gated_off(gate)) return;
    int32_t i = blockIdx.x * BLOCK + threadIdx.x;
    int32_t s = i < count_of(m, m_dev) ? pt_slot[i] : -1;
    bool p = s >= 0 && ds_idx[s] == i;  // only the winner of a slot resets it, after reading it
    int32_t tot;
    int32_t j = blk_off[blockIdx.x] + block_prefix(p, wsum, &tot);
    if (!p) return;
    ds_xyz[3 * (int64_t)j] = xyz[3 * (int64_t)i];
    ds_xyz[3 * (int64_t)j + 1] = xyz[3 * (int64_t)i + 1];
    ds_xyz[3 * (int64_t)j + 2] = xyz[3 * (int64_t)i + 2];
    ds_keys[s] = -1;
    ds_idx[s] = INT_MAX;
}

}  // namespace

hipError_t launch_frame_append(lw_ctx *c, const float *sx, const uint32_t *sa, const float *st, int32_t n,
                               float inv_dt, int32_t k_min, int32_t k_open, bool reset) {
    if (n <= 0) return hipSuccess;
    if (n <= ONE_THREADS * ONE_ROUNDS) {  // one launch
        hipLaunchKernelGGL(k_append_one, dim3(1), dim3(ONE_THREADS), 0, c->stream, sx, sa, st, n, inv_dt, k_min,
                           k_open, &c->ctr[6], reset ? 1 : 0, (int32_t)c->frame_cap, c->f_raw, c->f_attr, c->f_t,
                           c->f_frame);
        return hipGetLastError();
    }
    int32_t nb = (n + BLOCK - 1) / BLOCK;
    if (nb > c->max_blocks) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_append_count, dim3(nb), dim3(BLOCK), 0, c->stream, st, n, inv_dt, k_min, k_open,
                       c->blk_cnt);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_THREADS), 0, c->stream, c->blk_cnt, nb, c->blk_off,
                       &c->ctr[6], reset ? 0 : 1, nullptr);
    hipLaunchKernelGGL(k_append_write, dim3(nb), dim3(BLOCK), 0, c->stream, sx, sa, st, n, inv_dt, k_min, k_open,
                       c->blk_off, (int32_t)c->frame_cap, c->f_raw, c->f_attr, c->f_t, c->f_frame);
    return hipGetLastError();
}

// m < 0: the open frame's count on the device (its appended points, at most frame_cap); the grids then cover
// frame_cap points and the blocks past the count do nothing.
hipError_t launch_deskew(lw_ctx *c, int32_t m, float t_mid, float inv_dt, const float *xi, const float *xi_dev,
                         const int32_t *gate) {
    const int32_t *m_dev = m < 0 ? &c->ctr[6] : nullptr;
    if (m < 0) m = (int32_t)c->frame_cap;
    if (m == 0) return hipSuccess;
    const float z[6] = {0, 0, 0, 0, 0, 0};
    if (!xi) xi = z;
    hipLaunchKernelGGL(k_deskew, dim3((m + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, c->stream, c->f_raw, c->f_t, m, m_dev,
                       gate, xi_dev, t_mid, inv_dt, xi[0], xi[1], xi[2], xi[3], xi[4], xi[5], c->f_xyz);
    return hipGetLastError();
}

hipError_t launch_thin(lw_ctx *c, int32_t m, float inv_v, float min_range, float max_range, int32_t ret_mask,
                       const int32_t *gate) {
    const int32_t *m_dev = m < 0 ? &c->ctr[6] : nullptr;
    if (m < 0) m = (int32_t)c->frame_cap;
    if (m == 0) {
        hipLaunchKernelGGL(k_set_i32_one, dim3(1), dim3(1), 0, c->stream, &c->ctr[5], 0);
        return hipGetLastError();
    }
    int32_t nb = (m + BLOCK - 1) / BLOCK;  // blocks past the device count count no points
    hipLaunchKernelGGL(k_thin_claim, dim3(nb), dim3(BLOCK), 0, c->stream, c->f_xyz, c->f_attr, m, m_dev, inv_v,
                       min_range, max_range, ret_mask, (int32_t)(c->ds_cap - 1), c->ds_keys, c->ds_idx, c->pt_slot,
                       gate);
    hipLaunchKernelGGL(k_thin_count, dim3(nb), dim3(BLOCK), 0, c->stream, c->pt_slot, c->ds_idx, m, m_dev,
                       c->blk_cnt, gate);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(SCAN_THREADS), 0, c->stream, c->blk_cnt, nb, c->blk_off,
                       &c->ctr[5], 0, gate);
    hipLaunchKernelGGL(k_thin_write, dim3(nb), dim3(BLOCK), 0, c->stream, c->f_xyz, c->pt_slot, c->ds_idx,
                       c->ds_keys, m, m_dev, c->blk_off, c->ds_xyz, gate);
    return hipGetLastError();
}

}  // namespace lw
