// Bulk voxelisation of a prior map: the reference's voxel_downsample (python/livox_warp/prior_map.py:55-123,
// 169-203) -- every point's offset from its voxel's centre, its colour and intensity added to the voxel's
// sums in an open-addressed table (the registration map's keys and hash, 128 probes), then the claimed voxels
// compacted in slot order into dense means (flag, exclusive scan, emit).
//
// Layout in HBM: one 64-byte record a slot (VoxRec: key, count, the three offset sums, colour and intensity
// sums), so a point's claim and its eight atomics touch one cache line; the reference keeps four parallel
// arrays. Beside it an occupancy bitmap (one bit a slot, set at the claim): the compaction reads cap / 8
// bytes of bitmap and only the claimed records, instead of every key.
//
// Insert: points arrive in scan order, so neighbouring lanes often fall in one voxel. A wave first combines
// each run of consecutive lanes with the same key (a segmented shuffle reduction of the sums and the count),
// and the run's first lane claims the slot and makes one set of atomics for the whole run. The sums are
// float32 in another order than the reference's serial loop (parity within float32 summation order); the
// claims, counts and drops are exact.
//
// Compaction: a workgroup owns 4096 bitmap words (131072 slots): counts (popcounts), one workgroup scans the
// counts, then each thread emits the set bits of its 16 words in slot order at its offset.
#include <climits>
#include <new>

#include "../../include/livox_warp_amd.h"
#include "lw_common.h"
#include "lw_gn.h"

namespace lw {

namespace {

constexpr int VOX_PROBES = 128;  // prior_map.py:36
constexpr int VOX_BLOCK = 256;
constexpr int VOX_WORDS = 16;                              // bitmap words a thread, in the compaction
constexpr int VOX_SPAN = VOX_BLOCK * VOX_WORDS;            // bitmap words a workgroup
constexpr int VOX_SCAN = 1024;                             // threads of the scan workgroup

struct VoxRec {
    int64_t key;
    int32_t cnt;
    float r[3];  // offsets from the voxel centre, summed
    float c[4];  // r, g, b, intensity, summed
    int32_t pad[6];
};
static_assert(sizeof(VoxRec) == 64, "one 64-byte record a slot");

// prior_map.py:50-52
__device__ inline void vox_centre(int32_t ix, int32_t iy, int32_t iz, float voxel, float &x, float &y, float &z) {
    x = ((float)ix + 0.5f) * voxel;
    y = ((float)iy + 0.5f) * voxel;
    z = ((float)iz + 0.5f) * voxel;
}

__global__ void k_vox_clear(VoxRec *recs, uint32_t *occ, int64_t cap, int32_t *stats) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < cap; s += stride) {
        VoxRec r{};
        r.key = -1;
        recs[s] = r;
    }
    for (int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; w < cap / 32; w += stride) occ[w] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 2) stats[threadIdx.x] = 0;
}

// prior_map.py:55-91, with each wave's runs of one voxel combined before the atomics.
__global__ __launch_bounds__(VOX_BLOCK) void k_vox_insert(const float *__restrict__ xyz, const float *inten,
                                                          const float *rgb, int64_t n, float inv_v, float voxel,
                                                          int64_t mask, VoxRec *recs, uint32_t *occ,
                                                          int32_t *stats) {
    const int64_t i = blockIdx.x * (int64_t)VOX_BLOCK + threadIdx.x;
    const int lane = lane_id();
    const bool live = i < n;
    float ox = 0.0f, oy = 0.0f, oz = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
    int64_t key = -2 - lane;  // a dead lane's key matches no other lane's (voxel keys are non-negative)
    if (live) {
        const float px = xyz[3 * i], py = xyz[3 * i + 1], pz = xyz[3 * i + 2];
        const int32_t ix = ifloor(px * inv_v), iy = ifloor(py * inv_v), iz = ifloor(pz * inv_v);
        key = voxel_key_ijk(ix, iy, iz);
        float cx, cy, cz;
        vox_centre(ix, iy, iz, voxel, cx, cy, cz);
        ox = px - cx;
        oy = py - cy;
        oz = pz - cz;
        if (rgb) {
            c0 = rgb[3 * i];
            c1 = rgb[3 * i + 1];
            c2 = rgb[3 * i + 2];
        }
        if (inten) c3 = inten[i];
    }
    // runs of equal keys: a run starts where the key differs from the lane below's
    const int64_t below = __shfl_up(key, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || below != key);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const int head = 63 - __clzll(heads & upto);
    const uint64_t after = heads & ~upto;
    const int last = after ? __ffsll((long long)after) - 2 : 63;
    int32_t cnt = live ? 1 : 0;
    LW_UNROLL for (int off = 1; off < 64; off <<= 1) {  // the run's totals end up at its head
        const float a0 = __shfl_down(ox, off, 64), a1 = __shfl_down(oy, off, 64), a2 = __shfl_down(oz, off, 64);
        const float b0 = __shfl_down(c0, off, 64), b1 = __shfl_down(c1, off, 64), b2 = __shfl_down(c2, off, 64);
        const float b3 = __shfl_down(c3, off, 64);
        const int32_t oc = __shfl_down(cnt, off, 64);
        if (lane + off <= last) {
            ox = ox + a0;
            oy = oy + a1;
            oz = oz + a2;
            c0 = c0 + b0;
            c1 = c1 + b1;
            c2 = c2 + b2;
            c3 = c3 + b3;
            cnt += oc;
        }
    }
    if (lane != head || !live) return;
    const int64_t h = (int64_t)slot_hash(key) & mask;
    for (int probe = 0; probe < VOX_PROBES; ++probe) {
        const int64_t s = (h + probe) & mask;
        VoxRec *r = &recs[s];
        const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long *>(&r->key),
                                                  0xFFFFFFFFFFFFFFFFull, (unsigned long long)key);
        if ((int64_t)prev == -1 || (int64_t)prev == key) {
            if ((int64_t)prev == -1) {
                atomicAdd(&stats[0], 1);
                atomicOr(&occ[s >> 5], 1u << (s & 31));
            }
            unsafeAtomicAdd(&r->r[0], ox);
            unsafeAtomicAdd(&r->r[1], oy);
            unsafeAtomicAdd(&r->r[2], oz);
            unsafeAtomicAdd(&r->c[0], c0);
            unsafeAtomicAdd(&r->c[1], c1);
            unsafeAtomicAdd(&r->c[2], c2);
            unsafeAtomicAdd(&r->c[3], c3);
            atomicAdd(&r->cnt, cnt);
            return;
        }
    }
    atomicAdd(&stats[1], cnt);  // no slot along the probe chain: the run's points are dropped
}

// The claimed slots of each workgroup's VOX_SPAN bitmap words.
__global__ __launch_bounds__(VOX_BLOCK) void k_vox_count(const uint32_t *occ, int64_t words, int32_t *counts) {
    __shared__ int32_t part[VOX_BLOCK / 64];
    const int64_t w0 = (int64_t)blockIdx.x * VOX_SPAN;
    int32_t c = 0;
    for (int k = threadIdx.x; k < VOX_SPAN; k += VOX_BLOCK)
        if (w0 + k < words) c += __popc(occ[w0 + k]);
    LW_UNROLL for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
    if (lane_id() == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

// Exclusive scan of the workgroups' counts, in place, by one workgroup.
__global__ __launch_bounds__(VOX_SCAN) void k_vox_scan(int32_t *counts, int32_t nblk) {
    __shared__ int32_t part[VOX_SCAN];
    const int per = (nblk + VOX_SCAN - 1) / VOX_SCAN;
    const int b0 = threadIdx.x * per;
    int32_t s = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < nblk) s += counts[b0 + k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < VOX_SCAN; off <<= 1) {  // inclusive scan of the threads' sums (Hillis-Steele)
        const int32_t v = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    int32_t run = part[threadIdx.x] - s;  // exclusive prefix of this thread's first block
    for (int k = 0; k < per; ++k) {
        if (b0 + k < nblk) {
            const int32_t c = counts[b0 + k];
            counts[b0 + k] = run;
            run += c;
        }
    }
}

// prior_map.py:100-123: each thread's VOX_WORDS words in slot order, at the workgroup's offset plus the
// threads before it.
__global__ __launch_bounds__(VOX_BLOCK) void k_vox_emit(const uint32_t *occ, int64_t words, const int32_t *offs,
                                                        const VoxRec *recs, float voxel, float *out_xyz,
                                                        float *out_col, int32_t *out_cnt) {
    __shared__ int32_t pre[VOX_BLOCK];
    const int64_t w0 = (int64_t)blockIdx.x * VOX_SPAN + (int64_t)threadIdx.x * VOX_WORDS;
    uint32_t bits[VOX_WORDS];
    int32_t c = 0;
    LW_UNROLL for (int k = 0; k < VOX_WORDS; ++k) {
        bits[k] = w0 + k < words ? occ[w0 + k] : 0u;
        c += __popc(bits[k]);
    }
    pre[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < VOX_BLOCK; off <<= 1) {  // inclusive scan of the threads' counts
        const int32_t v = threadIdx.x >= off ? pre[threadIdx.x - off] : 0;
        __syncthreads();
        pre[threadIdx.x] += v;
        __syncthreads();
    }
    int64_t j = (int64_t)offs[blockIdx.x] + pre[threadIdx.x] - c;
    for (int k = 0; k < VOX_WORDS; ++k) {
        uint32_t b = bits[k];
        while (b) {
            const int bit = __ffs(b) - 1;
            b &= b - 1u;
            const int64_t s = ((w0 + k) << 5) + bit;
            const VoxRec r = recs[s];
            int32_t ix, iy, iz;
            voxel_key_unpack(r.key, ix, iy, iz);
            float cx, cy, cz;
            vox_centre(ix, iy, iz, voxel, cx, cy, cz);
            const float fc = (float)r.cnt;
            out_xyz[3 * j] = cx + r.r[0] / fc;
            out_xyz[3 * j + 1] = cy + r.r[1] / fc;
            out_xyz[3 * j + 2] = cz + r.r[2] / fc;
            LW_UNROLL for (int a = 0; a < 4; ++a) out_col[4 * j + a] = r.c[a] / fc;
            out_cnt[j] = r.cnt;
            ++j;
        }
    }
}

// A table's records into plain arrays (slot order), for readers on the host.
__global__ void k_vox_unpack(const VoxRec *recs, int64_t cap, int64_t *keys, float *rsum, float *csum,
                             int32_t *cnt) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < cap; s += (int64_t)gridDim.x * blockDim.x) {
        const VoxRec r = recs[s];
        keys[s] = r.key;
        LW_UNROLL for (int a = 0; a < 3; ++a) rsum[3 * s + a] = r.r[a];
        LW_UNROLL for (int a = 0; a < 4; ++a) csum[4 * s + a] = r.c[a];
        cnt[s] = r.cnt;
    }
}

unsigned grid_for(int64_t n, int64_t cap_blocks = 1 << 20) {
    const int64_t b = (n + VOX_BLOCK - 1) / VOX_BLOCK;
    return (unsigned)(b < 1 ? 1 : (b > cap_blocks ? cap_blocks : b));
}

}  // namespace

}  // namespace lw

struct lw_vox {
    hipStream_t stream = nullptr;
    int64_t cap = 0;
    lw::VoxRec *recs = nullptr;
    uint32_t *occ = nullptr;
    int32_t *stats = nullptr;   // [voxels claimed, points dropped]
    int32_t *counts = nullptr;  // the compaction's per-workgroup counts, then offsets
    int32_t nblk = 0;
};

namespace {

#define VOX_CK(x)                              \
    do {                                       \
        hipError_t e_ = (x);                   \
        if (e_ != hipSuccess) return -(int)e_; \
    } while (0)

void vox_free(lw_vox *v) {
    void *bufs[] = {v->recs, v->occ, v->stats, v->counts};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
}

}  // namespace

extern "C" {

int lw_vox_create(lw_vox **out, int32_t slots_log2, void *stream) {
    if (!out || slots_log2 < 10 || slots_log2 > 30) return LW_EINVAL;
    lw_vox *v = new (std::nothrow) lw_vox();
    if (!v) return LW_EINVAL;
    v->stream = (hipStream_t)stream;
    v->cap = 1ll << slots_log2;
    const int64_t words = v->cap / 32;
    v->nblk = (int32_t)((words + lw::VOX_SPAN - 1) / lw::VOX_SPAN);
    hipError_t e = hipMalloc((void **)&v->recs, (size_t)v->cap * sizeof(lw::VoxRec));
    if (e == hipSuccess) e = hipMalloc((void **)&v->occ, (size_t)words * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&v->stats, 2 * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&v->counts, (size_t)v->nblk * sizeof(int32_t));
    if (e != hipSuccess) {
        vox_free(v);
        delete v;
        return -(int)e;
    }
    *out = v;
    return lw_vox_clear(v);
}

int lw_vox_destroy(lw_vox *v) {
    if (!v) return LW_EINVAL;
    (void)hipStreamSynchronize(v->stream);
    vox_free(v);
    delete v;
    return 0;
}

int lw_vox_clear(lw_vox *v) {
    if (!v) return LW_EINVAL;
    hipLaunchKernelGGL(lw::k_vox_clear, dim3(lw::grid_for(v->cap, 65536)), dim3(lw::VOX_BLOCK), 0, v->stream,
                       v->recs, v->occ, v->cap, v->stats);
    VOX_CK(hipGetLastError());
    return 0;
}

int lw_vox_insert(lw_vox *v, const float *xyz, const float *intensity, const float *rgb, int64_t n, double voxel) {
    if (!v || n < 0 || (n > 0 && !xyz) || !(voxel > 0.0) || n > (1ll << 31) * lw::VOX_BLOCK) return LW_EINVAL;
    if (n == 0) return 0;
    // the reference's launch arguments: 1.0 / voxel and voxel, each rounded to float32 (prior_map.py:187)
    const float inv_v = (float)(1.0 / voxel), vf = (float)voxel;
    hipLaunchKernelGGL(lw::k_vox_insert, dim3((unsigned)((n + lw::VOX_BLOCK - 1) / lw::VOX_BLOCK)),
                       dim3(lw::VOX_BLOCK), 0, v->stream, xyz, intensity, rgb, n, inv_v, vf, v->cap - 1, v->recs,
                       v->occ, v->stats);
    VOX_CK(hipGetLastError());
    return 0;
}

int lw_vox_stats(lw_vox *v, int32_t *out) {
    if (!v || !out) return LW_EINVAL;
    VOX_CK(hipMemcpyAsync(out, v->stats, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, v->stream));
    VOX_CK(hipStreamSynchronize(v->stream));
    return 0;
}

int lw_vox_emit(lw_vox *v, double voxel, float *out_xyz, float *out_col, int32_t *out_cnt, int64_t out_cap,
                int64_t *count) {
    if (!v || !count || !(voxel > 0.0) || out_cap < 0 || (out_cap > 0 && (!out_xyz || !out_col || !out_cnt)))
        return LW_EINVAL;
    int32_t st[2];
    int rc = lw_vox_stats(v, st);
    if (rc) return rc;
    *count = st[0];
    if (st[0] > out_cap) return LW_ERANGE;
    if (st[0] == 0) return 0;
    const int64_t words = v->cap / 32;
    hipLaunchKernelGGL(lw::k_vox_count, dim3((unsigned)v->nblk), dim3(lw::VOX_BLOCK), 0, v->stream, v->occ, words,
                       v->counts);
    hipLaunchKernelGGL(lw::k_vox_scan, dim3(1), dim3(lw::VOX_SCAN), 0, v->stream, v->counts, v->nblk);
    hipLaunchKernelGGL(lw::k_vox_emit, dim3((unsigned)v->nblk), dim3(lw::VOX_BLOCK), 0, v->stream, v->occ, words,
                       (const int32_t *)v->counts, v->recs, (float)voxel, out_xyz, out_col, out_cnt);
    VOX_CK(hipGetLastError());
    return 0;
}

int lw_vox_read(lw_vox *v, int64_t *keys, float *rsum, float *csum, int32_t *cnt) {
    if (!v || !keys || !rsum || !csum || !cnt) return LW_EINVAL;
    hipLaunchKernelGGL(lw::k_vox_unpack, dim3(lw::grid_for(v->cap, 65536)), dim3(lw::VOX_BLOCK), 0, v->stream,
                       v->recs, v->cap, keys, rsum, csum, cnt);
    VOX_CK(hipGetLastError());
    return 0;
}

}  // extern "C"
