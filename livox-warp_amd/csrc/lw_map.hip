// Registration-map build: the MI355X form of k_reg_insert (python/livox_warp/odom.py:186-230) and of the
// map reset (odom.py:561-567).
//
// Layout in HBM (cap slots, V voxels):
//   slots[cap]    32 B {int64 key, int32 voxel id, float count, float mean[3]}: the reference's
//                 open-addressed key table, probed with its hash (gpu.py:247-265), carrying the voxel's
//                 mean so the closest-mean search reads nothing else per neighbour;
//   occ[cap/32]   occupancy bits (bit s set <=> key[s] != -1), 1/256 of the slot table, L2-resident;
//   vox_key[V], vox_slot[V], vox_mom[V] 40 B {sum r, count, sum r r^T (6)}: the moments, dense by voxel id;
//   vox_prec[V] 64 B: the residual precision of each voxel, read once per matched point.
//
// One insert is five short kernels:
//   1 claim    probe and CAS-claim each point's slot (a claimer hands out the voxel id);
//   2 count    per-voxel incoming count, and the list of voxels this insert touches;
//   3 quota    voxels that would pass the 4096-point cap (odom.py:225) rank their incoming points in
//              index order, so exactly the first points contribute, as in a serial insert;
//   4 accum    float32 atomic sums into the moments;
//   5 finish   reset the per-insert counters, queue the touched voxels for a precision refresh (or, when
//              the caller passes the precision parameters, refresh them at once).
#include <climits>

#include "lw_internal.h"
#include "lw_prec.h"

namespace lw {

namespace {

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_clear_slots(Slot *s, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        Slot z;
        z.key = -1;
        z.vid = -1;
        z.cnt = 0.0f;
        z.mu[0] = z.mu[1] = z.mu[2] = 0.0f;
        z.disp = 0;
        s[i] = z;
    }
}

__device__ inline bool insert_point(const float *xyz, const uint32_t *attr, int32_t i, const Mat4f &T, float inv_v,
                                    float max_range, int32_t ret_mask, int32_t &ix, int32_t &iy, int32_t &iz,
                                    float &wx, float &wy, float &wz) {
    float px = xyz[3 * (int64_t)i], py = xyz[3 * (int64_t)i + 1], pz = xyz[3 * (int64_t)i + 2];
    uint32_t a = attr[i];
    if (length3(px, py, pz) > max_range || tag_of(a) != 0 || !ret_ok(a, ret_mask)) return false;
    transform_point(T.m, px, py, pz, wx, wy, wz);
    ix = ifloor(wx * inv_v);
    iy = ifloor(wy * inv_v);
    iz = ifloor(wz * inv_v);
    return true;
}

// The claim of point i: probe and CAS-claim its voxel's slot (a claimer hands out the voxel id).
__device__ inline void claim_point(const MapRefs &M, const float *xyz, const uint32_t *attr, int32_t i, const Mat4f &T,
                                   float inv_v, float max_range, int32_t ret_mask) {
    int32_t ix, iy, iz;
    float wx, wy, wz;
    if (!insert_point(xyz, attr, i, T, inv_v, max_range, ret_mask, ix, iy, iz, wx, wy, wz)) {
        M.pt_slot[i] = -1;
        return;
    }
    int64_t key = voxel_key_ijk(ix, iy, iz);
    int32_t h = slot_hash(key) & M.mask;
    for (int probe = 0; probe < PROBES; ++probe) {
        int32_t s = (h + probe) & M.mask;
        int64_t k = __hip_atomic_load(&M.slots[s].key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (k == key) {
            M.pt_slot[i] = s;
            return;
        }
        if (k == -1) {
            unsigned long long prev = atomicCAS((unsigned long long *)&M.slots[s].key, 0xFFFFFFFFFFFFFFFFull,
                                                (unsigned long long)key);
            if ((int64_t)prev == -1) {
                int32_t vid = atomicAdd(&M.ctr[2], 1);
                if (vid >= M.vox_cap) {  // cannot happen while vox_cap == cap; keep the slot unusable
                    M.slots[s].vid = -1;
                    M.pt_slot[i] = -1;
                    atomicAdd(&M.ctr[1], 1);
                    return;
                }
                M.slots[s].vid = vid;
                M.vox_key[vid] = key;
                M.vox_slot[vid] = s;
                VoxMom z;
                for (int a = 0; a < 3; ++a) z.s[a] = 0.0f;
                for (int a = 0; a < 6; ++a) z.o[a] = 0.0f;
                z.cnt = 0;
                M.vox_mom[vid] = z;
                atomicOr(&M.occ[s >> 5], 1u << (s & 31));
                if (probe > 0) atomicMax(&M.slots[h].disp, probe);  // the home slot bounds this key's probe
                atomicAdd(&M.ctr[0], 1);
                M.pt_slot[i] = s;
                return;
            }
            if ((int64_t)prev == key) {
                M.pt_slot[i] = s;
                return;
            }
        }
    }
    atomicAdd(&M.ctr[1], 1);  // no slot within the probe window (odom.py:230)
    M.pt_slot[i] = -1;
}

// Each insert kernel's work for block b of nb (the single-map kernels: blockIdx.x of gridDim.x; the batched
// ones: the same over x, the map from blockIdx.y).

// 1: claim. A plain read first: a slot's key only ever goes from -1 to its final value.
__device__ inline void claim_blocks(const MapRefs &M, const float *xyz, const uint32_t *attr, int32_t cnt,
                                    const Mat4f &T, float inv_v, float max_range, int32_t ret_mask, int b, int nb) {
    if (b == 0 && threadIdx.x == 0) M.ctr[4] = 0;  // touched count of this insert (read by the next kernel)
    for (int32_t i = b * BLOCK + threadIdx.x; i < cnt; i += nb * BLOCK)
        claim_point(M, xyz, attr, i, T, inv_v, max_range, ret_mask);
}

// 2: count the incoming points per voxel; the first one lists the voxel as touched. counts, when given,
// receives the map's counts (ctr[0..1]), final once the claims are done.
__device__ inline void count_blocks(const MapRefs &M, int32_t cnt, int32_t *counts, int b, int nb) {
    if (counts && b == 0 && threadIdx.x < 2) counts[threadIdx.x] = M.ctr[threadIdx.x];
    for (int32_t i = b * BLOCK + threadIdx.x; i < cnt; i += nb * BLOCK) {
        int32_t s = M.pt_slot[i];
        int32_t vid = s >= 0 ? M.slots[s].vid : -1;
        M.pt_vid[i] = vid;
        if (vid >= 0 && atomicAdd(&M.vox_nin[vid], 1) == 0) M.touched[atomicAdd(&M.ctr[4], 1)] = vid;
    }
}

// 3: per touched voxel, its quota under the cap; voxels that overflow rank their points in index order.
__device__ inline void quota_blocks(const MapRefs &M, int32_t n, int b, int nb) {
    __shared__ int32_t wsum[BLOCK / 64];
    __shared__ int32_t s_vid, s_quota, s_nin;
    int32_t nt = M.ctr[4];
    for (int32_t j = b; j < nt; j += nb) {
        if (threadIdx.x == 0) {
            s_vid = M.touched[j];
            s_nin = M.vox_nin[s_vid];
            s_quota = REG_CAP - M.vox_mom[s_vid].cnt;
        }
        __syncthreads();
        int32_t vid = s_vid, quota = s_quota, nin = s_nin;
        __syncthreads();
        if (nin <= quota) {
            if (threadIdx.x == 0) M.vox_quota[vid] = INT_MAX;
            continue;
        }
        if (threadIdx.x == 0) M.vox_quota[vid] = quota;
        if (quota <= 0) continue;
        // rank this voxel's points in index order
        int32_t base = 0;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int32_t t0 = 0; t0 < n; t0 += BLOCK) {
            int32_t i = t0 + threadIdx.x;
            bool match = i < n && M.pt_vid[i] == vid;
            unsigned long long bal = __ballot(match);
            int32_t lp = __popcll(bal & ((1ull << lane) - 1ull));
            if (lane == 0) wsum[wave] = __popcll(bal);
            __syncthreads();
            int32_t off = 0, tot = 0;
            for (int w = 0; w < BLOCK / 64; ++w) {
                off += (w < wave) ? wsum[w] : 0;
                tot += wsum[w];
            }
            if (match) M.pt_take[i] = (base + off + lp) < quota ? 1 : 0;
            base += tot;
            __syncthreads();
            if (base >= nin) break;  // every point of the voxel is ranked (uniform: base and nin are)
        }
    }
}

// 4: accumulate (odom.py:217, 225-228): r relative to the voxel centre, float32 atomics.
// The moments of the taken points. Points arrive in scan order, so neighbouring lanes often add to one voxel:
// each wave first combines the runs of consecutive lanes with the same voxel id (a segmented shuffle
// reduction of the nine sums and the count), and the run's first lane makes one set of atomics for it. The
// float32 sums come out in another order than one atomic a point (the moments are order-dependent at the
// float32 level either way); the counts are exact.
__device__ inline void accum_blocks(const MapRefs &M, const float *xyz, const uint32_t *attr, int32_t cnt,
                                    const Mat4f &T, float inv_v, float voxel, float max_range, int32_t ret_mask,
                                    int b, int nb) {
    const int lane = lane_id();
    for (int32_t base = b * BLOCK; base < cnt; base += nb * BLOCK) {  // the same trip count for a whole wave
        const int32_t i = base + (int32_t)threadIdx.x;
        int32_t vid = i < cnt ? M.pt_vid[i] : -1;
        if (vid >= 0) {
            const int32_t q = M.vox_quota[vid];
            if (q != INT_MAX && !(q > 0 && M.pt_take[i])) vid = -1;
        }
        float v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (vid >= 0) {
            int32_t ix, iy, iz;
            float wx, wy, wz;
            insert_point(xyz, attr, i, T, inv_v, max_range, ret_mask, ix, iy, iz, wx, wy, wz);
            const float r0 = wx - center_of(ix, voxel), r1 = wy - center_of(iy, voxel), r2 = wz - center_of(iz, voxel);
            v[0] = r0;
            v[1] = r1;
            v[2] = r2;
            v[3] = r0 * r0;
            v[4] = r0 * r1;
            v[5] = r0 * r2;
            v[6] = r1 * r1;
            v[7] = r1 * r2;
            v[8] = r2 * r2;
        }
        // runs of one voxel: a run starts where the id differs from the lane below's (a lane adding nothing
        // has an id of its own)
        const int32_t key = vid >= 0 ? vid : -2 - lane;
        const int32_t below = __shfl_up(key, 1, 64);
        const uint64_t heads = __ballot(lane == 0 || below != key);
        const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const uint64_t after = heads & ~upto;
        const int last = after ? __ffsll((long long)after) - 2 : 63;
        int32_t c = vid >= 0 ? 1 : 0;
        LW_UNROLL for (int off = 1; off < 64; off <<= 1) {  // the run's totals end up at its first lane
            float o[9];
            LW_UNROLL for (int k = 0; k < 9; ++k) o[k] = __shfl_down(v[k], off, 64);
            const int32_t oc = __shfl_down(c, off, 64);
            if (lane + off <= last) {
                LW_UNROLL for (int k = 0; k < 9; ++k) v[k] = v[k] + o[k];
                c += oc;
            }
        }
        const bool head = (heads >> lane) & 1ull;
        if (vid < 0 || !head) continue;
        VoxMom *m = &M.vox_mom[vid];
        LW_UNROLL for (int k = 0; k < 3; ++k) unsafeAtomicAdd(&m->s[k], v[k]);
        LW_UNROLL for (int k = 0; k < 6; ++k) unsafeAtomicAdd(&m->o[k], v[3 + k]);
        atomicAdd(&m->cnt, c);
    }
}

// The mean and precision of voxel vid from its moments (odom.py:336-368), into its slot and vox_prec.
__device__ inline void refresh_voxel(int32_t vid, const int64_t *vox_key, const int32_t *vox_slot,
                                     const VoxMom *vox_mom, Slot *slots, VoxPrec *vox_prec, const PrecParams &q) {
    VoxMom m = vox_mom[vid];
    Slot *sl = &slots[vox_slot[vid]];
    if (m.cnt <= 0) {
        sl->cnt = 0.0f;
        return;
    }
    int32_t ix, iy, iz;
    voxel_key_unpack(vox_key[vid], ix, iy, iz);
    float mu[3];
    voxel_mean(ix, iy, iz, m.s, m.cnt, q.voxel, mu);
    sl->mu[0] = mu[0];
    sl->mu[1] = mu[1];
    sl->mu[2] = mu[2];
    sl->cnt = (float)m.cnt;
    VoxPrec P;
    P.shape = (float)voxel_precision(m.s, m.o, m.cnt, q.voxel, q.noise2, q.tight2, q.alpha, P.M, P.Mg, P.q);
    vox_prec[vid] = P;
}

// 5: finish: clear the per-insert counters; queue the touched voxels for a refresh, or (refresh set) refresh
// them here with q.
__device__ inline void finish_blocks(const MapRefs &M, int32_t refresh, const PrecParams &q, int b, int nb) {
    int32_t nt = M.ctr[4];
    for (int32_t j = b * BLOCK + threadIdx.x; j < nt; j += nb * BLOCK) {
        int32_t vid = M.touched[j];
        M.vox_nin[vid] = 0;
        M.vox_quota[vid] = INT_MAX;
        if (refresh)
            refresh_voxel(vid, M.vox_key, M.vox_slot, M.vox_mom, M.slots, M.vox_prec, q);
        else if (atomicExch(&M.vox_dirty[vid], 1) == 0)
            M.dirty_list[atomicAdd(&M.ctr[3], 1)] = vid;
    }
}

__global__ __launch_bounds__(BLOCK) void k_ins_claim(MapRefs M, const float *xyz, const uint32_t *attr, int32_t n,
                                                     const int32_t *n_dev, Mat4f T, float inv_v, float max_range,
                                                     int32_t ret_mask, const int32_t *gate, const float *T_dev) {
    if (gated_off(gate)) return;
    if (T_dev) T = load_mat4f(T_dev);
    claim_blocks(M, xyz, attr, count_of(n, n_dev), T, inv_v, max_range, ret_mask, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_count(MapRefs M, int32_t n, const int32_t *n_dev, const int32_t *gate,
                                                     int32_t *counts) {
    if (gated_off(gate)) return;
    count_blocks(M, count_of(n, n_dev), counts, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_quota(MapRefs M, int32_t n, const int32_t *n_dev, const int32_t *gate) {
    if (gated_off(gate)) return;
    quota_blocks(M, count_of(n, n_dev), blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_accum(MapRefs M, const float *xyz, const uint32_t *attr, int32_t n,
                                                     const int32_t *n_dev, Mat4f T, float inv_v, float voxel,
                                                     float max_range, int32_t ret_mask, const int32_t *gate,
                                                     const float *T_dev) {
    if (gated_off(gate)) return;
    if (T_dev) T = load_mat4f(T_dev);
    accum_blocks(M, xyz, attr, count_of(n, n_dev), T, inv_v, voxel, max_range, ret_mask, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_finish(MapRefs M, const int32_t *gate, int32_t refresh, PrecParams q) {
    if (gated_off(gate)) return;
    finish_blocks(M, refresh, q, blockIdx.x, gridDim.x);
}

// The batched forms: one insert per sequence (blockIdx.y), each with its own map, points, pose and gate.
__global__ __launch_bounds__(BLOCK) void k_ins_claim_seq(const InsSeq *D, float inv_v, float max_range,
                                                         int32_t ret_mask) {
    const InsSeq d = rebased(D, D[blockIdx.y]);
    if (gated_off(d.gate)) return;
    claim_blocks(d.M, d.xyz, d.attr, count_of(d.n, d.n_dev), load_mat4f(d.T_dev), inv_v, max_range, ret_mask,
                 blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_count_seq(const InsSeq *D) {
    const InsSeq d = rebased(D, D[blockIdx.y]);
    if (gated_off(d.gate)) return;
    count_blocks(d.M, count_of(d.n, d.n_dev), d.counts, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_quota_seq(const InsSeq *D) {
    const InsSeq d = rebased(D, D[blockIdx.y]);
    if (gated_off(d.gate)) return;
    quota_blocks(d.M, count_of(d.n, d.n_dev), blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_accum_seq(const InsSeq *D, float inv_v, float voxel, float max_range,
                                                         int32_t ret_mask) {
    const InsSeq d = rebased(D, D[blockIdx.y]);
    if (gated_off(d.gate)) return;
    accum_blocks(d.M, d.xyz, d.attr, count_of(d.n, d.n_dev), load_mat4f(d.T_dev), inv_v, voxel, max_range, ret_mask,
                 blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(BLOCK) void k_ins_finish_seq(const InsSeq *D, PrecParams q) {
    const InsSeq d = rebased(D, D[blockIdx.y]);
    if (gated_off(d.gate)) return;
    finish_blocks(d.M, 1, q, blockIdx.x, gridDim.x);
}

// Per-voxel mean and precision for the queued voxels (all == 1: every voxel handed out).
__global__ __launch_bounds__(BLOCK) void k_refresh(int32_t all, int32_t *ctr, const int32_t *dirty_list,
                                                   int32_t *vox_dirty, const int64_t *vox_key,
                                                   const int32_t *vox_slot, const VoxMom *vox_mom, Slot *slots,
                                                   VoxPrec *vox_prec, PrecParams q, int64_t vox_cap) {
    int32_t cnt = all ? ctr[2] : ctr[3];
    if (cnt > vox_cap) cnt = (int32_t)vox_cap;
    for (int32_t j = blockIdx.x * BLOCK + threadIdx.x; j < cnt; j += gridDim.x * BLOCK) {
        int32_t vid = all ? j : dirty_list[j];
        vox_dirty[vid] = 0;
        refresh_voxel(vid, vox_key, vox_slot, vox_mom, slots, vox_prec, q);
    }
}

__global__ void k_set_ctr(int32_t *ctr, int32_t idx, int32_t v) { ctr[idx] = v; }

__global__ __launch_bounds__(BLOCK) void k_map_find(const Slot *slots, const uint32_t *occ, int32_t mask,
                                                    const int64_t *keys, int32_t n, int32_t *out) {
    for (int32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += gridDim.x * BLOCK) {
        int64_t key = keys[i];
        int32_t h = slot_hash(key) & mask, res = -1;
        for (int probe = 0; probe < PROBES; ++probe) {
            int32_t s = (h + probe) & mask;
            if (!((occ[s >> 5] >> (s & 31)) & 1u)) break;
            if (slots[s].key == key) {
                res = s;
                break;
            }
        }
        out[i] = res;
    }
}

}  // namespace

MapRefs map_refs(lw_ctx *c) {
    MapRefs M;
    M.slots = c->slots;
    M.occ = c->occ;
    M.vox_key = c->vox_key;
    M.vox_slot = c->vox_slot;
    M.vox_mom = c->vox_mom;
    M.vox_prec = c->vox_prec;
    M.ctr = c->ctr;
    M.pt_slot = c->pt_slot;
    M.pt_vid = c->pt_vid;
    M.pt_take = c->pt_take;
    M.vox_nin = c->vox_nin;
    M.touched = c->touched;
    M.vox_quota = c->vox_quota;
    M.vox_dirty = c->vox_dirty;
    M.dirty_list = c->dirty_list;
    M.vox_cap = c->vox_cap;
    M.mask = c->mask;
    return M;
}

// The localisation's neighbourhood bitmap (NB_SCALE bits a map slot; mask = its bits - 1): each voxel id handed
// out (ctr[2] of them, at most vox_cap) sets two bits, slot_hash and nb_hash2 (each & mask), for each of its 27
// neighbours' keys, formed as voxel_key_ijk forms them (the axis indices decoded from the key, so the 21-bit wrap
// is the search's). A key with either bit clear has no voxel of the map among its own 27 neighbours; both bits
// set may be other keys' (then the search runs as it would without the bitmap).
__global__ __launch_bounds__(BLOCK) void k_nb_build(const int64_t *vox_key, const int32_t *ctr, int32_t vox_cap,
                                                   int32_t mask, uint32_t *nb) {
    const int32_t V = min(ctr[2], vox_cap);
    for (int32_t v = blockIdx.x * BLOCK + threadIdx.x; v < V; v += gridDim.x * BLOCK) {
        int32_t ix, iy, iz;
        voxel_key_unpack(vox_key[v], ix, iy, iz);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const int64_t k = voxel_key_ijk(ix + dx, iy + dy, iz + dz);
                    const int32_t h = slot_hash(k) & mask, h2 = nb_hash2(k) & mask;
                    atomicOr(&nb[h >> 5], 1u << (h & 31));
                    atomicOr(&nb[h2 >> 5], 1u << (h2 & 31));
                }
    }
}

namespace {

// Workgroups for n items, capped at max_grid: every kernel launched with it strides over its items by the grid.
inline int grid_for(int64_t n, int block = BLOCK, int64_t max_grid = 65535) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_grid) g = max_grid;
    return (int)g;
}

}  // namespace

hipError_t launch_fill_i32(lw_ctx *c, int32_t *p, int64_t n, int32_t v) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, BLOCK, 4096)), dim3(BLOCK), 0, c->stream, p, n, v);
    return hipGetLastError();
}

hipError_t launch_map_reset(lw_ctx *c) {
    c->nb_stale = true;
    hipLaunchKernelGGL(k_clear_slots, dim3(grid_for(c->cap, BLOCK, 8192)), dim3(BLOCK), 0, c->stream, c->slots,
                       c->cap);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(c->occ_words, BLOCK, 4096)), dim3(BLOCK), 0, c->stream, c->occ,
                       c->occ_words, 0u);
    // the per-voxel bookkeeping: every voxel id gets its records initialised when it is handed out
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c->vox_cap, BLOCK, 8192)), dim3(BLOCK), 0, c->stream, c->vox_nin,
                       c->vox_cap, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c->vox_cap, BLOCK, 8192)), dim3(BLOCK), 0, c->stream,
                       c->vox_quota, c->vox_cap, INT_MAX);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(c->vox_cap, BLOCK, 8192)), dim3(BLOCK), 0, c->stream,
                       c->vox_dirty, c->vox_cap, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, c->stream, c->ctr, (int64_t)5, 0);
    return hipGetLastError();
}

hipError_t launch_nb_build(lw_ctx *c) {
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(NB_SCALE * c->occ_words, BLOCK, 4096)), dim3(BLOCK), 0, c->stream,
                       c->nb_occ, NB_SCALE * c->occ_words, 0u);
    hipLaunchKernelGGL(k_nb_build, dim3(grid_for(c->vox_cap, BLOCK, 4096)), dim3(BLOCK), 0, c->stream, c->vox_key,
                       c->ctr, (int32_t)c->vox_cap, (int32_t)(NB_SCALE * c->cap - 1), c->nb_occ);
    return hipGetLastError();
}

hipError_t launch_insert_head(lw_ctx *c, const InsertArgs &a) {
    if (a.n <= 0) return hipSuccess;
    c->nb_stale = true;
    const int g = grid_for(a.n);
    const MapRefs M = map_refs(c);
    hipLaunchKernelGGL(k_ins_claim, dim3(g), dim3(BLOCK), 0, c->stream, M, a.xyz, a.attr, a.n, a.n_dev, a.T, a.inv_v,
                       a.max_range, a.ret_mask, a.gate, a.T_dev);
    hipLaunchKernelGGL(k_ins_count, dim3(g), dim3(BLOCK), 0, c->stream, M, a.n, a.n_dev, a.gate, a.counts);
    return hipGetLastError();
}

hipError_t launch_insert_tail(lw_ctx *c, const InsertArgs &a) {
    if (a.n <= 0) return hipSuccess;
    const int g = grid_for(a.n);
    const MapRefs M = map_refs(c);
    hipLaunchKernelGGL(k_ins_quota, dim3(1024), dim3(BLOCK), 0, c->stream, M, a.n, a.n_dev, a.gate);
    hipLaunchKernelGGL(k_ins_accum, dim3(g), dim3(BLOCK), 0, c->stream, M, a.xyz, a.attr, a.n, a.n_dev, a.T, a.inv_v,
                       a.voxel, a.max_range, a.ret_mask, a.gate, a.T_dev);
    const PrecParams q = a.refresh ? *a.refresh : PrecParams{};
    hipLaunchKernelGGL(k_ins_finish, dim3(256), dim3(BLOCK), 0, c->stream, M, a.gate, a.refresh ? 1 : 0, q);
    if (!a.refresh) c->dirty_maybe = true;
    return hipGetLastError();
}

hipError_t launch_insert_seq(hipStream_t s, const InsSeq *D, int32_t S, int32_t n_bound, float inv_v, float voxel,
                             float max_range, int32_t ret_mask, const PrecParams &q) {
    if (S <= 0 || n_bound <= 0) return hipSuccess;
    const int g = grid_for(n_bound);
    hipLaunchKernelGGL(k_ins_claim_seq, dim3(g, S), dim3(BLOCK), 0, s, D, inv_v, max_range, ret_mask);
    hipLaunchKernelGGL(k_ins_count_seq, dim3(g, S), dim3(BLOCK), 0, s, D);
    hipLaunchKernelGGL(k_ins_quota_seq, dim3(64, S), dim3(BLOCK), 0, s, D);
    hipLaunchKernelGGL(k_ins_accum_seq, dim3(g, S), dim3(BLOCK), 0, s, D, inv_v, voxel, max_range, ret_mask);
    hipLaunchKernelGGL(k_ins_finish_seq, dim3(8, S), dim3(BLOCK), 0, s, D, q);
    return hipGetLastError();
}

hipError_t launch_map_insert(lw_ctx *c, const float *xyz, const uint32_t *attr, int32_t n, const int32_t *n_dev,
                             const Mat4f &T, float inv_v, float voxel, float max_range, int32_t ret_mask,
                             const int32_t *gate, const float *T_dev, const PrecParams *refresh) {
    const InsertArgs a{xyz, attr, n, n_dev, T, inv_v, voxel, max_range, ret_mask, gate, T_dev, refresh, nullptr};
    hipError_t e = launch_insert_head(c, a);
    return e != hipSuccess ? e : launch_insert_tail(c, a);
}

hipError_t launch_refresh(lw_ctx *c, bool all, const PrecParams &q) {
    hipLaunchKernelGGL(k_refresh, dim3(all ? 4096 : 512), dim3(BLOCK), 0, c->stream, all ? 1 : 0, c->ctr,
                       c->dirty_list, c->vox_dirty, c->vox_key, c->vox_slot, c->vox_mom, c->slots, c->vox_prec, q,
                       c->vox_cap);
    hipLaunchKernelGGL(k_set_ctr, dim3(1), dim3(1), 0, c->stream, c->ctr, 3, 0);
    c->dirty_maybe = false;
    return hipGetLastError();
}

hipError_t launch_map_find(lw_ctx *c, const int64_t *keys_dev, int32_t n, int32_t *out_dev) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_map_find, dim3(grid_for(n)), dim3(BLOCK), 0, c->stream, c->slots, c->occ, c->mask, keys_dev,
                       n, out_dev);
    return hipGetLastError();
}

}  // namespace lw
