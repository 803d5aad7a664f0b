// Batched global localisation: many pose hypotheses x query scans registered against one prior map in one
// persistent launch. The reference's counterpart is GlobalLocalizer._icp (python/livox_warp/localize.py:517-553,
// kernels k_icp_batch / k_reduce_batch :126-179), which registers up to 16 hypotheses with one host solve per
// hypothesis and iteration; here the correspondence and accumulation are odom.py's (closest voxel mean in the
// 27 neighbours, point-to-distribution residual, odom.py:275-375) and the step is localize.py's
// (A = H + diag(1e-3 diag H + 1e-9), solve, left twist; skipped below 30 matches; localize.py:531-547).
//
// Work: (query, hypothesis) units, queries longest first; a workgroup takes a unit from an atomic queue and
// runs all its iterations alone: the lanes test the query scan's points against the map's neighbourhood bitmap
// (launch_nb_build) and queue those that can match; every lane then takes a queued point, searches the 27
// voxels around it (bitmap words of a chunk loaded together, then its slots), forms the float32 terms of
// H and b as the reference rounds them and adds them, widened, to float64 running sums; the workgroup
// reduces (wave shuffles, then LDS) and one lane solves the 6x6 step in float64. Units never wait on each
// other, so the grid needs no co-residency; the grid size, the block size and the chunk (neighbour keys a lane
// resolves together: 14, two chunks of the 27, or 7, four) are tuning parameters. A chunk's keys, hashes and
// slot records are live together with the 29 float64 running sums: 14 fills the register file at one wave
// per SIMD (and spills at 512 threads, where two are resident), 7 leaves room for two waves per SIMD.
#include "lw_internal.h"
#include "lw_loc.h"
#include "lw_search.h"

namespace lw {

namespace {

constexpr int LOC_ACC = 29;  // H (21), b (6), count, weight

struct LocArgs {
    const float *scans;
    const int32_t *scan_off;
    const int32_t *order;
    const double *hyp_T;
    double *out_T;
    double *out_score;
    int32_t *out_count;
    int32_t *out_iters;
    const Slot *slots;
    const uint32_t *occ;
    const uint32_t *nb_occ;  // the neighbourhood bitmap (launch_nb_build)
    int32_t nb_mask;         // its bits - 1
    const VoxPrec *vox_prec;
    uint32_t *queue;
    unsigned long long *prof;
    int32_t Q, H, mask, iters, min_count;
    float inv_v, max_d2;
    double damping;
    float kappa[16];
};

#define LW_H_PAIRS(X) \
    X(0, 0, 0) X(0, 1, 1) X(0, 2, 2) X(0, 3, 3) X(0, 4, 4) X(0, 5, 5) X(1, 1, 6) X(1, 2, 7) X(1, 3, 8) X(1, 4, 9) \
    X(1, 5, 10) X(2, 2, 11) X(2, 3, 12) X(2, 4, 13) X(2, 5, 14) X(3, 3, 15) X(3, 4, 16) X(3, 5, 17) X(4, 4, 18)   \
    X(4, 5, 19) X(5, 5, 20)

// localize.py:531-547 for one hypothesis: returns whether the pose moved by less than the batch tolerance.
__device__ inline bool loc_step(const double *v, double damping, int32_t min_count, double *T, double *moved) {
    *moved = 0.0;
    if (v[21 + 6] < (double)min_count) return false;  // too few matches: keep the pose (localize.py:535-536)
    double Hm[36], A[36], rhs[6], d[6];
    gn::sym6(v, Hm);
    LW_UNROLL for (int i = 0; i < 36; ++i) A[i] = Hm[i];
    LW_UNROLL for (int i = 0; i < 6; ++i) {
        A[7 * i] += damping * Hm[7 * i] + 1e-9;
        rhs[i] = -v[21 + i];
    }
    const bool all[6] = {true, true, true, true, true, true};
    if (!gn::solve6(A, rhs, all, d)) return false;  // LinAlgError: keep the pose (localize.py:541-543)
    gn::apply_twist(T, d);
    const double nr = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double nt = sqrt(d[3] * d[3] + d[4] * d[4] + d[5] * d[5]);
    *moved = nr > 5.0 * nt ? nr : 5.0 * nt;  // localize.py:545
    return true;
}

// The sum of the first k entries of c (k <= W), in order.
template <int W>
__device__ inline int32_t wave_prefix(const int32_t *c, int k) {
    int32_t t = 0;
    for (int w = 0; w < k; ++w) t += c[w];
    return t;
}

template <int B, int C>
__global__ __launch_bounds__(B) void k_loc(LocArgs a) {
    constexpr int W = B / 64;
    __shared__ double red[W][LOC_ACC + 1];
    __shared__ double vsum[32];
    __shared__ double Td[16];
    __shared__ float Tf[16];
    __shared__ int32_t s_unit, s_stop;
    __shared__ int32_t q_idx[2 * B];  // points waiting for their search (fewer than 2B at any time)
    __shared__ int32_t wcnt[W];
    __shared__ NbrOff nbr[NBR_ROWS];
    nbr_fill(nbr);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int32_t hits = 0, matched_total = 0;
    unsigned long long pt_iters = 0;
    const int32_t total = a.Q * a.H;
    for (;;) {
        if (threadIdx.x == 0) s_unit = (int32_t)atomicAdd(a.queue, 1u);
        __syncthreads();
        const int32_t u = s_unit;
        __syncthreads();
        if (u >= total) break;  // every workgroup leaves once the queue is drained
        const int32_t q = a.order[u / a.H];
        const int32_t h = u % a.H;
        const int32_t unit = q * a.H + h;
        const int32_t p0 = a.scan_off[q], n = a.scan_off[q + 1] - p0;
        if (threadIdx.x < 16) Td[threadIdx.x] = a.hyp_T[(size_t)unit * 16 + threadIdx.x];
        if (threadIdx.x == 0) s_stop = 0;
        __syncthreads();
        double score = 0.0, count = 0.0;
        int32_t iters = 0;
        for (int it = 0; it < a.iters; ++it) {
            if (threadIdx.x < 16) Tf[threadIdx.x] = (float)Td[threadIdx.x];
            __syncthreads();
            const float kappa = a.kappa[it < 16 ? it : 15];
            double acc[LOC_ACC];
            LW_UNROLL for (int k = 0; k < LOC_ACC; ++k) acc[k] = 0.0;
            // The workgroup's points that the neighbourhood bitmap passes are queued in LDS, B at a time in
            // point order (wave by wave, lanes in order), so every lane of a round searches a point that can
            // match. A round tests the next B points and appends; once B are queued (or every point is tested)
            // the queue's first B (or all) are searched, one a lane, and the rest move to the front.
            int32_t base = 0, qn = 0;
            for (;;) {
                if (base < n) {
                    const int32_t i = base + (int32_t)threadIdx.x;
                    bool act = false;
                    if (i < n) {
                        const float *pp = a.scans + 3 * (int64_t)(p0 + i);
                        float wx, wy, wz;
                        transform_point(Tf, pp[0], pp[1], pp[2], wx, wy, wz);
                        // a clear bit: no voxel of the map among the 27, so the search would find nothing
                        const int64_t k0 =
                            voxel_key_ijk(ifloor(wx * a.inv_v), ifloor(wy * a.inv_v), ifloor(wz * a.inv_v));
                        const int32_t h0 = slot_hash(k0) & a.nb_mask, h1 = nb_hash2(k0) & a.nb_mask;
                        act = (((a.nb_occ[h0 >> 5] >> (h0 & 31)) & (a.nb_occ[h1 >> 5] >> (h1 & 31))) & 1u) != 0u;
                    }
                    const uint64_t bal = __ballot(act);
                    if (lane == 0) wcnt[wave] = __popcll(bal);
                    __syncthreads();
                    if (act) q_idx[qn + wave_prefix<W>(wcnt, wave) + __popcll(bal & ((1ull << lane) - 1ull))] = i;
                    qn += wave_prefix<W>(wcnt, W);
                    base += B;
                    __syncthreads();
                }
                const bool flush = base >= n;
                if (qn == 0 && flush) break;
                if (qn < B && !flush) continue;
                const int32_t take = qn < B ? qn : B, rest = qn - take;
                const int32_t i = (int32_t)threadIdx.x < take ? q_idx[threadIdx.x] : -1;
                const int32_t moved = (int32_t)threadIdx.x < rest ? q_idx[take + threadIdx.x] : 0;
                __syncthreads();
                if ((int32_t)threadIdx.x < rest) q_idx[threadIdx.x] = moved;
                qn = rest;
                __syncthreads();  // the moved entries in place before the next round reads or appends
                if (i < 0) continue;
                const float *pp = a.scans + 3 * (int64_t)(p0 + i);
                float wx, wy, wz;
                transform_point(Tf, pp[0], pp[1], pp[2], wx, wy, wz);
                Cand best{a.max_d2, -1, 27, 0.0f, 0.0f, 0.0f};
                search_part<1, C, true>(Table{a.slots, a.occ, a.mask}, nbr, a.inv_v, wx, wy, wz, 0, 27, best, hits);
                const int32_t s = best.s;
                if (s < 0) continue;
                const float mx = best.mx, my = best.my, mz = best.mz;
                ++matched_total;
                const VoxPrec P = a.vox_prec[best.vid];
                const float M[3][3] = {{P.M[0], P.M[1], P.M[2]}, {P.M[1], P.M[3], P.M[4]}, {P.M[2], P.M[4], P.M[5]}};
                const float r0 = wx - mx, r1 = wy - my, r2 = wz - mz;
                float Mr[3];
                LW_UNROLL for (int k = 0; k < 3; ++k) {
                    float t = M[k][0] * r0;
                    t = t + M[k][1] * r1;
                    Mr[k] = t + M[k][2] * r2;
                }
                float d2 = r0 * Mr[0];
                d2 = d2 + r1 * Mr[1];
                d2 = d2 + r2 * Mr[2];
                float wgt = 1.0f / (1.0f + d2 / (kappa * kappa));
                wgt = wgt * wgt;
#define LW_LOC_H(p, q, k) acc[k] += (double)(wgt * hval<p, q>(M, wx, wy, wz));
                LW_H_PAIRS(LW_LOC_H)
#undef LW_LOC_H
                acc[21] += (double)(wgt * bval<0>(Mr, wx, wy, wz));
                acc[22] += (double)(wgt * bval<1>(Mr, wx, wy, wz));
                acc[23] += (double)(wgt * bval<2>(Mr, wx, wy, wz));
                acc[24] += (double)(wgt * bval<3>(Mr, wx, wy, wz));
                acc[25] += (double)(wgt * bval<4>(Mr, wx, wy, wz));
                acc[26] += (double)(wgt * bval<5>(Mr, wx, wy, wz));
                acc[27] += 1.0;
                acc[28] += (double)wgt;
            }
            pt_iters += (threadIdx.x == 0) ? (unsigned long long)n : 0ull;
            // reduce: wave shuffles, then one LDS row per wave summed in wave order
            LW_UNROLL for (int k = 0; k < LOC_ACC; ++k) {
                double x = acc[k];
                LW_UNROLL for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
                if (lane == 0) red[wave][k] = x;
            }
            __syncthreads();
            if (threadIdx.x < LOC_ACC) {
                double t = 0.0;
                LW_UNROLL for (int wv = 0; wv < W; ++wv) t += red[wv][threadIdx.x];
                vsum[threadIdx.x] = t;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                score = vsum[28];
                count = vsum[27];
                double moved;
                const bool stepped = loc_step(vsum, a.damping, a.min_count, Td, &moved);
                if (stepped) iters = it + 1;
                s_stop = stepped && moved < 2e-4 ? 1 : 0;  // converged (localize.py:548-549, per hypothesis)
            }
            __syncthreads();
            if (s_stop) break;
        }
        if (threadIdx.x < 16) a.out_T[(size_t)unit * 16 + threadIdx.x] = Td[threadIdx.x];
        if (threadIdx.x == 0) {
            a.out_score[unit] = score;
            a.out_count[unit] = (int32_t)count;
            a.out_iters[unit] = iters;
        }
        __syncthreads();
    }
    if (a.prof) {
        LW_UNROLL for (int off = 32; off > 0; off >>= 1) {
            hits += __shfl_xor(hits, off, 64);
            matched_total += __shfl_xor(matched_total, off, 64);
        }
        if (lane == 0) {
            atomicAdd(&a.prof[2], (unsigned long long)hits);
            atomicAdd(&a.prof[3], (unsigned long long)matched_total);
        }
        if (threadIdx.x == 0) {
            atomicAdd(&a.prof[0], pt_iters);
            atomicAdd(&a.prof[1], pt_iters * 27ull);
        }
    }
}

__global__ void k_queue_reset(uint32_t *q) { *q = 0u; }

}  // namespace

hipError_t loc_tile_attributes(int32_t block, int32_t chunk, hipFuncAttributes *fa) {
    const void *f = nullptr;
    switch (block * 100 + chunk) {
    case 6414: f = reinterpret_cast<const void *>(&k_loc<64, 14>); break;
    case 12814: f = reinterpret_cast<const void *>(&k_loc<128, 14>); break;
    case 25614: f = reinterpret_cast<const void *>(&k_loc<256, 14>); break;
    case 51214: f = reinterpret_cast<const void *>(&k_loc<512, 14>); break;
    case 6407: f = reinterpret_cast<const void *>(&k_loc<64, 7>); break;
    case 12807: f = reinterpret_cast<const void *>(&k_loc<128, 7>); break;
    case 25607: f = reinterpret_cast<const void *>(&k_loc<256, 7>); break;
    case 51207: f = reinterpret_cast<const void *>(&k_loc<512, 7>); break;
    default: return hipErrorInvalidValue;
    }
    return hipFuncGetAttributes(fa, f);
}

hipError_t launch_loc(lw_ctx *c, const LocLaunch &L) {
    LocArgs a{};
    a.scans = L.scans;
    a.scan_off = L.scan_off;
    a.order = L.order;
    a.hyp_T = L.hyp_T;
    a.out_T = L.out_T;
    a.out_score = L.out_score;
    a.out_count = L.out_count;
    a.out_iters = L.out_iters;
    a.slots = c->slots;
    a.occ = c->occ;
    a.nb_occ = c->nb_occ;
    a.nb_mask = (int32_t)(NB_SCALE * c->cap - 1);
    a.vox_prec = c->vox_prec;
    a.queue = c->loc_queue;
    a.prof = L.prof;
    a.Q = L.Q;
    a.H = L.H;
    a.mask = c->mask;
    a.iters = L.iters;
    a.min_count = L.min_count;
    a.inv_v = L.inv_v;
    a.max_d2 = L.max_d2;
    a.damping = L.damping;
    for (int k = 0; k < 16; ++k) a.kappa[k] = L.kappa[k];
    hipLaunchKernelGGL(k_queue_reset, dim3(1), dim3(1), 0, c->stream, c->loc_queue);
    const dim3 grid((unsigned)(L.grid > 0 ? L.grid : 1));
    switch (L.block * 100 + L.chunk) {
    case 6414: hipLaunchKernelGGL((k_loc<64, 14>), grid, dim3(64), 0, c->stream, a); break;
    case 12814: hipLaunchKernelGGL((k_loc<128, 14>), grid, dim3(128), 0, c->stream, a); break;
    case 25614: hipLaunchKernelGGL((k_loc<256, 14>), grid, dim3(256), 0, c->stream, a); break;
    case 51214: hipLaunchKernelGGL((k_loc<512, 14>), grid, dim3(512), 0, c->stream, a); break;
    case 6407: hipLaunchKernelGGL((k_loc<64, 7>), grid, dim3(64), 0, c->stream, a); break;
    case 12807: hipLaunchKernelGGL((k_loc<128, 7>), grid, dim3(128), 0, c->stream, a); break;
    case 25607: hipLaunchKernelGGL((k_loc<256, 7>), grid, dim3(256), 0, c->stream, a); break;
    case 51207: hipLaunchKernelGGL((k_loc<512, 7>), grid, dim3(512), 0, c->stream, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace lw
