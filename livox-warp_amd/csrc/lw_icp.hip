// Scan-to-map accumulation and the Gauss-Newton loop on the device: k_icp + k_reduce
// (python/livox_warp/odom.py:275-384) and the host step of Odometry._register (odom.py:752-803).
//
// One pass: each thinned point gets S lanes (S = 2 for the 27-voxel search, 4 for the 125-voxel one; the
// persistent loop gives four at both); the lanes split the point's neighbours, in the reference's order (dx
// outer, dz inner; odom.py:315-332), into contiguous parts. A lane searches its part in chunks (lw_search.h):
// the chunk's home slots (32 B records that carry key, voxel id, count and mean) are loaded together, and the
// keys whose home slot holds another key continue their probes together, so a chunk costs one dependent
// memory round trip plus one per probe depth, and a lane one or two chunks. The lanes then merge their
// closest means by (distance, position in the order), which is exactly the serial loop's choice. The point's
// first two lanes form the 53 float32 terms of the ACC layout as the reference rounds them (odom.py:246-259,
// 269-271, 372-375).
//
// Reduction: the wave's terms are transposed through LDS and summed per entry in lane order; each
// workgroup writes one partial row; the rows are summed in batches of 32 (each batch in row order, then the
// batch sums in batch order), and in a registration the float64 step (lw_gn.h, on the whole wave when it
// applies) follows.
//
// Two drivers:
//   k_icp       one pass per launch (lw_icp_accum, and the registration fallback): the last workgroup to
//               finish (atomic ticket) reduces and steps;
//   k_register  the whole loop in one cooperative launch: a leader workgroup reduces the rows as they come
//               in, runs the step and publishes the pose; the searching workgroups wait for it (bounded
//               spin; co-residency is what the cooperative launch guarantees), so a registration is one
//               launch with no idle passes after convergence and reads the thinned count on the device.
//
// Accumulation forms (acc_mode): 0 float64 (the reference), 1 compensated float32 (Neumaier sums, two-sum
// merges), 2 naive float32.
#include "lw_gn_wave.h"
#include "lw_internal.h"
#include "lw_search.h"

namespace lw {

namespace {

__device__ inline void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

constexpr int RED_STRIDE = ACC_USED + 1;  // LDS row stride (odd, so column reads do not share banks)
constexpr int RED_ROWS = ICP_BLOCK / 2;   // one row per point of a lane group: 64 / S with S >= 2 lanes a point
constexpr int RED_BYTES = RED_ROWS * RED_STRIDE * 8;

// Each point of a lane group owns one LDS row of 53 running sums (and compensations for the compensated
// form): row lane / S, written by the point's first two lanes, to which a point's terms are added as soon as
// they are formed, so no accumulator stays in registers across the neighbour search.
template <int MODE, int S>
__device__ inline void row_zero(unsigned char *lds) {
    const int lane = lane_id();
    if (lane >= ICP_BLOCK / S) return;
    if constexpr (MODE == 0) {
        double *red = reinterpret_cast<double *>(lds);
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) red[lane * RED_STRIDE + k] = 0.0;
    } else {
        float *rs = reinterpret_cast<float *>(lds);
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) rs[lane * RED_STRIDE + k] = 0.0f;
        if constexpr (MODE == 1) {
            float *rc = rs + RED_ROWS * RED_STRIDE;
            LW_UNROLL for (int k = 0; k < ACC_USED; ++k) rc[lane * RED_STRIDE + k] = 0.0f;
        }
    }
}

// One term x of a point into entry k of its row: a float64 sum of the float32 terms (the reference's
// accumulation), or a compensated / naive float32 sum.
template <int MODE>
__device__ inline void row_put(unsigned char *lds, int row, int k, float x) {
    if constexpr (MODE == 0) {
        double *red = reinterpret_cast<double *>(lds);
        red[row * RED_STRIDE + k] += (double)x;
    } else if constexpr (MODE == 2) {
        float *rs = reinterpret_cast<float *>(lds);
        rs[row * RED_STRIDE + k] += x;
    } else {
        float *rs = reinterpret_cast<float *>(lds);
        float *rc = rs + RED_ROWS * RED_STRIDE;
        float t, e;
        two_sum(rs[row * RED_STRIDE + k], x, t, e);
        rs[row * RED_STRIDE + k] = t;
        rc[row * RED_STRIDE + k] += e;
    }
}

// Sum the wave's rows per entry, in row (= lane) order. Lane k < 53 returns entry k (sum, and compensation for
// the compensated form).
template <int MODE, int S>
__device__ inline void wave_sum(unsigned char *lds, double &sum, double &comp) {
    const int lane = lane_id();
    sum = 0.0;
    comp = 0.0;
    __syncthreads();
    if constexpr (MODE == 0) {
        double *red = reinterpret_cast<double *>(lds);
        if (lane < ACC_USED) {
            double t = 0.0;
            for (int r = 0; r < ICP_BLOCK / S; ++r) t += red[r * RED_STRIDE + lane];
            sum = t;
        }
    } else if constexpr (MODE == 2) {
        float *red = reinterpret_cast<float *>(lds);
        if (lane < ACC_USED) {
            float t = 0.0f;
            for (int r = 0; r < ICP_BLOCK / S; ++r) t += red[r * RED_STRIDE + lane];
            sum = (double)t;
        }
    } else {
        float *rs = reinterpret_cast<float *>(lds);
        float *rc = rs + RED_ROWS * RED_STRIDE;
        if (lane < ACC_USED) {
            float s = 0.0f, c = 0.0f;
            for (int r = 0; r < ICP_BLOCK / S; ++r) {
                float t, e;
                two_sum(s, rs[r * RED_STRIDE + lane], t, e);
                s = t;
                c = c + rc[r * RED_STRIDE + lane] + e;
            }
            sum = (double)s;
            comp = (double)c;
        }
    }
    __syncthreads();  // the LDS rows may be rewritten by the next pass
}

struct IcpArgs {
    const float *ds;
    const Slot *slots;
    const uint32_t *occ;
    const VoxPrec *vox_prec;
    RegState *st;
    double *partials;  // [G][128]: sums, compensations
    int32_t *best_slot;
    unsigned long long *prof;   // measurement counters (lw_ctx::prof_dev), or null
    unsigned long long *phase;  // the same counters when the phase clock stamps are on, or null
    int32_t n;
    int32_t mask;
    float inv_v, voxel, max_d2, kappa;
    int32_t it;
    int32_t gn;
    gn::StepParams sp;
};

// What one workgroup hands to another within a launch -- partial rows, the step, the done flag, the ticket --
// moves through device-coherent accesses (agent scope: they bypass this XCD's non-coherent L2 copy), and a
// writer releases before it raises the flag (ticket, generation) the reader waits on. No pass invalidates the
// L2, so the map the search reads stays cached there.
template <class T>
__device__ inline void put(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class T>
__device__ inline T get(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A writer's release before it raises a flag: the agent-scope release writes back this XCD's dirty L2 lines
// (few: the shared data goes through device-coherent stores) and waits for the stores. Readers need no
// acquire: they read the shared data with device-coherent loads, so nothing is invalidated.
__device__ inline void stores_done() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

// What changes from pass to pass (kept out of IcpArgs so the persistent loop never copies the arguments).
struct Pass {
    int32_t n;
    float max_d2, kappa;
    int32_t it;
};

#define LW_H_PAIRS(X) \
    X(0, 0, 0) X(0, 1, 1) X(0, 2, 2) X(0, 3, 3) X(0, 4, 4) X(0, 5, 5) X(1, 1, 6) X(1, 2, 7) X(1, 3, 8) X(1, 4, 9) \
    X(1, 5, 10) X(2, 2, 11) X(2, 3, 12) X(2, 4, 13) X(2, 5, 14) X(3, 3, 15) X(3, 4, 16) X(3, 5, 17) X(4, 4, 18)   \
    X(4, 5, 19) X(5, 5, 20)

// Lanes per point and chunk size per reach: 27 neighbours as 14 + 13 on two lanes; 125 as 32+32+32+29 on
// four lanes, each in two chunks of 16.
template <int R>
struct Split;
template <>
struct Split<1> {
    static constexpr int S = 2, C = 14;
};
template <>
struct Split<2> {
    static constexpr int S = 4, C = 16;
};
static_assert(Split<1>::S >= 2 && Split<2>::S >= 2, "at most RED_ROWS points a lane group");

// The reference's terms of one matched point (odom.py:335-375) into its row (row = lane / S), formed by the
// point's lanes together with the same code on different data: sub-lane 0 adds w J^T M J, sub-lane 1 w J^T Mg J
// (the same products on the other matrix), and sub-lane 0 then the vector and scalar entries. Each entry's
// value is the reference's float32 expression, and each row entry receives the points' terms in point order.
// P: the matched voxel's precision.
template <int MODE, int S>
__device__ inline void point_terms(const VoxPrec &P, float kappa, float px, float py, float pz, float wx, float wy,
                                   float wz, const Cand &b, unsigned char *lds) {
    static_assert(S >= 2, "two lanes a point at least");
    const int lane = lane_id(), row = lane / S, sub = lane % S;
    const float r0 = wx - b.mx, r1 = wy - b.my, r2 = wz - b.mz;
    const float M[3][3] = {{P.M[0], P.M[1], P.M[2]}, {P.M[1], P.M[3], P.M[4]}, {P.M[2], P.M[4], P.M[5]}};
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
    wgt = wgt * wgt;  // Geman-McClure (odom.py:370-371)
    if (sub < 2) {
        const bool g = sub == 1;
        const float X[3][3] = {{g ? P.Mg[0] : P.M[0], g ? P.Mg[1] : P.M[1], g ? P.Mg[2] : P.M[2]},
                               {g ? P.Mg[1] : P.M[1], g ? P.Mg[3] : P.M[3], g ? P.Mg[4] : P.M[4]},
                               {g ? P.Mg[2] : P.M[2], g ? P.Mg[4] : P.M[4], g ? P.Mg[5] : P.M[5]}};
        const int base = g ? ACC_HG : ACC_H;
#define LW_PUT_H(p, q, k) row_put<MODE>(lds, row, base + k, wgt * hval<p, q>(X, wx, wy, wz));
        LW_H_PAIRS(LW_PUT_H)
#undef LW_PUT_H
    }
    if (sub == 0) {
        float rn;
        if (P.shape > 0.5f) {
            rn = r0 * P.q[0];
            rn = rn + r1 * P.q[1];
            rn = rn + r2 * P.q[2];
        } else {
            rn = length3(r0, r1, r2);
        }
        row_put<MODE>(lds, row, ACC_RMS, wgt * rn * rn);
        row_put<MODE>(lds, row, ACC_B + 0, wgt * bval<0>(Mr, wx, wy, wz));
        row_put<MODE>(lds, row, ACC_B + 1, wgt * bval<1>(Mr, wx, wy, wz));
        row_put<MODE>(lds, row, ACC_B + 2, wgt * bval<2>(Mr, wx, wy, wz));
        row_put<MODE>(lds, row, ACC_B + 3, wgt * bval<3>(Mr, wx, wy, wz));
        row_put<MODE>(lds, row, ACC_B + 4, wgt * bval<4>(Mr, wx, wy, wz));
        row_put<MODE>(lds, row, ACC_B + 5, wgt * bval<5>(Mr, wx, wy, wz));
        row_put<MODE>(lds, row, ACC_COUNT, 1.0f);
        row_put<MODE>(lds, row, ACC_ERR, wgt * d2);
        row_put<MODE>(lds, row, ACC_WEIGHT, wgt);
        row_put<MODE>(lds, row, ACC_RANGE, length3(px, py, pz));
    }
}

// ---- the persistent loop's neighbourhood record --------------------------------------------------------
//
// Within a registration the map does not change. In the persistent loop a searcher keeps the points of its
// first lane group in registers from pass to pass, and both reaches give a point four lanes (Split<2>::S, and
// the loop's narrow split below), so the searcher serves the same points in every pass. Each wide pass (the
// 5x5x5 search) records, for every key of each point's 125, what the search found there -- the mean and the
// voxel id when the key holds a positive count -- at the key's position in the 5x5x5 order, and the point's
// voxel beside it. A narrow pass whose point lies within one voxel of that centre on every axis has its 27
// neighbours inside the recorded 125: each lane reads the entries of its part of the 27 at their positions,
// all in flight together, and takes the closest with the search's float32 distance and (distance, position)
// rule, so the closest mean and the hit count are the search's; only the round trips to the map go. A point
// further away (or a pass with no record) searches. The loop writes no per-point slots, so a taken entry marks
// its match with the voxel id in place of the slot (any non-negative value marks a match). Each point's first
// lane also keeps the precision record of the voxel it matched last, for the terms.
template <int R>
struct LoopSplit;  // the persistent loop's lanes per point and chunk size: four lanes at either reach
template <>
struct LoopSplit<1> {
    static constexpr int S = 4, C = 7;  // 27 as 7 + 7 + 7 + 6
};
template <>
struct LoopSplit<2> {
    static constexpr int S = Split<2>::S, C = Split<2>::C;
};
static_assert(LoopSplit<1>::S == LoopSplit<2>::S, "both reaches serve the same points");
constexpr int LS = LoopSplit<2>::S;
constexpr int CROWS = ICP_BLOCK / LS;  // points of a lane group
constexpr int REC_NONE = -1;           // entry w: not in the map
constexpr int REC_EMPTY = -2;          // entry w: in the map with no positive count (a hit, never a match)
struct NbrCache {
    float4 e[CROWS][128];         // [point][position in the 5x5x5 order]: mean x, y, z; w (voxel id or REC_*)
    int32_t cx[CROWS], cy[CROWS], cz[CROWS];  // the voxel the record was made around
    int32_t valid[CROWS];         // 1 once a wide pass has recorded the point
    float4 P[4][CROWS];           // a point's last matched precision record (VoxPrec, 64 B)
    int32_t pvid[CROWS];          // its voxel id; -1: none
};

// Every lane: no record, no precision record (a new registration).
__device__ inline void cache_clear(NbrCache *c) {
    const int lane = lane_id();
    if (lane < CROWS) {
        c->valid[lane] = 0;
        c->pvid[lane] = -1;
    }
}

// The wide search's recorder for the point of lane-group row `row`: every key at its position, what the search
// found there (a key not found at its home is written as absent, and again if a probe finds it).
struct CacheRec {
    NbrCache *c;
    int row;
    __device__ void add(bool ok, int32_t, int32_t j, const int4 &lo, const float4 &hi) const {
        const int w = ok ? (__int_as_float(lo.w) > 0.0f ? lo.z : REC_EMPTY) : REC_NONE;
        c->e[row][j] = make_float4(hi.x, hi.y, hi.z, __int_as_float(w));
    }
};

// This lane's part of a wide pass: the search, recording every key of its part; the point's voxel beside it.
template <int C>
__device__ inline void search_recorded(NbrCache *c, const Table &t, const NbrOff *tab, float inv_v, float wx,
                                       float wy, float wz, int j_lo, int j_hi, Cand &best, int32_t &hits) {
    const int lane = lane_id(), row = lane / LS;
    search_part<2, C, false, CacheRec>(t, tab, inv_v, wx, wy, wz, j_lo, j_hi, best, hits, CacheRec{c, row});
    if (lane % LS == 0) {
        c->cx[row] = ifloor(wx * inv_v);
        c->cy[row] = ifloor(wy * inv_v);
        c->cz[row] = ifloor(wz * inv_v);
        c->valid[row] = 1;
    }
}

// This lane's part [j_lo, j_hi) of a narrow pass (at most C keys): from the record when the point's 27
// neighbours lie inside the recorded 125, otherwise by the search.
template <int C>
__device__ inline void search_narrow(const NbrCache *c, const Table &t, const NbrOff *tab, float inv_v, float wx,
                                     float wy, float wz, int j_lo, int j_hi, Cand &best, int32_t &hits) {
    const int row = lane_id() / LS;
    const int32_t dx = ifloor(wx * inv_v) - c->cx[row], dy = ifloor(wy * inv_v) - c->cy[row],
                  dz = ifloor(wz * inv_v) - c->cz[row];
    if (c->valid[row] && dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) {
        float4 m[C];
        LW_UNROLL for (int k = 0; k < C; ++k) {
            const int j = min(j_lo + k, 26);  // a padded key reads a valid entry and is ignored below
            const int ox = j / 9 - 1 + dx, oy = (j / 3) % 3 - 1 + dy, oz = j % 3 - 1 + dz;
            m[k] = c->e[row][(ox + 2) * 25 + (oy + 2) * 5 + (oz + 2)];
        }
        LW_UNROLL for (int k = 0; k < C; ++k) {
            const int w = __float_as_int(m[k].w);
            const bool in = j_lo + k < j_hi;
            hits += in && w != REC_NONE ? 1 : 0;
            take_if_closer(in && w >= 0, dist2(wx, wy, wz, m[k].x, m[k].y, m[k].z), w, j_lo + k, m[k].x, m[k].y,
                           m[k].z, w, best);
        }
        return;
    }
    search_part<1, C, false>(t, tab, inv_v, wx, wy, wz, j_lo, j_hi, best, hits);
}

// The precision record of voxel vid for the point of lane-group row `row`: the kept one when it is that voxel's.
__device__ inline VoxPrec cached_prec(NbrCache *c, const VoxPrec *vox_prec, int row, int32_t vid) {
    VoxPrec P;
    float4 *q = reinterpret_cast<float4 *>(&P);
    if (c->pvid[row] == vid) {
        LW_UNROLL for (int k = 0; k < 4; ++k) q[k] = c->P[k][row];
    } else {
        P = vox_prec[vid];
        if (lane_id() % LS == 0) {
            LW_UNROLL for (int k = 0; k < 4; ++k) c->P[k][row] = q[k];
            c->pvid[row] = vid;
        }
    }
    return P;
}

// One workgroup's share of a pass: the points of lane groups g, g + G, ... (one group per workgroup whenever
// the grid is large enough). Lane k < 53 returns the workgroup's entry k. pre: when has_pre, this lane's point
// of group g, already loaded (the persistent loop keeps it in registers from pass to pass). cache: when given
// (reach 1 only), group g's candidate lists. With stamp set, t_found and t_terms receive the device clock after
// the last lane group's neighbour search and after its terms (measurement).
template <int MODE, int R, int S = Split<R>::S, int C = Split<R>::C>
__device__ inline void wave_pass(const IcpArgs &a, const Pass p, const float *Tf, const NbrOff *nbr, int g, int G,
                                 unsigned char *lds, double &sum, double &comp, int32_t &hits, int32_t &matched, float3 pre,
                                 bool has_pre, bool stamp, unsigned long long &t_found,
                                 unsigned long long &t_terms, NbrCache *cache = nullptr) {
    static_assert(S >= 2 && S <= 64 && 64 % S == 0, "whole lane groups of at most RED_ROWS points");
    constexpr int D = 2 * R + 1, K = D * D * D, KQ = (K + S - 1) / S;
    const int lane = lane_id();
    const int sub = lane % S;
    row_zero<MODE, S>(lds);
    for (int64_t base = (int64_t)g * ICP_BLOCK; base < (int64_t)p.n * S; base += (int64_t)G * ICP_BLOCK) {
        const int32_t i = (int32_t)((base + lane) / S);
        const bool first = base == (int64_t)g * ICP_BLOCK;
        const bool cached = S == LS && cache != nullptr && has_pre && first;
        float px = 0.0f, py = 0.0f, pz = 0.0f, wx = 0.0f, wy = 0.0f, wz = 0.0f;
        Cand best{p.max_d2, -1, K, 0.0f, 0.0f, 0.0f};
        if (i < p.n) {
            if (has_pre && first) {
                px = pre.x;
                py = pre.y;
                pz = pre.z;
            } else {
                px = a.ds[3 * (int64_t)i];
                py = a.ds[3 * (int64_t)i + 1];
                pz = a.ds[3 * (int64_t)i + 2];
            }
            transform_point(Tf, px, py, pz, wx, wy, wz);
            const Table t{a.slots, a.occ, a.mask};
            const int j_lo = sub * KQ, j_hi = min(K, (sub + 1) * KQ);
            if (!cached)
                search_part<R, C, false>(t, nbr, a.inv_v, wx, wy, wz, j_lo, j_hi, best, hits);
            else if constexpr (R == 2)
                search_recorded<C>(cache, t, nbr, a.inv_v, wx, wy, wz, j_lo, j_hi, best, hits);
            else
                search_narrow<C>(cache, t, nbr, a.inv_v, wx, wy, wz, j_lo, j_hi, best, hits);
        }
        if (stamp) t_found = wall_clock64();
        // merge the S parts: smaller distance wins, a tie goes to the earlier neighbour (the serial strict <)
        LW_UNROLL for (int off = 1; off < S; off <<= 1) {
            Cand o;
            o.d2 = __shfl_xor(best.d2, off, 64);
            o.s = __shfl_xor(best.s, off, 64);
            o.j = __shfl_xor(best.j, off, 64);
            o.mx = __shfl_xor(best.mx, off, 64);
            o.my = __shfl_xor(best.my, off, 64);
            o.mz = __shfl_xor(best.mz, off, 64);
            o.vid = __shfl_xor(best.vid, off, 64);
            if (o.s >= 0 && (best.s < 0 || o.d2 < best.d2 || (o.d2 == best.d2 && o.j < best.j))) best = o;
        }
        if (i < p.n) {  // every lane of the point holds the merged choice
            if (sub == 0 && a.best_slot) a.best_slot[i] = best.s;
            if (best.s >= 0) {
                matched += sub == 0 ? 1 : 0;
                const VoxPrec P = cached ? cached_prec(cache, a.vox_prec, lane / S, best.vid) : a.vox_prec[best.vid];
                point_terms<MODE, S>(P, p.kappa, px, py, pz, wx, wy, wz, best, lds);
            }
        }
        if (stamp) t_terms = wall_clock64();
    }
    wave_sum<MODE, S>(lds, sum, comp);
}

// ---- the cross-workgroup reduction ----------------------------------------------------------------
// The G partial rows are summed in batches of RED_BATCH rows, each batch in row order from zero, and the
// batch sums then in batch order. Both drivers use this order (the one-pass launch's last workgroup reads the
// batches one after another, the persistent loop's leader each batch as soon as its rows are in), so the two
// give the same sums bit for bit. Lane k < RED_COLS handles entry k: the 53 ACC entries, then the
// measurement counts (hits, matched points) in 53 and 54. Rows are read device-coherently.
constexpr int RED_COLS = ACC_USED + 2;
constexpr int RED_BATCH = 16;
constexpr int MAX_BATCHES = MAX_LOOP_ROWS / RED_BATCH;
static_assert(LOOP_TAGGED_ROWS == MAX_LOOP_ROWS + MAX_BATCHES, "the tagged rows, then one batch row per batch");
static_assert(MAX_BATCHES <= 32, "one bit per batch");

struct BSum {
    double e;    // float64 sum: the float64 form's entries, and every form's counts (integers: exact)
    float s, c;  // float32 sum and compensation of the float32 forms
};

// Row r (r < G), entry k, as the sum adds it; rows at or past G add nothing.
template <int MODE>
__device__ inline void add_row(BSum &r, bool ok, double x, double y) {
    if constexpr (MODE == 0) {
        r.e = ok ? r.e + x : r.e;
    } else if constexpr (MODE == 2) {
        r.s = ok ? r.s + (float)x : r.s;
        r.e = ok ? r.e + x : r.e;
    } else {
        float t, er;
        two_sum(r.s, (float)x, t, er);
        r.s = ok ? t : r.s;
        r.c = ok ? r.c + (float)y + er : r.c;
        r.e = ok ? r.e + x : r.e;
    }
}

// The sums of NB batches b[0..NB) (each RED_BATCH rows from row b[k] * RED_BATCH, rows at or past G adding
// nothing) of this lane's column, in row order from zero. Every load is issued before any is used (addresses
// clamped to row G - 1, so none is conditional).
template <int MODE, int NB>
__device__ inline void sum_batches(const double *partials, int G, const int (&b)[NB], BSum (&out)[NB]) {
    const int lane = threadIdx.x;
    const int cx = lane < RED_COLS ? lane : 0, cy = 64 + (lane < ACC_USED ? lane : 0);
    double x[NB][RED_BATCH], y[NB][RED_BATCH];
    LW_UNROLL for (int k = 0; k < NB; ++k)
        LW_UNROLL for (int u = 0; u < RED_BATCH; ++u) {
            const size_t row = (size_t)min(b[k] * RED_BATCH + u, G - 1) * 128;
            x[k][u] = get(partials + row + cx);
            if constexpr (MODE == 1) y[k][u] = get(partials + row + cy);
        }
    LW_UNROLL for (int k = 0; k < NB; ++k) {
        BSum r{0.0, 0.0f, 0.0f};
        LW_UNROLL for (int u = 0; u < RED_BATCH; ++u)
            add_row<MODE>(r, b[k] * RED_BATCH + u < G, x[k][u], MODE == 1 ? y[k][u] : 0.0);
        out[k] = r;
    }
}

template <int MODE>
__device__ inline void combine(BSum &acc, const BSum &b) {
    if constexpr (MODE == 1) {
        float t, er;
        two_sum(acc.s, b.s, t, er);
        acc.s = t;
        acc.c = acc.c + b.c + er;
    } else if constexpr (MODE == 2) {
        acc.s += b.s;
    }
    acc.e += b.e;
}

template <int MODE>
__device__ inline double entry_of(const BSum &r) {
    if (threadIdx.x >= RED_COLS) return 0.0;
    if constexpr (MODE == 0) return r.e;
    else if constexpr (MODE == 2) return threadIdx.x < ACC_USED ? (double)r.s : r.e;
    else return threadIdx.x < ACC_USED ? (double)r.s + (double)r.c : r.e;
}

// Every row, batch after batch (the one-pass launch's last workgroup, after its ticket).
template <int MODE>
__device__ inline double reduce_rows(const double *partials, int G) {
    BSum acc{0.0, 0.0f, 0.0f};
    for (int b = 0; b < (G + RED_BATCH - 1) / RED_BATCH; ++b) {
        const int bb[1] = {b};
        BSum out[1];
        sum_batches<MODE, 1>(partials, G, bb, out);
        combine<MODE>(acc, out[0]);
    }
    return entry_of<MODE>(acc);
}

// ---- the persistent loop's tagged rows ---------------------------------------------------------------------
//
// The one-pass launch's rows are plain doubles behind a release and a ticket. The persistent loop's rows are
// tagged instead: every value as two 64-bit words, (tag << 32) | one half of its bits, each stored and loaded
// as one single-copy-atomic access. A word is therefore current exactly when it carries the pass's tag, so a
// writer needs no release and no flag, and a reader takes a row as soon as every word it reads carries the
// tag. Row g: words
// 2k and 2k + 1 hold entry k (the 53 sums, then the two measurement counts), and in the compensated form words
// 128 + 2k and 129 + 2k the compensation of entry k. The tag is (epoch << 7) | (pass + 1): the epoch advances
// at every registration start (reg_state_begin) and is cleared with the rows before it could wrap, so no word
// left from an earlier pass or registration carries a current tag.
__device__ inline uint32_t row_tag(uint32_t epoch, int it) { return (epoch << 7) | (uint32_t)(it + 1); }

__device__ inline void put_tagged(uint64_t *w, uint32_t tag, double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v), t = (uint64_t)tag << 32;
    put(&w[0], (uint64_t)(t | (b & 0xffffffffu)));
    put(&w[1], (uint64_t)(t | (b >> 32)));
}

__device__ inline bool has_tag(uint64_t w, uint32_t tag) { return (uint32_t)(w >> 32) == tag; }

__device__ inline double join_tagged(uint64_t lo, uint64_t hi) {
    return __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffu)));
}

template <int MODE>
__device__ inline void write_row_tagged(uint64_t *rows, int g, uint32_t tag, double sum, double comp, int32_t hits,
                                        int32_t matched) {
    const int lane = threadIdx.x;
    uint64_t *row = rows + (size_t)g * LOOP_ROW_WORDS;
    LW_UNROLL for (int off = 32; off > 0; off >>= 1) {
        hits += __shfl_xor(hits, off, 64);
        matched += __shfl_xor(matched, off, 64);
    }
    if (lane < RED_COLS) {
        const double v = lane < ACC_USED ? sum : (lane == ACC_USED ? (double)hits : (double)matched);
        put_tagged(&row[2 * lane], tag, v);
        if constexpr (MODE == 1)
            if (lane < ACC_USED) put_tagged(&row[128 + 2 * lane], tag, comp);
    }
}

// The sums of batches b[0..NB) of the tagged rows, formed exactly as sum_batches forms them (row order from
// zero, rows at or past G adding nothing), and per batch whether every word it needs carried the tag (if not,
// a row is still in flight and the batch is read again later). Every load is issued before any is used.
template <int MODE, int NB>
__device__ inline void sum_tagged(const uint64_t *rows, int G, uint32_t tag, const int (&b)[NB], BSum (&out)[NB],
                                  bool (&ok)[NB]) {
    constexpr int W = MODE == 1 ? 4 : 2;
    const int lane = threadIdx.x;
    const int cx = lane < RED_COLS ? lane : 0, cy = lane < ACC_USED ? lane : 0;
    uint64_t w[NB][RED_BATCH][W];
    LW_UNROLL for (int k = 0; k < NB; ++k)
        LW_UNROLL for (int u = 0; u < RED_BATCH; ++u) {
            const uint64_t *row = rows + (size_t)min(b[k] * RED_BATCH + u, G - 1) * LOOP_ROW_WORDS;
            w[k][u][0] = get(row + 2 * cx);
            w[k][u][1] = get(row + 2 * cx + 1);
            if constexpr (MODE == 1) {
                w[k][u][2] = get(row + 128 + 2 * cy);
                w[k][u][3] = get(row + 129 + 2 * cy);
            }
        }
    LW_UNROLL for (int k = 0; k < NB; ++k) {
        BSum r{0.0, 0.0f, 0.0f};
        bool good = true;
        LW_UNROLL for (int u = 0; u < RED_BATCH; ++u) {
            const bool in = b[k] * RED_BATCH + u < G;
            bool cur = has_tag(w[k][u][0], tag) && has_tag(w[k][u][1], tag);
            double y = 0.0;
            if constexpr (MODE == 1) {
                cur = cur && has_tag(w[k][u][2], tag) && has_tag(w[k][u][3], tag);
                y = join_tagged(w[k][u][2], w[k][u][3]);
            }
            good = good && (cur || !in);
            add_row<MODE>(r, in, join_tagged(w[k][u][0], w[k][u][1]), y);
        }
        ok[k] = __ballot(!good) == 0ull;
        out[k] = r;
    }
}

constexpr uint32_t SPIN_LIMIT = 1u << 24;  // ~ seconds of polling: a wait that long means the loop is broken
constexpr int32_t STATUS_SPIN_TIMEOUT = -1;

// ---- the persistent loop's two-level reduction ------------------------------------------------------
//
// Batch b of a pass is rows [16b, 16b + 16) of [0, G). Its head, searcher 16b, waits for the batch's tagged
// rows, sums them in row order from zero (sum_tagged: sum_batches' arithmetic) and writes the sum as tagged
// batch row b; the leader waits for the batch rows and combines them in batch order. That is reduce_rows'
// order, so the one-pass launch and the persistent loop give the same sums bit for bit. Waiting on 16 rows
// at a time (the heads, side by side) and then on at most 16 batch rows (the leader, in one poll) takes the
// last row about two round trips after it is written.

// Batch row words: entry k (lane k) at words 4k..4k+3: the float64 sum's two halves, then (forms 1 and 2)
// the float32 sum and (form 1) its compensation.
template <int MODE>
__device__ inline void write_batch_tagged(uint64_t *row, uint32_t tag, const BSum &r) {
    const int lane = threadIdx.x;
    if (lane >= RED_COLS) return;
    put_tagged(&row[4 * lane], tag, r.e);
    const uint64_t t = (uint64_t)tag << 32;
    if constexpr (MODE != 0) put(&row[4 * lane + 2], (uint64_t)(t | __float_as_uint(r.s)));
    if constexpr (MODE == 1) put(&row[4 * lane + 3], (uint64_t)(t | __float_as_uint(r.c)));
}

// Batch b's head: waits until every word of the batch's rows carries the tag, then writes their sum as batch
// row b (brows + b * LOOP_ROW_WORDS). False after SPIN_LIMIT polls that found a row missing.
template <int MODE>
__device__ inline bool batch_head(const uint64_t *rows, uint64_t *brows, int G, int b, uint32_t tag) {
    const int bb[1] = {b};
    for (uint32_t spins = 0; spins <= SPIN_LIMIT; ++spins) {
        BSum out[1];
        bool ok[1];
        sum_tagged<MODE, 1>(rows, G, tag, bb, out, ok);
        if (ok[0]) {
            write_batch_tagged<MODE>(brows + (size_t)b * LOOP_ROW_WORDS, tag, out[0]);
            return true;
        }
    }
    return false;
}

// The leader: polls the ceil(G / 16) batch rows (every load in flight at once) until every word it needs
// carries the tag, then combines the batch sums in batch order into this lane's entry v. t_in: the device
// clock when the last batch row was seen (measurement). False after SPIN_LIMIT polls that found a batch
// missing.
template <int MODE>
__device__ inline bool leader_reduce(const uint64_t *brows, int G, uint32_t tag, double &v, unsigned long long &t_in,
                                     bool stamp) {
    constexpr int W = MODE == 0 ? 2 : MODE == 2 ? 3 : 4;
    const int lane = threadIdx.x;
    const int cx = lane < RED_COLS ? lane : 0;
    const int nb = (G + RED_BATCH - 1) / RED_BATCH;
    for (uint32_t spins = 0; spins <= SPIN_LIMIT; ++spins) {
        uint64_t w[MAX_BATCHES][W];
        LW_UNROLL for (int b = 0; b < MAX_BATCHES; ++b) {
            const uint64_t *row = brows + (size_t)min(b, nb - 1) * LOOP_ROW_WORDS + 4 * cx;
            LW_UNROLL for (int q = 0; q < W; ++q) w[b][q] = get(row + q);
        }
        bool good = true;
        LW_UNROLL for (int b = 0; b < MAX_BATCHES; ++b)
            LW_UNROLL for (int q = 0; q < W; ++q) good = good && (b >= nb || has_tag(w[b][q], tag));
        if (__ballot(!good) != 0ull) continue;
        if (stamp) t_in = wall_clock64();
        BSum acc{0.0, 0.0f, 0.0f};
        LW_UNROLL for (int b = 0; b < MAX_BATCHES; ++b) {
            BSum r{join_tagged(w[b][0], w[b][1]), 0.0f, 0.0f};
            if constexpr (W > 2) r.s = __uint_as_float((uint32_t)w[b][W > 2 ? 2 : 0]);
            if constexpr (W > 3) r.c = __uint_as_float((uint32_t)w[b][W > 3 ? 3 : 0]);
            if (b < nb) combine<MODE>(acc, r);
        }
        v = entry_of<MODE>(acc);
        return true;
    }
    return false;
}

// ---- the float64 step ------------------------------------------------------------------------------

// An accepted step (odom.py:794-796), lane 0: record it and move the pose T (LDS in the persistent loop, the
// registration state otherwise) by delta. True when the convergence test stops the loop (STEP_TOL,
// odom.py:88, 795-796).
__device__ inline bool accept_step(RegState *st, double *T, const double *delta, double rms, int32_t n_degen,
                                   int32_t it) {
    st->rms = rms;
    if (it == 0) st->rms0 = rms;
    st->n_degen = n_degen;
    gn::apply_twist(T, delta);
    st->iters = it + 1;
    const double nr = sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
    const double nt = sqrt(delta[3] * delta[3] + delta[4] * delta[4] + delta[5] * delta[5]);
    if (nr < 1e-4 && nt < 1e-4) {
        st->status = 3;
        return true;
    }
    return false;
}

// The float64 step after a pass (odom.py:769-796) in one lane: the general routine, for the steps the wave
// form does not take (a degenerate direction to hold, or an indefinite system). True when the loop stops.
// Not inlined: its 6x6 temporaries get their own register budget. The parameters come by value: a reference
// into the kernel's arguments would make the compiler copy the whole argument block to private memory, and
// every pointer in it would then be read back from there as a generic (flat) pointer.
__device__ __noinline__ bool gauss_newton(RegState *st, const double *v, const gn::StepParams sp, int32_t it,
                                          double *T, const double *Tp) {
    gn::StepOut o;
    int rc = gn::step(v, T, Tp, sp, o);
    st->corr = o.corr;
    if (rc == gn::STEP_FEW) {  // corr < min_corr: stop (odom.py:770-771)
        st->status = rc;
        return true;
    }
    for (int k = 0; k < 36; ++k) st->H[k] = o.H[k];
    if (rc == gn::STEP_SINGULAR) {  // LinAlgError: stop (odom.py:791-792)
        st->rms = o.rms;
        if (it == 0) st->rms0 = o.rms;
        st->status = rc;
        return true;
    }
    return accept_step(st, T, o.delta, o.rms, o.n_degen, it);
}

// The step on the whole calling wave (lw_gn_wave.h) when it applies, otherwise gauss_newton in its lane 0 (same
// results either way), on the reduced row v. xi: the twist from T to the prediction Tp (LDS, when the prior is on).
// Every lane returns whether the loop stops; st->done records it. With phase measurement on, a wave step's
// parts go to ph[16..19] (set-up, factorisations, substitution, pose update; device wall-clock ticks) and ph[20],
// and ph[5] / ph[6] count the steps taken on the wave / in one lane.
__device__ inline bool run_step(RegState *st, const double *v, const gn::StepParams &sp, int32_t it, double *sh,
                                double *T, const double *Tp, const double *xi, int32_t *s_flag,
                                unsigned long long *ph) {
    gnw::WaveOut o;
    unsigned long long t1 = 0ull, t2 = 0ull, t3 = 0ull;
    const unsigned long long t0 = ph ? wall_clock64() : 0ull;
    const int r = gnw::wave_step(v, T, xi, sp, sh, o, st->H, ph != nullptr, t1, t2, t3);
    if (lane_id() == 0) {
        bool stop;
        if (r == 0) {
            stop = gauss_newton(st, v, sp, it, T, Tp);
        } else {
            st->corr = o.corr;
            if (r < 0) {  // corr < min_corr: stop (odom.py:770-771)
                st->status = gn::STEP_FEW;
                stop = true;
            } else {
                stop = accept_step(st, T, o.delta, o.rms, 0, it);
            }
        }
        if (stop) st->done = 1;
        *s_flag = stop ? 1 : 0;
        if (ph) {
            atomicAdd(&ph[r == 0 ? 6 : 5], 1ull);
            if (r > 0) {
                const unsigned long long t4 = wall_clock64();
                atomicAdd(&ph[16], t1 - t0);
                atomicAdd(&ph[17], t2 - t1);
                atomicAdd(&ph[18], t3 - t2);
                atomicAdd(&ph[19], t4 - t3);
                atomicAdd(&ph[20], 1ull);
            }
        }
    }
    wave_sync();
    return *s_flag != 0;
}

// ---- partial rows ------------------------------------------------------------------------------------

// Write workgroup row g (its 53 sums and compensations, and its measurement counts in the spare entries 53
// and 54) with device-coherent stores, and wait until they are device-visible.
template <int MODE>
__device__ inline void write_row(const IcpArgs &a, int g, double sum, double comp, int32_t hits, int32_t matched) {
    const int lane = threadIdx.x;
    double *row = a.partials + (size_t)g * 128;
    if (lane < ACC_USED) {
        put(&row[lane], sum);
        if constexpr (MODE == 1) put(&row[64 + lane], comp);
    }
    LW_UNROLL for (int off = 32; off > 0; off >>= 1) {
        hits += __shfl_xor(hits, off, 64);
        matched += __shfl_xor(matched, off, 64);
    }
    if (lane == 0) {
        put(&row[ACC_USED], (double)hits);
        put(&row[ACC_USED + 1], (double)matched);
    }
    stores_done();
}

// The one-pass launch: the row, then a ticket; true for the last workgroup of the pass (which then sees
// every row).
template <int MODE>
__device__ inline bool row_and_ticket(const IcpArgs &a, double sum, double comp, int32_t hits, int32_t matched,
                                      int G, int32_t *s_last) {
    write_row<MODE>(a, blockIdx.x, sum, comp, hits, matched);
    if (threadIdx.x == 0) *s_last = (atomicAdd(&a.st->ticket, 1u) == (uint32_t)G - 1);
    __syncthreads();
    const bool last = *s_last != 0;
    __syncthreads();
    return last;
}

// The one-pass launch's last workgroup: reduce the rows, then (registration) the float64 step. With phase
// measurement on, lane 0 adds the pass's phases to the counters, in device wall-clock ticks: from the pass's
// start (st->t_rel) to this workgroup's ticket (t_ticket), the reduction, and the step.
template <int MODE>
__device__ inline void finish_pass(const IcpArgs &a, const Pass p, int K, int G, double *vs, double *xi,
                                   int32_t *s_flag, unsigned long long t_ticket, double *sh) {
    const int lane = threadIdx.x;
    RegState *st = a.st;
    const double v = reduce_rows<MODE>(a.partials, G);
    vs[lane] = v;
    st->acc[lane] = lane < ACC_USED ? v : 0.0;
    if (a.gn && lane == 0 && a.sp.prior_info > 0.0) gn::twist_between(st->T, st->T_pred, xi);
    __syncthreads();
    const unsigned long long t_red = a.phase ? wall_clock64() : 0ull;
    if (a.gn) {
        run_step(st, vs, a.sp, p.it, sh, st->T, st->T_pred, xi, s_flag, a.phase);
        if (lane == 0)  // the lane that moved the pose
            for (int k = 0; k < 16; ++k) st->Tf[k] = (float)st->T[k];
    }
    if (lane == 0) {
        put(&st->ticket, 0u);
        if (a.prof) {
            atomicAdd(&a.prof[0], (unsigned long long)p.n);
            atomicAdd(&a.prof[1], (unsigned long long)p.n * (unsigned long long)K);
            atomicAdd(&a.prof[2], (unsigned long long)vs[ACC_USED]);
            atomicAdd(&a.prof[3], (unsigned long long)vs[ACC_USED + 1]);
            atomicAdd(&a.prof[4], 1ull);
        }
        if (a.phase) {
            atomicAdd(&a.phase[8], t_ticket - st->t_rel);
            atomicAdd(&a.phase[9], t_red - t_ticket);
            atomicAdd(&a.phase[10], wall_clock64() - t_red);
            atomicAdd(&a.phase[11], 1ull);
        }
    }
}

template <int MODE, int R>
__global__ __launch_bounds__(ICP_BLOCK) void k_icp(IcpArgs a) {
    RegState *st = a.st;
    if (st->done) return;
    __shared__ float Tf[16];
    __shared__ int32_t s_last;
    __shared__ double vs[64], xi[6];
    __shared__ NbrOff nbr[NBR_ROWS];
    __shared__ __align__(16) unsigned char lds[RED_BYTES];
    if (threadIdx.x < 16) Tf[threadIdx.x] = st->Tf[threadIdx.x];
    nbr_fill(nbr);
    __syncthreads();
    double sum, comp;
    int32_t hits = 0, matched = 0;
    const Pass p{a.n, a.max_d2, a.kappa, a.it};
    const int G = (int)gridDim.x;
    if (a.phase && blockIdx.x == 0 && threadIdx.x == 0) a.st->t_rel = wall_clock64();
    unsigned long long t_found, t_terms;
    wave_pass<MODE, R>(a, p, Tf, nbr, blockIdx.x, G, lds, sum, comp, hits, matched, {}, false, false, t_found, t_terms);
    if (!row_and_ticket<MODE>(a, sum, comp, hits, matched, G, &s_last)) return;
    constexpr int D = 2 * R + 1;
    finish_pass<MODE>(a, p, D * D * D, G, vs, xi, &s_last, a.phase ? wall_clock64() : 0ull,
                      reinterpret_cast<double *>(lds));
}

// ---- the persistent registration ------------------------------------------------------------------
//
// Workgroup 0 is the leader, the others search. Per pass, searcher g (= blockIdx.x - 1) takes the point
// groups g, g + G, ... (G = the searchers the pass can use) and writes its tagged partial row; every 16th
// searcher also sums its batch's rows into a batch row; the leader, which searches nothing, combines the batch
// rows, runs the step, and publishes the new pose (float32, tagged; pub in RegState), which the searchers poll
// for. Only
// the leader keeps the float64 pose; the step's code and the reduction run in one workgroup that always
// has them in its caches, and the twist to the prediction is computed while the searchers work. Every wait
// is bounded (SPIN_LIMIT polls, then status STATUS_SPIN_TIMEOUT); co-residency is what the cooperative
// launch guarantees.

struct LoopArgs {
    IcpArgs a;                // pointers and the fixed scalars; n, max_d2, kappa, it are set per pass
    const int32_t *n_dev;     // thinned count on the device
    const int32_t *gate;      // when given, a device flag that must be nonzero for the loop to run
    const uint32_t *epoch;    // the row-tag epoch (advanced by the registration's start)
    uint64_t *rows;           // MAX_LOOP_ROWS tagged rows, then MAX_BATCHES batch rows, of LOOP_ROW_WORDS words
    int32_t max_iter, wide_iters, min_corr;
    float max_d2_r1, max_d2_r2;
};

constexpr uint32_t TAG_STOP = 0x80000000u;

__device__ inline int searchers_for(int64_t n, int S) {
    return (int)min((int64_t)gridDim.x - 1, (n * S + ICP_BLOCK - 1) / ICP_BLOCK);
}

template <int MODE>
__device__ inline void leader(const LoopArgs &L, int32_t n, uint32_t epoch, unsigned char *lds) {
    const IcpArgs &a = L.a;
    RegState *st = a.st;
    const int lane = threadIdx.x;
    __shared__ double sT[16], sTp[16], vs[64], xi[6];
    __shared__ int32_t s_flag;
    static_assert(gnw::W_DOUBLES * sizeof(double) <= RED_BYTES, "the step's scratch fits the searchers' LDS rows");
    double *sh = reinterpret_cast<double *>(lds);
    if (lane < 16) {
        sT[lane] = st->T[lane];
        sTp[lane] = st->T_pred[lane];
    }
    unsigned long long t_rel = a.phase ? wall_clock64() : 0ull;
    if (a.phase && lane == 0) put(&st->t_rel, t_rel);
    __syncthreads();
    for (int it = 0; it < L.max_iter; ++it) {
        const int reach = it < L.wide_iters ? 2 : 1;
        const int G = searchers_for(n, LS);
        // the twist to the prediction, while the pass's search runs (odom.py:775-781)
        if (lane == 0 && a.sp.prior_info > 0.0) gn::twist_between(sT, sTp, xi);
        double v = 0.0;
        unsigned long long t_in = 0ull;
        if (!leader_reduce<MODE>(L.rows + (size_t)MAX_LOOP_ROWS * LOOP_ROW_WORDS, G, row_tag(epoch, it), v, t_in,
                                 a.phase != nullptr)) {
            if (lane == 0) st->status = STATUS_SPIN_TIMEOUT;
            if (lane < 16) put(&st->pub[lane], (unsigned long long)(TAG_STOP | (uint32_t)(it + 1)) << 32);
            return;
        }
        const unsigned long long t_red = a.phase ? wall_clock64() : 0ull;
        if (a.phase && it > 0 && lane == 0) {  // the pass's slowest searcher, from the pose's publication
            atomicAdd(&a.phase[28], atomicExch(&a.phase[27], 0ull));
            atomicAdd(&a.phase[29], 1ull);
        }
        vs[lane] = v;
        st->acc[lane] = lane < ACC_USED ? v : 0.0;
        __syncthreads();
        const bool stop = run_step(st, vs, a.sp, it, sh, sT, sTp, xi, &s_flag, a.phase);
        unsigned long long t_pub = 0ull;
        if (a.phase) {  // the publication time, device-visible before the pose (measurement only)
            t_pub = wall_clock64();
            if (lane == 0) put(&st->t_rel, t_pub);
            stores_done();
        }
        // publish the pose (moved by an accepted step) to the searchers
        if (lane < 16) {
            const uint32_t tag = (uint32_t)(it + 1) | (stop ? TAG_STOP : 0u);
            put(&st->pub[lane], ((unsigned long long)tag << 32) | (unsigned long long)__float_as_uint((float)sT[lane]));
        }
        // off the searchers' path: the host's copies and the measurement
        if (lane < 16) {
            st->T[lane] = sT[lane];
            st->Tf[lane] = (float)sT[lane];
        }
        if (lane == 0) {
            if (a.prof) {
                const int K = reach == 2 ? 125 : 27;
                atomicAdd(&a.prof[0], (unsigned long long)n);
                atomicAdd(&a.prof[1], (unsigned long long)n * (unsigned long long)K);
                atomicAdd(&a.prof[2], (unsigned long long)vs[ACC_USED]);
                atomicAdd(&a.prof[3], (unsigned long long)vs[ACC_USED + 1]);
                atomicAdd(&a.prof[4], 1ull);
            }
            if (a.phase) {
                atomicAdd(&a.phase[8], t_in - t_rel);
                atomicAdd(&a.phase[9], t_red - t_in);
                atomicAdd(&a.phase[10], t_pub - t_red);
                atomicAdd(&a.phase[11], 1ull);
            }
        }
        t_rel = t_pub;
        if (stop) break;
    }
}

template <int MODE>
__device__ inline void searcher(const LoopArgs &L, int32_t n, uint32_t epoch, unsigned char *lds) {
    const IcpArgs &a = L.a;
    RegState *st = a.st;
    const int lane = threadIdx.x;
    const int g = (int)blockIdx.x - 1;
    __shared__ float Tf[16];
    __shared__ NbrOff nbr[NBR_ROWS];
    __shared__ NbrCache cache;
    if (lane < 16) Tf[lane] = st->Tf[lane];
    nbr_fill(nbr);
    cache_clear(&cache);
    // this lane's point in its first lane group: the same every pass, at either reach
    float3 pt = {};
    {
        const int64_t i = ((int64_t)g * ICP_BLOCK + lane) / LS;
        if (i < n) pt = make_float3(a.ds[3 * i], a.ds[3 * i + 1], a.ds[3 * i + 2]);
    }
    __syncthreads();
    const bool stamp = a.phase && g == 0;  // the first searcher's view of the passes
    unsigned long long t_wake = stamp ? wall_clock64() : 0ull;
    // every searcher's view (passes after the first): wake-up and row (with its batch row, for a head) after the
    // pose's publication
    unsigned long long t_pub = 0ull, t_woke = 0ull, sum_wake = 0ull, sum_done = 0ull, n_timed = 0ull;
    for (int it = 0; it < L.max_iter; ++it) {
        const int reach = it < L.wide_iters ? 2 : 1;
        const Pass p{n, reach == 2 ? L.max_d2_r2 : L.max_d2_r1, st->kappa[it], it};
        const int G = searchers_for(n, LS);
        if (g < G) {
            double sum, comp;
            int32_t hits = 0, matched = 0;
            unsigned long long t_found = 0ull, t_terms = 0ull;
            if (reach == 2)
                wave_pass<MODE, 2, LoopSplit<2>::S, LoopSplit<2>::C>(a, p, Tf, nbr, g, G, lds, sum, comp, hits, matched,
                                                                     pt, true, stamp, t_found, t_terms, &cache);
            else
                wave_pass<MODE, 1, LoopSplit<1>::S, LoopSplit<1>::C>(a, p, Tf, nbr, g, G, lds, sum, comp, hits, matched,
                                                                     pt, true, stamp, t_found, t_terms, &cache);
            const unsigned long long t_summed = stamp ? wall_clock64() : 0ull;
            if (a.phase && it > 0 && lane == 0)  // the pass's latest row so far
                atomicMax(&a.phase[27], wall_clock64() - t_pub);
            const uint32_t tag = row_tag(epoch, it);
            write_row_tagged<MODE>(L.rows, g, tag, sum, comp, hits, matched);
            if (g % RED_BATCH == 0 &&
                !batch_head<MODE>(L.rows, L.rows + (size_t)MAX_LOOP_ROWS * LOOP_ROW_WORDS, G, g / RED_BATCH, tag)) {
                if (lane == 0) st->status = STATUS_SPIN_TIMEOUT;
                return;
            }
            if (a.phase && it > 0) {
                sum_wake += t_woke - t_pub;
                sum_done += wall_clock64() - t_pub;
                ++n_timed;
            }
            if (stamp && lane == 0) {
                atomicAdd(&a.phase[13], t_summed - t_wake);
                atomicAdd(&a.phase[14], wall_clock64() - t_summed);
                atomicAdd(&a.phase[15], 1ull);
                atomicAdd(&a.phase[21], t_found - t_wake);
                atomicAdd(&a.phase[22], t_terms - t_found);
                atomicAdd(&a.phase[23], t_summed - t_terms);
                if (reach == 1) {  // the narrow passes' neighbour phase on its own
                    atomicAdd(&a.phase[30], t_found - t_wake);
                    atomicAdd(&a.phase[31], 1ull);
                }
            }
        }
        // wait for the leader's pose of pass it + 1: 16 tagged words, each read with its own tag
        uint32_t spins = 0;
        uint32_t tag0;
        while (true) {
            const unsigned long long w = lane < 16 ? get(&st->pub[lane]) : 0ull;
            const uint32_t tag = (uint32_t)(w >> 32);
            const bool in = lane >= 16 || (tag & ~TAG_STOP) >= (uint32_t)(it + 1);
            if (__ballot(in) == ~0ull) {
                tag0 = __shfl(tag, 0, 64);
                if (lane < 16) Tf[lane] = __uint_as_float((uint32_t)w);
                break;
            }
            __builtin_amdgcn_s_sleep(4);
            if (++spins > SPIN_LIMIT) {
                if (lane == 0) st->status = STATUS_SPIN_TIMEOUT;
                return;
            }
        }
        if (a.phase) {
            t_woke = wall_clock64();
            t_pub = get(&st->t_rel);
            if (stamp) {
                t_wake = t_woke;
                if (lane == 0) atomicAdd(&a.phase[12], t_wake - t_pub);
            }
        }
        __syncthreads();
        if (tag0 & TAG_STOP) break;
    }
    if (a.phase && lane == 0) {
        atomicAdd(&a.phase[24], sum_wake);
        atomicAdd(&a.phase[25], sum_done);
        atomicAdd(&a.phase[26], n_timed);
    }
}

template <int MODE>
__global__ __launch_bounds__(ICP_BLOCK) void k_register(LoopArgs L) {
    RegState *st = L.a.st;
    __shared__ __align__(16) unsigned char lds[RED_BYTES];
    if (gated_off(L.gate)) return;  // the state stays as the previous registration left it
    const int32_t n = *L.n_dev;
    if (blockIdx.x == 0 && threadIdx.x == 0) st->n = n;
    if (n < L.min_corr) return;  // odom.py:756-757: the host sees n and keeps the prediction
    const uint32_t epoch = *L.epoch;  // advanced by this registration's start (an earlier launch)
    if (blockIdx.x == 0) {
        leader<MODE>(L, n, epoch, lds);
        return;
    }
    // searchers that no pass can use (the widest pass needs ceil(n * 4 / 64)) leave now
    if ((int64_t)(blockIdx.x - 1) * ICP_BLOCK >= (int64_t)n * Split<2>::S) return;
    searcher<MODE>(L, n, epoch, lds);
}

// ---- a whole registration in one workgroup (the batched odometry) ---------------------------------------
//
// Independent sequences need no cross-workgroup exchange at all: one workgroup of SEQ_WAVES waves runs a
// sequence's whole Gauss-Newton loop. Per pass the waves take the point groups wave, wave + SEQ_WAVES, ...
// (wave_pass), their rows meet in LDS and are added in wave order (the first RED_BATCH rows of the
// persistent loop's reduction, in the same arithmetic), and wave 0 takes the float64 step while the others
// wait at the workgroup barrier. A pass is bound by the latency of each lane group's dependent gathers, so
// the workgroup runs two waves a SIMD: eight waves of at most 256 registers, each lane taking its keys in
// chunks of 8 at the 5x5x5 reach and 7 at the 3x3x3 (chunks of 16 need more registers than a wave of the eight
// has). The rows of the eight waves use most of the CU's LDS, so a CU runs one sequence at a time and the grid
// is the batch.
constexpr int SEQ_WAVES = 8;
constexpr int SEQ_C2 = 8, SEQ_C1 = 7;  // keys a chunk at the 5x5x5 and 3x3x3 reaches (lanes: Split<R>::S)
static_assert(SEQ_WAVES <= RED_BATCH, "the wave rows are one batch of the reduction");
static_assert(gnw::W_DOUBLES * sizeof(double) <= RED_BYTES, "the step's scratch fits one wave's rows");

template <int MODE>
__global__ __launch_bounds__(64 * SEQ_WAVES) void k_register_seq(const SeqReg *D, RegLoop L,
                                                                unsigned long long *prof) {
    __shared__ __align__(16) unsigned char lds[SEQ_WAVES][RED_BYTES];
    __shared__ NbrOff nbr[NBR_ROWS];
    __shared__ float Tf[16];
    __shared__ double sT[16], sTp[16], vs[64], xi[6], rows[SEQ_WAVES][2][64];
    __shared__ int32_t s_flag;
    const SeqReg d = rebased(D, D[blockIdx.x]);
    if (gated_off(d.gate)) return;  // the state stays as the frame preparation left it
    const int lane = lane_id(), wave = (int)(threadIdx.x >> 6);
    RegState *st = d.st;
    const int32_t n = *d.n_dev;
    if (threadIdx.x == 0) st->n = n;
    if (n < L.min_corr) return;  // odom.py:756-757: the host sees n and keeps the prediction
    IcpArgs a{};
    a.ds = d.ds;
    a.slots = d.slots;
    a.occ = d.occ;
    a.vox_prec = d.vox_prec;
    a.st = st;
    a.mask = d.mask;
    a.inv_v = L.inv_v;
    a.voxel = L.voxel;
    a.gn = 1;
    a.sp = L.sp;
    if (threadIdx.x < 16) {
        sT[threadIdx.x] = st->T[threadIdx.x];
        sTp[threadIdx.x] = st->T_pred[threadIdx.x];
        Tf[threadIdx.x] = st->Tf[threadIdx.x];
    }
    nbr_fill(nbr);
    __syncthreads();
    unsigned long long t_found, t_terms;
    for (int it = 0; it < L.max_iter; ++it) {
        const int reach = it < L.wide_iters ? 2 : 1;
        const Pass p{n, reach == 2 ? L.max_d2_r2 : L.max_d2_r1, st->kappa[it], it};
        double sum, comp;
        int32_t hits = 0, matched = 0;
        // the home records directly, as the single-stream loop reads them: reading the occupancy bitmap first
        // (k_loc's form) measured 3% slower here -- a workgroup's passes are bound by the latency of its
        // dependent loads, not by their bytes
        if (reach == 2)
            wave_pass<MODE, 2, Split<2>::S, SEQ_C2>(a, p, Tf, nbr, wave, SEQ_WAVES, lds[wave], sum, comp, hits,
                                                    matched, {}, false, false, t_found, t_terms);
        else
            wave_pass<MODE, 1, Split<1>::S, SEQ_C1>(a, p, Tf, nbr, wave, SEQ_WAVES, lds[wave], sum, comp, hits,
                                                    matched, {}, false, false, t_found, t_terms);
        LW_UNROLL for (int off = 32; off > 0; off >>= 1) {
            hits += __shfl_xor(hits, off, 64);
            matched += __shfl_xor(matched, off, 64);
        }
        rows[wave][0][lane] = lane < ACC_USED ? sum : (lane == ACC_USED ? (double)hits : (double)matched);
        rows[wave][1][lane] = comp;
        __syncthreads();
        if (wave == 0) {
            BSum r{0.0, 0.0f, 0.0f};
            LW_UNROLL for (int w = 0; w < SEQ_WAVES; ++w) add_row<MODE>(r, true, rows[w][0][lane], rows[w][1][lane]);
            const double v = entry_of<MODE>(r);
            vs[lane] = v;
            st->acc[lane] = lane < ACC_USED ? v : 0.0;
            if (lane == 0 && L.sp.prior_info > 0.0) gn::twist_between(sT, sTp, xi);
            wave_sync();
            // the step's scratch: wave 1's rows, free until the next pass
            run_step(st, vs, L.sp, it, reinterpret_cast<double *>(lds[1]), sT, sTp, xi, &s_flag, nullptr);
            if (lane < 16) {
                st->T[lane] = sT[lane];
                Tf[lane] = (float)sT[lane];
                st->Tf[lane] = Tf[lane];
            }
            if (lane == 0 && prof) {
                const int K = reach == 2 ? 125 : 27;
                atomicAdd(&prof[0], (unsigned long long)n);
                atomicAdd(&prof[1], (unsigned long long)n * (unsigned long long)K);
                atomicAdd(&prof[2], (unsigned long long)vs[ACC_USED]);
                atomicAdd(&prof[3], (unsigned long long)vs[ACC_USED + 1]);
                atomicAdd(&prof[4], 1ull);
            }
        }
        __syncthreads();
        if (s_flag) break;
    }
    // odom.py:797-800: the condition number of the last accepted H, here (a 6x6 eigen problem in one lane: a
    // batch's hundreds of them would otherwise queue up on the host)
    if (threadIdx.x == 0) st->cond = st->corr >= L.min_corr ? gn::cond_of(st->H) : 1.0;
}

// Tp: 16 doubles (the prediction) followed by 64 floats (the per-iteration robust-kernel scale); T_dev, when
// given, holds the prediction instead. Also clears the persistent loop's published pose and advances its row-tag
// epoch. Does
// nothing when gate is given and holds 0.
__global__ void k_reg_begin(RegState *st, const double *Tp, uint32_t *epoch, const int32_t *gate, const double *T_dev) {
    if (gated_off(gate)) return;
    reg_state_begin(st, T_dev ? T_dev : Tp, reinterpret_cast<const float *>(Tp + 16), epoch, threadIdx.x, blockDim.x);
}

// The persistent loop's tagged rows cleared and its row-tag epoch reset (loop_epoch_guard).
__global__ void k_loop_clear(uint64_t *rows, int64_t n, uint32_t *epoch) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        rows[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) *epoch = 0u;
}

template <int MODE>
hipError_t launch_mode(lw_ctx *c, const IcpArgs &a, int reach) {
    if (reach == 1) {
        const int g = (int)(((int64_t)a.n * Split<1>::S + ICP_BLOCK - 1) / ICP_BLOCK);
        hipLaunchKernelGGL((k_icp<MODE, 1>), dim3(g > 0 ? g : 1), dim3(ICP_BLOCK), 0, c->stream, a);
    } else if (reach == 2) {
        const int g = (int)(((int64_t)a.n * Split<2>::S + ICP_BLOCK - 1) / ICP_BLOCK);
        hipLaunchKernelGGL((k_icp<MODE, 2>), dim3(g > 0 ? g : 1), dim3(ICP_BLOCK), 0, c->stream, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

IcpArgs base_args(lw_ctx *c) {
    IcpArgs a{};
    a.ds = c->ds_xyz;
    a.slots = c->slots;
    a.occ = c->occ;
    a.vox_prec = c->vox_prec;
    a.st = c->reg;
    a.partials = c->partials;
    a.best_slot = nullptr;
    a.prof = c->prof_on ? c->prof_dev : nullptr;
    a.phase = c->prof_on && c->prof_phases ? c->prof_dev : nullptr;
    a.mask = c->mask;
    return a;
}

template <int MODE>
const void *loop_kernel() {
    return reinterpret_cast<const void *>(&k_register<MODE>);
}

}  // namespace

hipError_t launch_reg_begin(lw_ctx *c, const double *T_pred_dev, const int32_t *gate, const double *T_dev) {
    hipLaunchKernelGGL(k_reg_begin, dim3(1), dim3(64), 0, c->stream, c->reg, T_pred_dev, c->loop_epoch, gate, T_dev);
    return hipGetLastError();
}

hipError_t loop_epoch_guard(lw_ctx *c, uint32_t starts) {
    if (c->reg_starts + starts <= LOOP_EPOCH_LIMIT) {
        c->reg_starts += starts;
        return hipSuccess;
    }
    hipLaunchKernelGGL(k_loop_clear, dim3(256), dim3(256), 0, c->stream, c->loop_rows,
                       (int64_t)LOOP_TAGGED_ROWS * LOOP_ROW_WORDS, c->loop_epoch);
    c->reg_starts = starts;
    return hipGetLastError();
}

int32_t icp_max_groups(int64_t frame_cap) { return (int32_t)((frame_cap * 4 + ICP_BLOCK - 1) / ICP_BLOCK); }

hipError_t launch_icp_pass(lw_ctx *c, const IcpPass &p, int acc_mode) {
    if (p.n > c->frame_cap) return hipErrorInvalidValue;
    IcpArgs a = base_args(c);
    a.best_slot = p.want_slots ? c->best_slot : nullptr;
    a.n = p.n;
    a.inv_v = p.inv_v;
    a.voxel = p.voxel;
    a.max_d2 = p.max_d2;
    a.kappa = p.kappa;
    a.it = p.it;
    a.gn = p.gn;
    a.sp = p.sp;
    const bool timed = c->prof_on && c->prof_used < c->prof_cap;
    if (timed) {
        hipError_t e = hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream);
        if (e != hipSuccess) return e;
    }
    hipError_t e;
    switch (acc_mode) {
    case 0: e = launch_mode<0>(c, a, p.reach); break;
    case 1: e = launch_mode<1>(c, a, p.reach); break;
    case 2: e = launch_mode<2>(c, a, p.reach); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (timed) {
        e = hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream);
        c->prof_used++;
    }
    return e;
}

hipError_t launch_register_seq(hipStream_t s, const SeqReg *D, int32_t S, const RegLoop &L, int acc_mode,
                               unsigned long long *prof) {
    if (S <= 0) return hipSuccess;
    if (L.max_iter > 64) return hipErrorInvalidValue;
    switch (acc_mode) {
    case 0: hipLaunchKernelGGL(k_register_seq<0>, dim3(S), dim3(64 * SEQ_WAVES), 0, s, D, L, prof); break;
    case 1: hipLaunchKernelGGL(k_register_seq<1>, dim3(S), dim3(64 * SEQ_WAVES), 0, s, D, L, prof); break;
    case 2: hipLaunchKernelGGL(k_register_seq<2>, dim3(S), dim3(64 * SEQ_WAVES), 0, s, D, L, prof); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_register_loop(lw_ctx *c, const int32_t *n_dev, int32_t n_bound, const RegLoop &R, int acc_mode,
                                const int32_t *gate) {
    if (acc_mode < 0 || acc_mode > 2 || R.max_iter > 64 || n_bound > c->frame_cap) return hipErrorNotSupported;
    const void *fn = acc_mode == 0 ? loop_kernel<0>() : acc_mode == 1 ? loop_kernel<1>() : loop_kernel<2>();
    if (c->loop_max_groups[acc_mode] == 0) {  // co-resident workgroups, once per context
        int per_cu = 0, cus = 0, coop = 0;
        if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device) != hipSuccess || !coop)
            return hipErrorNotSupported;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, ICP_BLOCK, 0) != hipSuccess || per_cu <= 0)
            return hipErrorNotSupported;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
            return hipErrorNotSupported;
        // co-resident limit, and at most two workgroups per CU: the loop never needs more than one lane
        // group per searcher (8192 thinned points at the 5x5x5 reach), and every extra workgroup only
        // polls for the pose
        c->loop_max_groups[acc_mode] = per_cu < 2 ? per_cu * cus : 2 * cus;
        c->loop_capacity[acc_mode] = per_cu * cus;
        if (c->loop_max_groups[acc_mode] < 2) return hipErrorNotSupported;
    }
    LoopArgs L;
    L.a = base_args(c);
    L.a.inv_v = R.inv_v;
    L.a.voxel = R.voxel;
    L.a.gn = 1;
    L.a.sp = R.sp;
    L.n_dev = n_dev;
    L.gate = gate;
    L.epoch = c->loop_epoch;
    L.rows = c->loop_rows;
    L.max_iter = R.max_iter;
    L.wide_iters = R.wide_iters;
    L.min_corr = R.min_corr;
    L.max_d2_r1 = R.max_d2_r1;
    L.max_d2_r2 = R.max_d2_r2;
    // the leader, and one searcher per lane group of the widest pass, up to what can be co-resident
    int64_t g = ((int64_t)(n_bound > 0 ? n_bound : 1) * Split<2>::S + ICP_BLOCK - 1) / ICP_BLOCK;
    if (g > c->loop_max_groups[acc_mode] - 1) g = c->loop_max_groups[acc_mode] - 1;
    if (g > MAX_LOOP_ROWS) g = MAX_LOOP_ROWS;
    if (g > icp_max_groups(c->frame_cap)) g = icp_max_groups(c->frame_cap);
    if (g < 1) g = 1;
    g += 1;
    void *args[] = {&L};
    const bool timed = c->prof_on && c->prof_used < c->prof_cap;
    if (timed) {
        hipError_t e = hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream);
        if (e != hipSuccess) return e;
    }
    // A plain launch when the grid is at most half of what the device holds at once: every workgroup then
    // becomes resident as soon as the device is free of other work (a kernel on another stream delays some,
    // it cannot keep them out), which is what the loop's waits need, without the barriers the runtime puts
    // around a cooperative launch. Otherwise (or on request) the cooperative launch guarantees it.
    hipError_t e;
    if (c->loop_plain && 2 * g <= c->loop_capacity[acc_mode])
        e = hipLaunchKernel(fn, dim3((unsigned)g), dim3(ICP_BLOCK), args, 0, c->stream);
    else
        e = hipLaunchCooperativeKernel(fn, dim3((unsigned)g), dim3(ICP_BLOCK), args, 0, c->stream);
    if (e != hipSuccess) return e;
    if (timed) {
        e = hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream);
        c->prof_used++;
    }
    return e;
}

}  // namespace lw
