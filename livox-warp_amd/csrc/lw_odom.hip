// The decisions Odometry._finalize / _estimate take between device steps (odom.py:644-692), taken on the
// device so that a frame's whole chain -- deskew, registration, the re-deskew test and second registration,
// the keyframe test and insert -- is queued at once and read back once (lw_odom_frame in lw_abi.hip). Each
// decision is one lane of a tiny launch that reads the registration state and writes a FrameCtl flag the
// next kernels are gated on.
#include <algorithm>
#include <climits>

#include "lw_frame.h"
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

// What the frame check decides (FrameCtl's fields of the same names).
struct Check {
    double T1[16];
    float xi1[6];
    int32_t m, skipped, ok1, redo;
};

// After the first registration (one lane): the frame's count and skip test (odom.py:676-678) and the re-deskew
// test (odom.py:683-690) on xi_new = body twist of T_prev^-1 T, from the registration state and the frame's
// inputs alone. cnt: the frame's appended points.
__device__ inline void frame_check_of(const RegState *st, const FrameCtl *fc, int32_t cnt, Check &c) {
    const int32_t m = cnt < fc->frame_cap ? cnt : fc->frame_cap;
    c.m = m;
    c.skipped = m < fc->min_points ? 1 : 0;
    double T1[16];
    const bool ok1 = !c.skipped && reg_outcome(st, fc->T_pred, fc->min_corr, T1);
    for (int k = 0; k < 16; ++k) c.T1[k] = T1[k];
    c.ok1 = ok1 ? 1 : 0;
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
    c.redo = redo ? 1 : 0;
    for (int k = 0; k < 6; ++k) c.xi1[k] = redo ? xn[k] : fc->xi0[k];
}

__device__ inline void check_store(const Check &c, FrameCtl *fc) {
    fc->m = c.m;
    fc->skipped = c.skipped;
    for (int k = 0; k < 16; ++k) fc->T1[k] = c.T1[k];
    fc->ok1 = c.ok1;
    fc->redo = c.redo;
    for (int k = 0; k < 6; ++k) fc->xi1[k] = c.xi1[k];
}

__device__ inline void frame_check(const RegState *st, FrameCtl *fc, int32_t cnt) {
    Check c;
    frame_check_of(st, fc, cnt, c);
    check_store(c, fc);
}

// Keep the first registration's state in st1 (the second registration reuses st); every thread of the block
// takes part.
__device__ inline void keep_state(const RegState *st, RegState *st1) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(st);
    uint32_t *dst = reinterpret_cast<uint32_t *>(st1);
    for (size_t k = threadIdx.x; k < sizeof(RegState) / 4; k += blockDim.x) dst[k] = src[k];
}

__global__ void k_frame_check(const RegState *st, RegState *st1, FrameCtl *fc, const int32_t *ctr) {
    keep_state(st, st1);
    if (threadIdx.x == 0) frame_check(st, fc, ctr[6]);
}

// After the (possible) second registration: the frame's pose and trust, and the keyframe test
// (odom.py:650-651, 712-717) against the last keyframe, with the float32 pose an insert uses. counts gets the
// map's counts (ctr[0..1]) as they stand before the insert (which, when it runs, writes them again after).
__device__ inline void keyframe_decide(const RegState *st, FrameCtl *fc, const int32_t *ctr, int32_t *counts) {
    counts[0] = ctr[0];
    counts[1] = ctr[1];
    double T[16];
    bool ok;
    if (fc->warm) {  // odom.py:650-651, 669-670: a warm-up frame keeps its pose and is inserted when it has points
        for (int k = 0; k < 16; ++k) {
            fc->T[k] = fc->T_pred[k];
            fc->T32[k] = (float)fc->T_pred[k];
        }
        fc->ok = fc->skipped ? 0 : 1;
        fc->keyframe = fc->ok;
        return;
    }
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

__global__ void k_keyframe_check(const RegState *st, FrameCtl *fc, const int32_t *ctr, int32_t *counts) {
    keyframe_decide(st, fc, ctr, counts);
}

__global__ void k_keyframe_check_seq(const KeySeq *D, int32_t S) {
    const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const KeySeq d = rebased(D, D[s]);
    keyframe_decide(d.st, d.fc, d.ctr, d.counts);
}

// ---- the learned prior in the frame chain (lw_odom_frame_prior) --------------------------------------------
//
// LearnedPrior.guess and .accept, between the chain's steps. Before the first preparation (one lane): the
// constant-velocity prediction and twist the host uploaded are kept (T_cv, xi_cv); when the frame holds at least
// min_points points (guess() is asked only then), the network's output y becomes the increment
// d = hist0 + y sd_t + mu_t and the first registration starts from prior_T exp(d) -- written into the chain's
// prediction (tp) and T_pred -- with the points deskewed by d. After the first registration: a first residual
// above the gate (fallback x the running median of earlier first residuals, taken by the host) keeps that
// registration's state for the host and turns the frame back to the constant-velocity prediction and twist,
// and the preparation and registration gated on fb run from there (odom.py's prediction, as _estimate's
// fallback does).
__global__ void k_prior_guess(FrameCtl *fc, const float *y, double *tp, const int32_t *m_dev, int32_t cap) {
    for (int k = 0; k < 16; ++k) fc->T_cv[k] = fc->T_pred[k];
    for (int k = 0; k < 6; ++k) fc->xi_cv[k] = fc->xi0[k];
    fc->fb = 0;
    const int32_t m = count_of(cap, m_dev);
    fc->guess = fc->prior_on && m >= fc->min_points ? 1 : 0;
    if (!fc->guess) return;
    double d[6], R[9], Tw[16], Tg[16];
    for (int k = 0; k < 6; ++k) d[k] = fc->hist0[k] + (double)y[k] * fc->sd_t[k] + fc->mu_t[k];
    gn::rot_from_rotvec(d + 3, R);  // twist_matrix(d): the rotation of d's rotation vector, d's translation
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) Tw[4 * i + j] = R[3 * i + j];
        Tw[4 * i + 3] = d[i];
        Tw[12 + i] = 0.0;
    }
    Tw[15] = 1.0;
    mat4_mul(fc->prior_T, Tw, Tg);
    for (int k = 0; k < 16; ++k) {
        tp[k] = Tg[k];
        fc->T_pred[k] = Tg[k];
    }
    for (int k = 0; k < 6; ++k) {
        fc->d[k] = d[k];
        fc->xi0[k] = (float)d[k];
    }
}

__global__ void k_prior_fallback(FrameCtl *fc, const RegState *st, RegState *st_fb, double *tp) {
    __shared__ int32_t s_fb;
    if (threadIdx.x == 0) s_fb = fc->guess && !(st->rms0 <= fc->rms_gate) ? 1 : 0;
    __syncthreads();
    if (!s_fb) return;
    keep_state(st, st_fb);
    if (threadIdx.x == 0) {
        fc->fb = 1;
        for (int k = 0; k < 16; ++k) {
            tp[k] = fc->T_cv[k];
            fc->T_pred[k] = fc->T_cv[k];
        }
        for (int k = 0; k < 6; ++k) fc->xi0[k] = fc->xi_cv[k];
    }
}

// ---- one workgroup: deskew, thinning and the registration's start ------------------------------------
//
// A frame of at most PREP_MAX points is deskewed and thinned by one workgroup of PREP_THREADS threads in one
// launch, instead of the grid kernels' five (deskew, claim, count, scan, write), and the registration state is
// set up in the same launch. Point i is handled by thread i mod PREP_THREADS in round i / PREP_THREADS, so
// the rounds cover the points in index order. The thinning claims keys in a table in LDS (PREP_SLOTS slots,
// linear probing over PROBES slots like the global table), where each key keeps its lowest point index; the
// winners are then compacted in index order from per-(round, wave) counts. Where no key is dropped (every key
// finds a slot within PROBES of its home), the result does not depend on the table -- one point per thinning
// voxel, the one of lowest index, in index order -- so the LDS table uses its own hash over every key bit (the
// global hash's low bits leave out the z index: a column's voxels would share a home). The workgroup also
// shows that the global table drops nothing: it counts the keys homed in each aligned block of PREP_BLOCK
// global slots (block indices folded onto PREP_COUNTERS counters, which can only raise the counts), and at
// most PREP_BLOCK_KEYS homes per block leave every cluster of the global table shorter than 64 slots (a run of
// L >= 64 full slots holds L keys homed inside it, but meets at most (L - 1) / PREP_BLOCK + 2 blocks, which
// hold at most (L - 1) / 2 + 2 PREP_BLOCK_KEYS < L keys), so no key probes past its 64th slot, in any
// insertion order. When the check fails, or a key finds no free slot
// in the LDS table, the workgroup runs the claims again in the global table, exactly as the grid kernels do.
constexpr int PREP_THREADS = ONE_THREADS;
constexpr int PREP_ROUNDS = ONE_ROUNDS;
constexpr int PREP_SLOTS = 8192;
constexpr int PREP_WAVES = ONE_WAVES;
constexpr int PREP_BATCH = 4;  // rounds whose loads are issued together
constexpr int PREP_BLOCK = 32, PREP_BLOCK_KEYS = 16, PREP_COUNTERS = 8192;
static_assert(2 * PREP_BLOCK_KEYS <= PREP_BLOCK && 4 * PREP_BLOCK_KEYS <= PROBES, "no cluster of PROBES slots");
static_assert(PREP_ROUNDS % PREP_BATCH == 0 && PREP_ROUNDS <= 32, "one bit per round");


// The claim of one key in the LDS table: the slot, or -1 when none is free within PROBES slots; fresh when
// this claim put the key in.
__device__ inline int32_t lds_claim(int64_t *tkey, int32_t *tidx, int64_t key, int32_t i, bool &fresh) {
    const int32_t h = (int32_t)(((uint64_t)key * HASH_MUL) >> (64 - 13));
    static_assert(PREP_SLOTS == 1 << 13, "the LDS hash takes 13 bits");
    fresh = false;
    for (int probe = 0; probe < PROBES; ++probe) {
        const int32_t s = (h + probe) & (PREP_SLOTS - 1);
        const unsigned long long prev =
            atomicCAS(reinterpret_cast<unsigned long long *>(&tkey[s]), ~0ull, (unsigned long long)key);
        if ((int64_t)prev == -1 || (int64_t)prev == key) {
            atomicMin(&tidx[s], i);
            fresh = (int64_t)prev == -1;
            return s;
        }
    }
    return -1;
}

// The claim of one key in the global table (k_thin_claim's probe).
__device__ inline int32_t global_claim(int64_t *ds_keys, int32_t *ds_idx, int32_t ds_mask, int64_t key, int32_t i) {
    const int32_t h = slot_hash(key) & ds_mask;
    for (int probe = 0; probe < PROBES; ++probe) {
        const int32_t s = (h + probe) & ds_mask;
        const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long *>(&ds_keys[s]),
                                                  0xFFFFFFFFFFFFFFFFull, (unsigned long long)key);
        if ((int64_t)prev == -1 || (int64_t)prev == key) {
            atomicMin(&ds_idx[s], i);
            return s;
        }
    }
    return -1;
}

__device__ __forceinline__ void frame_prep(const PrepArgs &A) {
    __shared__ int64_t tkey[PREP_SLOTS];
    __shared__ int32_t tidx[PREP_SLOTS];
    __shared__ int32_t cnt[PREP_ROUNDS * PREP_WAVES];
    __shared__ int32_t blk[PREP_COUNTERS];
    __shared__ double s_T[16];
    __shared__ float s_xi[6];
    __shared__ int32_t s_go, s_over;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (A.st1) {  // after the first registration: keep its state; the skip and re-deskew tests
        keep_state(A.st, A.st1);
        if (tid == 0) {
            frame_check(A.st, A.fc, *A.m_dev);
            s_go = A.fc->redo;
            for (int k = 0; k < 16; ++k) s_T[k] = A.fc->T1[k];
            for (int k = 0; k < 6; ++k) s_xi[k] = A.fc->xi1[k];
        }
    } else if (tid == 0) {
        s_go = A.gate ? (*A.gate != 0 ? 1 : 0) : 1;
        if (!A.gate) A.fc->err = 0;
        for (int k = 0; k < 16; ++k) s_T[k] = A.Tp[k];
        for (int k = 0; k < 6; ++k) s_xi[k] = A.xi_dev ? A.xi_dev[k] : A.xi[k];
    }
    for (int s = tid; s < PREP_SLOTS; s += PREP_THREADS) {
        tkey[s] = -1;
        tidx[s] = INT_MAX;
    }
    for (int s = tid; s < PREP_COUNTERS; s += PREP_THREADS) blk[s] = 0;
    if (tid == 0) s_over = 0;
    __syncthreads();
    if (!s_go) return;
    reg_state_begin(A.st, s_T, reinterpret_cast<const float *>(A.Tp + 16), A.epoch, tid, PREP_THREADS);
    const int32_t m = count_of(A.frame_cap, A.m_dev);
    if (m > PREP_THREADS * PREP_ROUNDS) {  // more than the caller's bound: thin nothing, and say so
        if (tid == 0) {
            A.fc->err = 1;
            *A.n_out = 0;
        }
        return;
    }
    const int rounds = (m + PREP_THREADS - 1) / PREP_THREADS;
    float xi[6];
    for (int k = 0; k < 6; ++k) xi[k] = s_xi[k];
    // deskew, filters, claims in LDS; each point's slot stays in a register (slot[r] for round r) and the
    // loads of PREP_BATCH rounds are issued together
    int32_t slot[PREP_ROUNDS];
    LW_UNROLL for (int r = 0; r < PREP_ROUNDS; ++r) slot[r] = -1;
    LW_UNROLL for (int b = 0; b < PREP_ROUNDS; b += PREP_BATCH) {
        if (b >= rounds) break;
        float px[PREP_BATCH], py[PREP_BATCH], pz[PREP_BATCH], pt[PREP_BATCH];
        uint32_t pa[PREP_BATCH];
        LW_UNROLL for (int u = 0; u < PREP_BATCH; ++u) {
            const int32_t i = (b + u) * PREP_THREADS + tid;
            const int32_t ic = i < m ? i : m - 1;  // m >= 1 here; a clamped load is not used
            px[u] = A.raw[3 * ic];
            py[u] = A.raw[3 * ic + 1];
            pz[u] = A.raw[3 * ic + 2];
            pt[u] = A.t[ic];
            pa[u] = A.attr[ic];
        }
        LW_UNROLL for (int u = 0; u < PREP_BATCH; ++u) {
            const int32_t i = (b + u) * PREP_THREADS + tid;
            bool v = false;
            int64_t key = 0;
            if (i < m) {
                float qx, qy, qz;
                deskew_point(px[u], py[u], pz[u], pt[u], A.t_mid, A.inv_dt, xi, qx, qy, qz);
                A.xyz[3 * i] = qx;
                A.xyz[3 * i + 1] = qy;
                A.xyz[3 * i + 2] = qz;
                v = thin_key(qx, qy, qz, pa[u], A.inv_v, A.min_range, A.max_range, A.ret_mask, key);
            }
            if (v) {
                bool fresh;
                slot[b + u] = lds_claim(tkey, tidx, key, i, fresh);
                // no free slot, or one block of the global table with more homes than the bound allows
                if (slot[b + u] < 0 ||
                    (fresh && atomicAdd(&blk[((slot_hash(key) & A.ds_mask) / PREP_BLOCK) & (PREP_COUNTERS - 1)],
                                        1) == PREP_BLOCK_KEYS))
                    s_over = 1;
            }
        }
    }
    __syncthreads();
    const bool glob = s_over != 0;  // not shown drop-free: the claims again, in the global table
    if (glob) {
        LW_UNROLL for (int r = 0; r < PREP_ROUNDS; ++r) {
            if (r >= rounds) break;
            const int32_t i = r * PREP_THREADS + tid;
            slot[r] = -1;
            int64_t key;
            if (i < m && thin_key(A.xyz[3 * i], A.xyz[3 * i + 1], A.xyz[3 * i + 2], A.attr[i], A.inv_v, A.min_range,
                                  A.max_range, A.ret_mask, key))
                slot[r] = global_claim(A.ds_keys, A.ds_idx, A.ds_mask, key, i);
        }
        __threadfence();
        __syncthreads();
    }
    // winners (the lowest index of each claimed slot), counted per round and wave
    uint32_t win = 0u;
    LW_UNROLL for (int r = 0; r < PREP_ROUNDS; ++r) {
        if (r >= rounds) break;
        const int32_t i = r * PREP_THREADS + tid;
        const int32_t s = slot[r];
        bool w = false;
        if (s >= 0)
            w = (glob ? __hip_atomic_load(&A.ds_idx[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : tidx[s]) == i;
        win |= (w ? 1u : 0u) << r;
        const unsigned long long bw = __ballot(w);
        if (lane == 0) cnt[r * PREP_WAVES + wave] = __popcll(bw);
    }
    __syncthreads();
    // exclusive offsets over (round, wave) in that order
    if (wave == 0) {
        const int32_t total = scan_round_counts(cnt, rounds);
        if (lane == 0) *A.n_out = total;
    }
    __syncthreads();
    // the winners in index order, PREP_BATCH rounds' loads at a time
    LW_UNROLL for (int b = 0; b < PREP_ROUNDS; b += PREP_BATCH) {
        if (b >= rounds) break;
        float x[PREP_BATCH], y[PREP_BATCH], z[PREP_BATCH];
        LW_UNROLL for (int u = 0; u < PREP_BATCH; ++u) {
            const int32_t i = (b + u) * PREP_THREADS + tid;
            if ((win >> (b + u)) & 1u) {
                x[u] = A.xyz[3 * i];
                y[u] = A.xyz[3 * i + 1];
                z[u] = A.xyz[3 * i + 2];
            }
        }
        LW_UNROLL for (int u = 0; u < PREP_BATCH; ++u) {
            const bool w = ((win >> (b + u)) & 1u) != 0u;
            const unsigned long long bw = __ballot(w);
            if (w) {
                const int32_t j = cnt[(b + u) * PREP_WAVES + wave] + __popcll(bw & ((1ull << lane) - 1ull));
                A.ds_xyz[3 * j] = x[u];
                A.ds_xyz[3 * j + 1] = y[u];
                A.ds_xyz[3 * j + 2] = z[u];
                if (glob) {  // the global table is left clean, as the grid kernels leave it
                    A.ds_keys[slot[b + u]] = -1;
                    A.ds_idx[slot[b + u]] = INT_MAX;
                }
            }
        }
    }
}

__global__ __launch_bounds__(PREP_THREADS) void k_frame_prep(PrepArgs A) { frame_prep(A); }

// One frame per workgroup: frame blockIdx.x of a batch.
__global__ __launch_bounds__(PREP_THREADS) void k_frame_prep_seq(const PrepArgs *D) {
    const PrepArgs A = rebased(D, D[blockIdx.x]);
    frame_prep(A);
}

// ---- many workgroups: deskew, thinning and the registration's start --------------------------------------
//
// The same preparation over G workgroups of PG_THREADS threads (one launch, two grid barriers), for the
// single-stream chain, where one workgroup leaves the rest of the device idle. Workgroup g takes the contiguous
// indices [g chunk, (g + 1) chunk), chunk a multiple of PG_THREADS, in rounds of PG_THREADS, so index order is
// (workgroup, round, wave, lane) order. Every workgroup takes the frame check itself (the same reads, the same
// decision; workgroup 0 records it and keeps the first registration's state), then deskews its points and
// claims their keys in the global thinning table (gpu.py's probe, each slot keeping the lowest index, exactly as
// the grid kernels do). After the first barrier every claim is in: each point learns whether it holds its slot's
// lowest index, each workgroup counts its winners per (round, wave) and publishes its total, and workgroup 0
// starts the registration state (nothing reads the old one any more). After the second barrier each workgroup
// writes its winners at its offset (the totals of the workgroups before it) in index order and clears their
// slots. The result is k_frame_prep's: one point per thinning voxel, the lowest index, in index order.
//
// The barrier counts arrivals on one 64-bit counter that never resets: a launch reads the base it started
// from, waits for base + G and base + 2 G, and workgroup 0 moves the base on after the second. Workgroups are
// few (at most PG_MAX_G) and small; every wait is bounded (then the frame reports err 2). Data crossing
// workgroups (the claims, the totals) moves by device-coherent atomics; a workgroup's deskewed points are read
// back only by itself.
constexpr int PG_THREADS = 256, PG_WAVES = PG_THREADS / 64, PG_ROUNDS = 8;
static_assert(PG_ROUNDS * PG_WAVES <= 64, "one wave scans the per-(round, wave) counts");
constexpr uint32_t PG_SPIN_LIMIT = 1u << 24;

// Arrive at the grid barrier, then wait until target arrivals are in. Every thread of the workgroup calls it;
// false when the wait ran out.
__device__ inline bool grid_barrier(PrepSync *ps, unsigned long long target) {
    __shared__ int32_t s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this workgroup's claims, totals and stores are done
        __hip_atomic_fetch_add(&ps->count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int32_t ok = 0;
        for (uint32_t spins = 0; spins <= PG_SPIN_LIMIT; ++spins) {
            if (__hip_atomic_load(&ps->count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) {
                ok = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

__global__ __launch_bounds__(PG_THREADS) void k_frame_prep_grid(PrepArgs A, PrepSync *ps) {
    __shared__ double s_T[16];
    __shared__ float s_xi[6];
    __shared__ int32_t s_go, s_off;
    __shared__ unsigned long long s_base;
    __shared__ int32_t cnt[PG_ROUNDS * PG_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = (int)blockIdx.x, G = (int)gridDim.x;
    if (A.st1 && g == 0) keep_state(A.st, A.st1);  // the first registration's state, kept for the host
    if (tid == 0) {
        s_base = __hip_atomic_load(&ps->base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (A.st1) {  // after the first registration: the skip and re-deskew tests
            Check c;
            frame_check_of(A.st, A.fc, *A.m_dev, c);
            if (g == 0) check_store(c, A.fc);
            s_go = c.redo;
            for (int k = 0; k < 16; ++k) s_T[k] = c.T1[k];
            for (int k = 0; k < 6; ++k) s_xi[k] = c.xi1[k];
        } else {
            s_go = A.gate ? (*A.gate != 0 ? 1 : 0) : 1;
            if (g == 0 && !A.gate) A.fc->err = 0;
            for (int k = 0; k < 16; ++k) s_T[k] = A.Tp[k];
            for (int k = 0; k < 6; ++k) s_xi[k] = A.xi_dev ? A.xi_dev[k] : A.xi[k];
        }
    }
    __syncthreads();
    if (!s_go) return;  // every workgroup decides alike, before any barrier
    const int32_t m = count_of(A.frame_cap, A.m_dev);
    const int32_t chunk = (((m + G - 1) / G) + PG_THREADS - 1) / PG_THREADS * PG_THREADS;
    const int rounds = chunk / PG_THREADS;
    if (rounds > PG_ROUNDS) {  // more than the launch's bound (every workgroup alike): thin nothing, and say so
        if (g == 0 && tid == 0) {
            A.fc->err = 1;
            *A.n_out = 0;
        }
        return;
    }
    const int32_t lo = g * chunk;
    float xi[6];
    for (int k = 0; k < 6; ++k) xi[k] = s_xi[k];
    int32_t slot[PG_ROUNDS];
    float qx[PG_ROUNDS], qy[PG_ROUNDS], qz[PG_ROUNDS];
    LW_UNROLL for (int r = 0; r < PG_ROUNDS; ++r) slot[r] = -1;
    LW_UNROLL for (int r = 0; r < PG_ROUNDS; ++r) {
        if (r >= rounds) break;
        const int32_t i = lo + r * PG_THREADS + tid;
        if (i < m) {
            deskew_point(A.raw[3 * i], A.raw[3 * i + 1], A.raw[3 * i + 2], A.t[i], A.t_mid, A.inv_dt, xi, qx[r],
                         qy[r], qz[r]);
            A.xyz[3 * i] = qx[r];
            A.xyz[3 * i + 1] = qy[r];
            A.xyz[3 * i + 2] = qz[r];
            int64_t key;
            if (thin_key(qx[r], qy[r], qz[r], A.attr[i], A.inv_v, A.min_range, A.max_range, A.ret_mask, key))
                slot[r] = global_claim(A.ds_keys, A.ds_idx, A.ds_mask, key, i);
        }
    }
    const bool ok1 = grid_barrier(ps, s_base + (unsigned long long)G);
    if (g == 0) reg_state_begin(A.st, s_T, reinterpret_cast<const float *>(A.Tp + 16), A.epoch, tid, PG_THREADS);
    // winners, counted per (round, wave)
    uint32_t win = 0u;
    LW_UNROLL for (int r = 0; r < PG_ROUNDS; ++r) {
        if (r >= rounds) break;
        const int32_t i = lo + r * PG_THREADS + tid;
        const int32_t s = slot[r];
        const bool w = s >= 0 && __hip_atomic_load(&A.ds_idx[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == i;
        win |= (w ? 1u : 0u) << r;
        const unsigned long long bw = __ballot(w);
        if (lane == 0) cnt[r * PG_WAVES + wave] = __popcll(bw);
    }
    __syncthreads();
    if (wave == 0) {  // exclusive offsets over (round, wave) in that order; the workgroup's total published
        const int n = rounds * PG_WAVES;
        const int32_t v = lane < n ? cnt[lane] : 0;
        int32_t incl = v;
        for (int d = 1; d < 64; d <<= 1) {
            const int32_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane < n) cnt[lane] = incl - v;
        const int32_t total = __shfl(incl, 63, 64);
        if (lane == 0) __hip_atomic_store(&ps->total[g], total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const bool ok2 = grid_barrier(ps, s_base + 2ull * (unsigned long long)G);
    if (!ok1 || !ok2) {
        if (tid == 0) A.fc->err = 2;
        return;
    }
    if (wave == 0) {  // this workgroup's offset; workgroup 0 also the frame's count and the next launch's base
        int32_t before = 0, all = 0;
        for (int k = lane; k < G; k += 64) {
            const int32_t t = __hip_atomic_load(&ps->total[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            all += t;
            before += k < g ? t : 0;
        }
        for (int d = 32; d > 0; d >>= 1) {
            before += __shfl_xor(before, d, 64);
            all += __shfl_xor(all, d, 64);
        }
        if (lane == 0) {
            s_off = before;
            if (g == 0) {
                *A.n_out = all;
                __hip_atomic_store(&ps->base, s_base + 2ull * (unsigned long long)G, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    __syncthreads();
    LW_UNROLL for (int r = 0; r < PG_ROUNDS; ++r) {
        if (r >= rounds) break;
        const bool w = ((win >> r) & 1u) != 0u;
        const unsigned long long bw = __ballot(w);
        if (w) {
            const int32_t j = s_off + cnt[r * PG_WAVES + wave] + __popcll(bw & ((1ull << lane) - 1ull));
            A.ds_xyz[3 * j] = qx[r];
            A.ds_xyz[3 * j + 1] = qy[r];
            A.ds_xyz[3 * j + 2] = qz[r];
            A.ds_keys[slot[r]] = -1;  // the table left clean, as the grid kernels leave it
            A.ds_idx[slot[r]] = INT_MAX;
        }
    }
}

}  // namespace

int32_t frame_prep_max() { return PREP_THREADS * PREP_ROUNDS; }

int32_t frame_prep_grid_max() { return PG_MAX_G * PG_THREADS * PG_ROUNDS; }

hipError_t launch_frame_prep_seq(hipStream_t s, const PrepArgs *D, int32_t S) {
    if (S <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_frame_prep_seq, dim3(S), dim3(PREP_THREADS), 0, s, D);
    return hipGetLastError();
}

hipError_t launch_keyframe_check_seq(hipStream_t s, const KeySeq *D, int32_t S) {
    if (S <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_keyframe_check_seq, dim3((S + 63) / 64), dim3(64), 0, s, D, S);
    return hipGetLastError();
}

namespace {

PrepArgs prep_args(lw_ctx *c, const double *Tp, float t_mid, float inv_dt, const float *xi, float inv_v,
                   float min_range, float max_range, int32_t ret_mask, bool check) {
    PrepArgs A{};
    A.raw = c->f_raw;
    A.t = c->f_t;
    A.attr = c->f_attr;
    A.xyz = c->f_xyz;
    A.m_dev = &c->ctr[6];
    A.frame_cap = (int32_t)c->frame_cap;
    A.t_mid = t_mid;
    A.inv_dt = inv_dt;
    for (int k = 0; k < 6; ++k) A.xi[k] = xi ? xi[k] : 0.0f;
    A.inv_v = inv_v;
    A.min_range = min_range;
    A.max_range = max_range;
    A.ret_mask = ret_mask;
    A.ds_mask = (int32_t)(c->ds_cap - 1);
    A.ds_keys = c->ds_keys;
    A.ds_idx = c->ds_idx;
    A.ds_xyz = c->ds_xyz;
    A.n_out = &c->ctr[5];
    A.st = c->reg;
    A.Tp = Tp;
    A.epoch = c->loop_epoch;
    A.st1 = check ? c->reg1 : nullptr;
    A.fc = c->fctl;
    return A;
}

}  // namespace

hipError_t launch_frame_prep(lw_ctx *c, const double *Tp, float t_mid, float inv_dt, const float *xi, float inv_v,
                             float min_range, float max_range, int32_t ret_mask, bool check, const float *xi_dev) {
    PrepArgs A = prep_args(c, Tp, t_mid, inv_dt, xi, inv_v, min_range, max_range, ret_mask, check);
    A.xi_dev = xi_dev;
    if (c->prep_sync && c->frame_bound > 0 && c->frame_bound <= frame_prep_grid_max()) {
        // one workgroup per PG_THREADS points the frame can hold
        const int G = (int)std::min<int64_t>(PG_MAX_G, (c->frame_bound + PG_THREADS - 1) / PG_THREADS);
        hipLaunchKernelGGL(k_frame_prep_grid, dim3(G), dim3(PG_THREADS), 0, c->stream, A, c->prep_sync);
    } else {
        hipLaunchKernelGGL(k_frame_prep, dim3(1), dim3(PREP_THREADS), 0, c->stream, A);
    }
    return hipGetLastError();
}

hipError_t launch_prior_guess(lw_ctx *c, const float *y) {
    hipLaunchKernelGGL(k_prior_guess, dim3(1), dim3(1), 0, c->stream, c->fctl, y, c->chain_tp, &c->ctr[6],
                       (int32_t)c->frame_cap);
    return hipGetLastError();
}

hipError_t launch_prior_fallback(lw_ctx *c) {
    hipLaunchKernelGGL(k_prior_fallback, dim3(1), dim3(256), 0, c->stream, c->fctl, c->reg, c->reg_fb, c->chain_tp);
    return hipGetLastError();
}

hipError_t launch_frame_prep_fallback(lw_ctx *c, float t_mid, float inv_dt, float inv_v, float min_range,
                                      float max_range, int32_t ret_mask) {
    PrepArgs A = prep_args(c, c->chain_tp, t_mid, inv_dt, nullptr, inv_v, min_range, max_range, ret_mask, false);
    A.xi_dev = c->fctl->xi0;
    A.gate = &c->fctl->fb;
    if (c->prep_sync && c->frame_bound > 0 && c->frame_bound <= frame_prep_grid_max()) {
        const int G = (int)std::min<int64_t>(PG_MAX_G, (c->frame_bound + PG_THREADS - 1) / PG_THREADS);
        hipLaunchKernelGGL(k_frame_prep_grid, dim3(G), dim3(PG_THREADS), 0, c->stream, A, c->prep_sync);
    } else {
        hipLaunchKernelGGL(k_frame_prep, dim3(1), dim3(PREP_THREADS), 0, c->stream, A);
    }
    return hipGetLastError();
}

hipError_t launch_frame_check(lw_ctx *c) {
    hipLaunchKernelGGL(k_frame_check, dim3(1), dim3(64), 0, c->stream, c->reg, c->reg1, c->fctl, c->ctr);
    return hipGetLastError();
}

hipError_t launch_keyframe_check(lw_ctx *c) {
    hipLaunchKernelGGL(k_keyframe_check, dim3(1), dim3(1), 0, c->stream, c->reg, c->fctl, c->ctr, c->chain_counts);
    return hipGetLastError();
}

}  // namespace lw
