// Context and launch helpers shared by the kernel translation units and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "lw_common.h"
#include "lw_gn.h"

namespace lw {

constexpr int BLOCK = 256;           // threads per block of the streaming kernels
constexpr int ICP_BLOCK = 64;        // one wave per accumulation workgroup
constexpr int ICP_MAX_GROUPS = 2048; // partial rows kept for the cross-workgroup reduction
constexpr int LW_PROF_COUNTERS = 32; // device measurement counters (lw_ctx::prof_dev)
constexpr int MAX_LOOP_ROWS = 256;   // searching workgroups (partial rows) of the persistent registration
constexpr int LOOP_ROW_WORDS = 256;  // 64-bit words of one tagged partial row of the persistent registration
constexpr int LOOP_TAGGED_ROWS = MAX_LOOP_ROWS + MAX_LOOP_ROWS / 16;  // its rows, then one batch row per 16 rows
// Registration starts between two clears of the persistent registration's row tags: the tag's epoch has 25
// bits and advances at most once a start, so it never wraps between clears.
constexpr uint32_t LOOP_EPOCH_LIMIT = 1u << 24;
constexpr int SCAN_THREADS = 1024;

// Device-resident state of one registration (one Gauss-Newton loop).
struct RegState {
    double T[16];       // current pose, float64
    double T_pred[16];  // the prediction the prior pulls toward
    double H[36];       // H of the last accepted accumulation (for the condition number)
    double acc[64];     // last reduced accumulator (ACC layout)
    // the persistent loop's pose for its searching workgroups: entry k = tag << 32 | float bits of Tf[k],
    // tag = passes completed, with bit 31 set when the loop stops (one 64-bit store each, so a reader that
    // sees the tag sees the value written with it)
    unsigned long long pub[16];
    double rms;
    double rms0;        // the residual of the first pass (at the prediction)
    double cond;
    float Tf[16];       // the current pose rounded to float32, as the accumulation kernels use it
    int32_t done;       // 1: converged or stopped, later passes return at once
    int32_t iters;
    int32_t corr;
    int32_t n_degen;
    int32_t status;     // last step status (gn::STEP_*)
    uint32_t ticket;    // workgroups finished in the current pass (one launch per pass)
    int32_t n;          // thinned points the registration ran on
    unsigned long long t_rel;  // device wall clock at the current pass's start (measurement only)
    float kappa[64];    // robust-kernel scale per iteration (odom.py:767)
};

// One frame's chain of device steps and the decisions between them (lw_odom_frame): the inputs the host
// uploads, then what the decision kernels write.
struct FrameCtl {
    double T_pred[16];      // the prediction
    double T_prev_inv[16];  // the inverse of the previous frame's pose (has_prev)
    double T_key_inv[16];   // the inverse of the last keyframe's pose (has_key)
    double kf_dist, kf_angle;           // keyframe thresholds: metres, degrees
    double redeskew_rot, redeskew_trans;
    float xi0[6];           // the twist the frame is deskewed with first
    int32_t min_points, min_corr, has_prev, has_key, redeskew, frame_cap;
    // the batched chain's (lw_batch_frame): a warm-up frame (inserted at T_pred unregistered), the first
    // registration's gate (1 unless warm), the frame's point count
    int32_t warm, run1, m_in, pad_in;
    // the learned prior's frame (lw_odom_frame_prior): T_pred and xi0 above hold the constant-velocity
    // prediction and twist; the guess is prior_T exp(d), d = hist0 + y sd_t + mu_t (y: the network's output);
    // a first residual (rms0) above rms_gate hands the frame back to the constant-velocity prediction
    double prior_T[16];
    double hist0[6], sd_t[6], mu_t[6];
    double rms_gate;
    int32_t prior_on, pad_p;
    // written on the device
    double T_cv[16];        // the constant-velocity prediction, kept while T_pred holds the guess
    float xi_cv[6];
    int32_t guess, fb;      // the first registration started from the learned guess; the fallback ran
    double d[6];            // the learned increment
    double T1[16];          // the first registration's pose (the second's prediction)
    double T[16];           // the frame's pose
    float xi1[6];           // the twist the frame ends deskewed with
    float T32[16];          // the pose a keyframe insert uses
    int32_t m, skipped, ok1, redo, ok, keyframe;
    int32_t err;            // 1: the frame held more points than the preparation takes; 2: a grid barrier's wait ran out
};
constexpr size_t FRAME_CTL_INPUT = offsetof(FrameCtl, T_cv);  // bytes the host uploads

// One frame's deskew, thinning and registration start in one workgroup (k_frame_prep, lw_odom.hip).
struct PrepArgs {
    const float *raw, *t;  // the frame: raw points and times
    const float *xi_dev;   // given: the deskew twist on the device (in place of xi)
    const int32_t *gate;   // given: the preparation runs only when *gate is nonzero
    const uint32_t *attr;
    float *xyz;            // the deskewed points
    const int32_t *m_dev;  // the frame's appended points (capped at frame_cap)
    int32_t frame_cap;
    float t_mid, inv_dt;
    float xi[6];           // the deskew twist (the first preparation)
    float inv_v, min_range, max_range;
    int32_t ret_mask;
    int32_t ds_mask;       // the global thinning table (the fallback)
    int64_t *ds_keys;
    int32_t *ds_idx;
    float *ds_xyz;         // the thinned points, in index order
    int32_t *n_out;        // their count
    struct RegState *st;   // the registration state to set up
    const double *Tp;      // the prediction (16 doubles) and the robust-kernel scales (64 floats)
    uint32_t *epoch;       // the persistent loop's row-tag epoch (null: the batched chain, which has no loop)
    struct RegState *st1;  // given: the frame check runs first, and the rest only when it asks for a re-deskew
    struct FrameCtl *fc;
};

// The many-workgroup frame preparation's grid barrier (k_frame_prep_grid, lw_odom.hip): arrivals counted on
// one counter that never resets, the base the next launch starts from, and each workgroup's thinned points.
constexpr int PG_MAX_G = 64;
struct PrepSync {
    unsigned long long base, count;
    int32_t total[PG_MAX_G];
};

// One sequence's keyframe decision in a batch (k_keyframe_check's inputs).
struct KeySeq {
    const RegState *st;
    FrameCtl *fc;
    const int32_t *ctr;
    int32_t *counts;
};

__device__ inline PrepArgs rebased(const void *b, PrepArgs A) {
    A.raw = via_arg(b, A.raw);
    A.t = via_arg(b, A.t);
    A.attr = via_arg(b, A.attr);
    A.xyz = via_arg(b, A.xyz);
    A.m_dev = via_arg(b, A.m_dev);
    A.ds_keys = via_arg(b, A.ds_keys);
    A.ds_idx = via_arg(b, A.ds_idx);
    A.ds_xyz = via_arg(b, A.ds_xyz);
    A.n_out = via_arg(b, A.n_out);
    A.st = via_arg(b, A.st);
    A.Tp = via_arg(b, A.Tp);
    A.epoch = via_arg(b, A.epoch);
    A.st1 = via_arg(b, A.st1);
    A.fc = via_arg(b, A.fc);
    return A;
}

__device__ inline KeySeq rebased(const void *b, KeySeq d) {
    d.st = via_arg(b, d.st);
    d.fc = via_arg(b, d.fc);
    d.ctr = via_arg(b, d.ctr);
    d.counts = via_arg(b, d.counts);
    return d;
}

// The parameters the per-voxel precision is computed with (odom.py:336-368).
struct PrecParams {
    float voxel, noise2, tight2, alpha;
};

}  // namespace lw

struct lw_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t cap = 0;  // map slots (power of two)
    int32_t mask = 0;
    int64_t frame_cap = 0;
    int64_t ds_cap = 0;
    int64_t vox_cap = 0;
    // map
    lw::Slot *slots = nullptr;
    uint32_t *occ = nullptr;
    int64_t occ_words = 0;
    int64_t *vox_key = nullptr;
    lw::VoxMom *vox_mom = nullptr;
    int32_t *vox_slot = nullptr;  // slot of each voxel id
    lw::VoxPrec *vox_prec = nullptr;
    int32_t *vox_nin = nullptr;
    int32_t *vox_quota = nullptr;
    int32_t *vox_dirty = nullptr;
    int32_t *dirty_list = nullptr;
    int32_t *touched = nullptr;
    int32_t *ctr = nullptr;  // [0] voxels claimed (stats[0]) [1] points without a slot (stats[1])
                             // [2] voxel ids handed out [3] dirty count [4] touched count [5] thinned count
                             // [6] frame count [7] scan base
    // per-point scratch, frame_cap long
    int32_t *pt_slot = nullptr;
    int32_t *pt_vid = nullptr;
    uint8_t *pt_take = nullptr;
    // frame buffers (bound by the caller)
    float *f_raw = nullptr, *f_xyz = nullptr, *f_t = nullptr;
    uint32_t *f_attr = nullptr;
    int32_t *f_frame = nullptr;
    // thinning
    int64_t *ds_keys = nullptr;
    int32_t *ds_idx = nullptr;
    float *ds_xyz = nullptr;
    // block scan scratch
    int32_t *blk_cnt = nullptr;
    int32_t *blk_off = nullptr;
    int64_t max_blocks = 0;
    // registration
    lw::RegState *reg = nullptr;
    lw::RegState *reg1 = nullptr;   // the first registration of a frame, kept by lw_odom_frame's check
    lw::FrameCtl *fctl = nullptr;   // lw_odom_frame's inputs and decisions
    // reg, reg1, chain_tp (the prediction and robust-kernel scales lw_odom_frame uploads, right before fctl),
    // fctl and chain_counts (the map's counts after lw_odom_frame's insert) live in this one block, so that
    // lw_odom_frame uploads once and reads back once
    char *chain = nullptr;
    double *chain_tp = nullptr;
    int32_t *chain_counts = nullptr;
    int64_t frame_bound = 0;  // points appended since the last lw_frame_reset (at most frame_cap): bounds ctr[6]
    bool count_reset = false;  // lw_frame_reset's zeroing of ctr[6], left to the next append or reader of it
    hipEvent_t chain_ev = nullptr;  // lw_odom_frame's readback (the host waits for it, not for the insert's tail)
    double *partials = nullptr;
    uint32_t *loop_epoch = nullptr;  // the persistent registration's row-tag epoch (one word)
    lw::PrepSync *prep_sync = nullptr;  // the many-workgroup frame preparation's barrier and counts
    lw::RegState *reg_fb = nullptr;     // lw_odom_frame_prior: the guess's registration, kept when the fallback runs
    uint64_t *loop_rows = nullptr;   // LOOP_TAGGED_ROWS x LOOP_ROW_WORDS: the persistent registration's tagged rows
    uint32_t reg_starts = 0;         // launches that may start a registration since the rows were last cleared
    int32_t *best_slot = nullptr;
    int64_t *dtmp = nullptr;  // small uploads: poses, keys (frame_cap + 8 entries)
    // pinned host staging
    void *pin = nullptr;
    size_t pin_bytes = 0;
    // live measurement of the accumulation launches: HIP event pairs around each launch, and device
    // counters [0] point-iterations [1] neighbour lookups [2] occupied neighbours hit [3] matched points
    // [4] launches that did work; with prof_phases set (their clock reads and waits slow the loop a
    // little) also [5] [6] steps on the wave / in one lane [8..11] pass phases (search, reduce, step,
    // passes) [12..15] the first searcher's phases (wake, search and terms, row and flag, passes)
    // [16..20] wave-step parts (set-up, factorisations, substitution, pose update, steps) [21..23] the
    // first searcher's neighbour search, terms, row sum [24..26] every searcher's wake-up and row flag after
    // the pose's publication, and the searcher-passes timed [27] the current pass's latest row (running max)
    // [28] [29] the latest row per pass, summed, and the passes
    bool prof_on = false;
    bool prof_phases = false;
    hipEvent_t *prof_ev = nullptr;
    int32_t prof_cap = 0, prof_used = 0;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
    unsigned long long *prof_dev = nullptr;  // LW_PROF_COUNTERS measurement counters
    // batched localisation: the work-queue counter, and the last captured launch (replayed while the
    // arguments stay the same)
    uint32_t *loc_queue = nullptr;
    hipGraph_t loc_graph = nullptr;
    hipGraphExec_t loc_exec = nullptr;
    unsigned char loc_key[512] = {0};
    // the localisation's neighbourhood bitmap (launch_nb_build, NB_SCALE bits a slot), allocated by the first
    // localisation launch and rebuilt before one whenever the map changed since the last build
    uint32_t *nb_occ = nullptr;
    bool nb_stale = true;
    // the persistent registration: co-resident workgroups per accumulation form (0 = not yet queried), and
    // whether cooperative launches work on this device
    int32_t loop_max_groups[3] = {0, 0, 0};
    int32_t loop_capacity[3] = {0, 0, 0};  // workgroups of each form the device holds at once
    bool loop_ok = true;
    bool loop_plain = true;  // plain launches of the loop where its grid is at most half the capacity
    // parameters the cached per-voxel precision was computed with, and whether an insert may have queued
    // voxels for a refresh since the last one
    bool dirty_maybe = false;
    bool prec_valid = false;
    float prec_voxel = 0, prec_noise2 = 0, prec_tight2 = 0, prec_alpha = 0;
};

namespace lw {

struct Mat4f {
    float m[16];
};

__device__ inline Mat4f load_mat4f(const float *p) {
    Mat4f T;
    for (int k = 0; k < 16; ++k) T.m[k] = p[k];
    return T;
}

// A registration's initial state (odom.py:758-763): the pose P (16 doubles) as current pose and prediction,
// the per-iteration robust-kernel scale, H = I, counters cleared; with the persistent loop's published pose
// cleared and (epoch given) its row-tag epoch advanced. Thread tid of nth takes part.
__device__ inline void reg_state_begin(RegState *st, const double *P, const float *kappa, uint32_t *epoch, int tid,
                                       int nth) {
    if (tid < 64) st->kappa[tid] = kappa[tid];
    if (epoch && tid == 0) *epoch += 1u;
    if (tid < 16) st->pub[tid] = 0ull;
    if (tid != 0) return;
    for (int k = 0; k < 16; ++k) {
        st->T[k] = P[k];
        st->T_pred[k] = P[k];
        st->Tf[k] = (float)P[k];
    }
    for (int k = 0; k < 36; ++k) st->H[k] = (k % 7 == 0) ? 1.0 : 0.0;  // np.eye(6) (odom.py:763)
    st->rms = 0.0;
    st->rms0 = 0.0;
    st->cond = 1.0;
    st->done = 0;
    st->iters = 0;
    st->corr = 0;
    st->n_degen = 0;
    st->status = 0;
    st->ticket = 0;
    st->n = 0;
}

hipError_t launch_map_reset(lw_ctx *c);
constexpr int64_t NB_SCALE = 2;  // bits of the neighbourhood bitmap per map slot
// The neighbourhood bitmap c->nb_occ (NB_SCALE * occ_words words): cleared, then, for every key k within one voxel
// on each axis of a voxel of the map, the bits slot_hash(k) and nb_hash2(k), each & (NB_SCALE * cap - 1), set.
hipError_t launch_nb_build(lw_ctx *c);
// n_dev: when given, the points to insert are its count capped at n (n sizes the grids). gate: when given, a
// device flag that must be nonzero for the insert to happen. T_dev: the pose on the device (replaces T).
// refresh: when given, the touched voxels' precision is computed at once with these parameters instead of
// being queued for the next refresh.
hipError_t launch_map_insert(lw_ctx *c, const float *xyz, const uint32_t *attr, int32_t n, const int32_t *n_dev,
                             const Mat4f &T, float inv_v, float voxel, float max_range, int32_t ret_mask,
                             const int32_t *gate = nullptr, const float *T_dev = nullptr,
                             const PrecParams *refresh = nullptr);
// The same insert in two parts: the head (claims, per-voxel counts) leaves the map's counts final, and writes
// them to counts when given; the tail (quotas, accumulation, finish) completes the moments and precision.
struct InsertArgs {
    const float *xyz;
    const uint32_t *attr;
    int32_t n;
    const int32_t *n_dev;
    Mat4f T;
    float inv_v, voxel, max_range;
    int32_t ret_mask;
    const int32_t *gate;
    const float *T_dev;
    const PrecParams *refresh;
    int32_t *counts;
};
// The arrays of one registration map and its insert scratch, as the insert kernels take them.
struct MapRefs {
    Slot *slots;
    uint32_t *occ;
    int64_t *vox_key;
    int32_t *vox_slot;
    VoxMom *vox_mom;
    VoxPrec *vox_prec;
    int32_t *ctr;
    int32_t *pt_slot, *pt_vid;
    uint8_t *pt_take;
    int32_t *vox_nin, *touched, *vox_quota, *vox_dirty, *dirty_list;
    int64_t vox_cap;
    int32_t mask;
};
MapRefs map_refs(lw_ctx *c);
// Every pointer of a descriptor re-derived from the descriptor array b (a kernel argument) with via_arg.
__device__ inline MapRefs rebased(const void *b, MapRefs M) {
    M.slots = via_arg(b, M.slots);
    M.occ = via_arg(b, M.occ);
    M.vox_key = via_arg(b, M.vox_key);
    M.vox_slot = via_arg(b, M.vox_slot);
    M.vox_mom = via_arg(b, M.vox_mom);
    M.vox_prec = via_arg(b, M.vox_prec);
    M.ctr = via_arg(b, M.ctr);
    M.pt_slot = via_arg(b, M.pt_slot);
    M.pt_vid = via_arg(b, M.pt_vid);
    M.pt_take = via_arg(b, M.pt_take);
    M.vox_nin = via_arg(b, M.vox_nin);
    M.touched = via_arg(b, M.touched);
    M.vox_quota = via_arg(b, M.vox_quota);
    M.vox_dirty = via_arg(b, M.vox_dirty);
    M.dirty_list = via_arg(b, M.dirty_list);
    return M;
}
// One sequence's insert in a batch: its map, its points (n, or the count at n_dev capped at n), its float32
// pose on the device, a gate that must be nonzero, and where the map's counts go once the claims are done.
struct InsSeq {
    MapRefs M;
    const float *xyz;
    const uint32_t *attr;
    int32_t n;
    const int32_t *n_dev;
    const float *T_dev;
    const int32_t *gate;
    int32_t *counts;
};
__device__ inline InsSeq rebased(const void *b, InsSeq d) {
    d.M = rebased(b, d.M);
    d.xyz = via_arg(b, d.xyz);
    d.attr = via_arg(b, d.attr);
    d.n_dev = via_arg(b, d.n_dev);
    d.T_dev = via_arg(b, d.T_dev);
    d.gate = via_arg(b, d.gate);
    d.counts = via_arg(b, d.counts);
    return d;
}
// Every sequence's insert (D: S descriptors on the device), touched voxels refreshed with q at once; n_bound
// bounds every sequence's points.
hipError_t launch_insert_seq(hipStream_t s, const InsSeq *D, int32_t S, int32_t n_bound, float inv_v, float voxel,
                             float max_range, int32_t ret_mask, const PrecParams &q);
hipError_t launch_insert_head(lw_ctx *c, const InsertArgs &a);
// S frame preparations (one workgroup each, D: S PrepArgs on the device) and S keyframe decisions.
hipError_t launch_frame_prep_seq(hipStream_t s, const PrepArgs *D, int32_t S);
hipError_t launch_keyframe_check_seq(hipStream_t s, const KeySeq *D, int32_t S);
hipError_t launch_insert_tail(lw_ctx *c, const InsertArgs &a);
hipError_t launch_refresh(lw_ctx *c, bool all, const PrecParams &q);
hipError_t launch_map_find(lw_ctx *c, const int64_t *keys_dev, int32_t n, int32_t *out_dev);

// reset: the frame's count starts from 0 (a reset the caller deferred), else it grows from its value.
hipError_t launch_frame_append(lw_ctx *c, const float *sx, const uint32_t *sa, const float *st, int32_t n,
                               float inv_dt, int32_t k_min, int32_t k_open, bool reset);
// m < 0: the open frame's point count on the device (ctr[6], capped at frame_cap). gate: when given, a device
// flag that must be nonzero for the kernels to do anything. xi_dev: the twist on the device (replaces xi).
hipError_t launch_deskew(lw_ctx *c, int32_t m, float t_mid, float inv_dt, const float *xi,
                         const float *xi_dev = nullptr, const int32_t *gate = nullptr);
hipError_t launch_thin(lw_ctx *c, int32_t m, float inv_v, float min_range, float max_range, int32_t ret_mask,
                       const int32_t *gate = nullptr);
hipError_t launch_fill_i32(lw_ctx *c, int32_t *p, int64_t n, int32_t v);

struct IcpPass {
    int32_t n;
    int32_t reach;
    float inv_v, voxel, max_d2, kappa;
    int32_t it;
    int32_t gn;  // 1: run the Gauss-Newton step in the last workgroup
    gn::StepParams sp;
    int32_t want_slots;
};
hipError_t launch_icp_pass(lw_ctx *c, const IcpPass &p, int acc_mode);

// The whole Gauss-Newton loop in one cooperative launch (see lw_icp.hip); n_dev: device count of thinned
// points, n_bound: an upper bound known on the host (sizes the grid). Returns hipErrorNotSupported when the
// device cannot run it (the caller then queues one launch per pass).
struct RegLoop {
    int32_t max_iter, wide_iters, min_corr;
    float inv_v, voxel, max_d2_r1, max_d2_r2;
    float kappa[64];
    gn::StepParams sp;
};
hipError_t launch_register_loop(lw_ctx *c, const int32_t *n_dev, int32_t n_bound, const RegLoop &L, int acc_mode,
                                const int32_t *gate = nullptr);
// One sequence's registration in a batch: its thinned points and their count, its map, its registration state
// (set up by the frame preparation), and a gate that must be nonzero for it to run.
struct SeqReg {
    const float *ds;
    const int32_t *n_dev;
    const Slot *slots;
    const uint32_t *occ;
    const VoxPrec *vox_prec;
    RegState *st;
    const int32_t *gate;
    int32_t mask;
    int32_t pad;
};
__device__ inline SeqReg rebased(const void *b, SeqReg d) {
    d.ds = via_arg(b, d.ds);
    d.n_dev = via_arg(b, d.n_dev);
    d.slots = via_arg(b, d.slots);
    d.occ = via_arg(b, d.occ);
    d.vox_prec = via_arg(b, d.vox_prec);
    d.st = via_arg(b, d.st);
    d.gate = via_arg(b, d.gate);
    return d;
}
// S registrations (D: S descriptors on the device), one workgroup each, every Gauss-Newton pass and step inside
// the workgroup. prof: the measurement counters, or null.
hipError_t launch_register_seq(hipStream_t s, const SeqReg *D, int32_t S, const RegLoop &L, int acc_mode,
                               unsigned long long *prof);
int32_t icp_max_groups(int64_t frame_cap);  // partial rows one pass over frame_cap points can write
hipError_t launch_reg_begin(lw_ctx *c, const double *T_pred_dev, const int32_t *gate = nullptr,
                            const double *T_dev = nullptr);
// Before launches that may start up to `starts` registrations of c (each advances the row-tag epoch once at
// most): clears the tagged rows and the epoch on c's stream once LOOP_EPOCH_LIMIT starts have accumulated.
hipError_t loop_epoch_guard(lw_ctx *c, uint32_t starts);
// lw_odom.hip: after a frame's first registration, and after its second
hipError_t launch_frame_check(lw_ctx *c);
hipError_t launch_keyframe_check(lw_ctx *c);
// lw_odom.hip: deskew (with xi), thinning and the registration's start from Tp (the prediction and the
// robust-kernel scales) for the open frame in one single-workgroup launch; the frame may hold at most
// frame_prep_max() points. check: after the first registration -- the frame check first, then the rest
// (deskew with the check's twist, the registration's start from its pose) only when it asks for a re-deskew.
int32_t frame_prep_max();
// lw_odom.hip: the learned prior in the frame chain: its guess into the chain's prediction (y: the network's
// output on the device), the residual gate after the first registration, and the preparation gated on it
// (deskew with the constant-velocity twist, thinning, the registration's start from the constant-velocity
// prediction).
hipError_t launch_prior_guess(lw_ctx *c, const float *y);
hipError_t launch_prior_fallback(lw_ctx *c);
hipError_t launch_frame_prep_fallback(lw_ctx *c, float t_mid, float inv_dt, float inv_v, float min_range,
                                      float max_range, int32_t ret_mask);
int32_t frame_prep_grid_max();  // the many-workgroup form's bound (launch_frame_prep takes it when c->prep_sync is set
                                // and the frame's bound is within it)
hipError_t launch_frame_prep(lw_ctx *c, const double *Tp, float t_mid, float inv_dt, const float *xi, float inv_v,
                             float min_range, float max_range, int32_t ret_mask, bool check,
                             const float *xi_dev = nullptr);

}  // namespace lw
