/*
 * livox_warp_amd: MI355X (gfx950) scan-to-map odometry kernels behind a plain C ABI.
 *
 * The reference (nhaumann/livox-warp) has no C boundary on this path: its surface is the Python class
 * livox_warp.odom.Odometry (python/livox_warp/odom.py:458-803), which launches Warp kernels. Each
 * entry point below replaces one of those launches (or one host step) and cites it. The Python class
 * that keeps the reference's call surface binds these entry points with ctypes.
 *
 * Conventions
 *   - Every function returns 0 on success, a negative value for a HIP error (-(hipError_t)) or for a
 *     bad argument (LW_EINVAL), and never throws.
 *   - Point, attribute and timestamp pointers are DEVICE pointers, borrowed for the call. Matrices and
 *     results are HOST memory.
 *   - One context owns its map, its frame-sized work buffers and one HIP stream; all work is queued on
 *     that stream. A context is driven by one thread at a time. Nothing is allocated after lw_create
 *     (readbacks go through pinned buffers allocated there).
 *   - Row-major 4x4 matrices: sensor-to-world poses, as in the reference.
 */
#ifndef LIVOX_WARP_AMD_H
#define LIVOX_WARP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LW_EINVAL (-1000)
#define LW_ENOTBOUND (-1001)
#define LW_ELOOP (-1002) /* the persistent registration loop stopped waiting (a bounded wait ran out) */
#define LW_ENOLOOP (-1003) /* no cooperative launch on this device: drive the frame step by step instead */
#define LW_ERANGE (-1004)  /* a size beyond what an allocation-free entry point's fixed scratch holds */
#define LW_ACC_STRIDE 64

typedef struct lw_ctx lw_ctx;

/* Registration parameters: the Odometry tuning attributes of odom.py:499-525 that _register reads. */
typedef struct lw_reg_params {
    double damping;    /* Levenberg-Marquardt fraction of the diagonal (odom.py:512) */
    double degen_thr;  /* degenerate-direction threshold, 0 = off (odom.py:516) */
    double prior_info; /* pull toward the prediction, 0 = off (odom.py:517) */
    float reg_voxel;   /* registration voxel edge, m (odom.py:501) */
    float min_range;   /* thinning range filter, m (odom.py:507-508) */
    float max_range;
    float noise;       /* range noise sigma, m (odom.py:509) */
    float alpha;       /* weight of loose directions (odom.py:511) */
    int32_t ret_mask;  /* returns used (odom.py:510) */
    int32_t max_iter;  /* odom.py:502 */
    int32_t min_corr;  /* odom.py:506 */
    int32_t wide_iters; /* iterations with the 5x5x5 search (odom.py:86) */
    int32_t acc_mode;  /* 0: float64 accumulation; 1: compensated float32; 2: naive float32 */
} lw_reg_params;

/* Registration outcome: odom.py:797-803 plus the thinned count (odom.py:754-755). */
typedef struct lw_reg_result {
    double T[16];
    double rms;
    double cond;
    int32_t ok;
    int32_t iters;
    int32_t corr;
    int32_t n_ds;
    int32_t degen;
    int32_t weak;
    double rms0; /* the residual of the first pass, at the prediction (no counterpart in the reference: the
                    learned prior's fallback test) */
    int32_t m;   /* the frame points registered: m as passed, or, for m < 0, the open frame's count read back
                    with the result (no counterpart in the reference: it saves the host a read of f_count) */
    int32_t pad;
} lw_reg_result;

/* Library version (major * 10000 + minor * 100 + patch). */
int lw_version(void);

/* Odometry.__init__ buffers (odom.py:461-497): a registration map of reg_slots slots (a power of two,
 * odom.py:468; 64 to 2^27, a 4 GiB table) and frame-sized work buffers of frame_cap points (up to 2^26). stream: a hipStream_t, or NULL for the
 * default stream. The map starts empty (odom.py:527 -> reset). */
int lw_create(lw_ctx **out, int64_t reg_slots, int64_t frame_cap, void *stream);
int lw_destroy(lw_ctx *ctx);
int lw_set_stream(lw_ctx *ctx, void *stream);
/* Wait for everything queued on the context's stream (Odometry.synchronize, odom.py:534-537). */
int lw_synchronize(lw_ctx *ctx);

/* The caller-owned frame buffers the frame kernels fill (odom.py:477-481): f_raw and f_xyz are
 * float32 x 3 per point, f_attr uint32, f_t float32, f_frame int32, each frame_cap long. */
int lw_bind_frame(lw_ctx *ctx, float *f_raw, float *f_xyz, uint32_t *f_attr, float *f_t, int32_t *f_frame);

/* Forget the map (odom.py:563-567). */
int lw_map_reset(lw_ctx *ctx);
/* f_count = 0 (odom.py:579, 589, 614, 636). */
int lw_frame_reset(lw_ctx *ctx);

/* k_frame_append (odom.py:97-123): append the staged points with k_min <= floor(max(t,0)*inv_dt) <=
 * k_open to the open frame. Points keep their input order (Warp's serial CPU order); the count keeps
 * counting past frame_cap, as the reference's does, and the surplus is dropped. */
int lw_frame_append(lw_ctx *ctx, const float *sx, const uint32_t *sa, const float *st, int64_t n, float inv_dt,
                    int32_t k_min, int32_t k_open);
/* f_count read on the stream (odom.py:539-543 with odom.py:607, 645). */
int lw_frame_count(lw_ctx *ctx, int32_t *out);

/* k_deskew (odom.py:163-178, gpu.py:268-284): f_xyz[i] = Rot(theta s) f_raw[i] + rho s,
 * s = clamp((f_t[i] - t_mid) * inv_dt, -1, 1); xi = (rho, theta). m < 0: the open frame's point count as it
 * stands on the device (f_count capped at frame_cap), with no read back to the host. */
int lw_deskew(lw_ctx *ctx, int32_t m, float t_mid, float inv_dt, const float *xi);

/* Odometry._downsample (odom.py:731-737 with k_ds_insert, odom.py:126-160): keep the first point (in
 * index order) of each thinning voxel of edge 1/inv_v in the sensor frame, after the range / tag /
 * return filters; the kept points stay in index order. Writes the count to *n_ds (waits on the stream). */
int lw_thin(lw_ctx *ctx, int32_t m, float inv_v, float min_range, float max_range, int32_t ret_mask,
            int32_t *n_ds);
/* Copy the first n thinned points (float32 x 3) to host memory. */
int lw_thin_read(lw_ctx *ctx, float *out, int32_t n);
/* Upload n points (float32 x 3, host) as the thinned set (tests and batch drivers). */
int lw_thin_load(lw_ctx *ctx, const float *xyz, int32_t n);

/* k_reg_insert (odom.py:186-230), Odometry._insert (odom.py:719-727): add n points posed by T to the
 * map. xyz/attr are device pointers, or NULL for the bound frame buffers f_xyz / f_attr; with the frame
 * buffers, n < 0 takes the open frame's count on the device. A voxel stops accumulating at 4096 points;
 * exactly its first points in index order contribute. */
int lw_map_insert(lw_ctx *ctx, const float *xyz, const uint32_t *attr, int64_t n, const float *T, float inv_v,
                  float voxel, float max_range, int32_t ret_mask);
/* reg_stats (odom.py:474, read at odom.py:698): out[0] voxels claimed, out[1] points without a slot. */
int lw_map_stats(lw_ctx *ctx, int32_t *out);
/* The map in the reference's per-slot layout (odom.py:470-473), copied to host memory: keys[cap],
 * rsum[cap*3], router[cap*9], rcnt[cap]. Any pointer may be NULL. */
int lw_map_read(lw_ctx *ctx, int64_t *keys, float *rsum, float *router, int32_t *rcnt);
/* The map's voxels in the order they were claimed (a voxel keeps its place until the map is reset), in the
 * reference's record layout: keys[v], rsum[v*3], router[v*9], rcnt[v] for the first min(voxels, cap); *count
 * receives the number of voxels. Any array may be NULL. Waits for the context's stream. No counterpart in the
 * reference (a compact form of lw_map_read: V records instead of one per slot). */
int lw_map_read_voxels(lw_ctx *ctx, int32_t cap, int64_t *keys, float *rsum, float *router, int32_t *rcnt,
                       int32_t *count);
/* Slot of a key, or -1 (gpu.find_slot, gpu.py:254-265), evaluated on the device for n keys. */
int lw_map_find(lw_ctx *ctx, const int64_t *keys_host, int32_t n, int32_t *slots_host);

/* One k_icp + k_reduce pass (odom.py:275-384, Odometry._accumulate odom.py:739-750) over the n thinned
 * points at pose T (float32 4x4): out receives the 64 reduced float64 entries of the ACC layout
 * (odom.py:72-82). best_slot (host, may be NULL) receives each point's chosen slot or -1. */
int lw_icp_accum(lw_ctx *ctx, int32_t n, const float *T, float inv_v, float voxel, int32_t reach, float max_d2,
                 float noise2, float tight2, float alpha, float kappa, int32_t acc_mode, double *out,
                 int32_t *best_slot);

/* One frame of Odometry._finalize after the warm-up (odom.py:644-692, 712-727), queued as one chain with one
 * readback: deskew the open frame (its count on the device) with xi, register it from T_pred; when that is
 * trusted, the re-deskew test against the previous pose (xi_new, the body twist of T_prev^-1 T, in float32,
 * against redeskew_rot / redeskew_trans) and, if it fires, deskew with xi_new and register again from the first
 * pose; then the keyframe test against the last keyframe (kf_dist metres, kf_angle degrees; every trusted frame
 * when has_key is 0) and, if it fires, the insert of the frame at its pose. A frame under min_points points
 * keeps T_pred untrusted (its registration runs and is discarded). The host supplies the inverses of the
 * previous and keyframe poses as it computes them. No counterpart in the reference (its host decides between
 * device steps). */
typedef struct {
    double T_pred[16];
    double T_prev_inv[16]; /* valid when has_prev */
    double T_key_inv[16];  /* valid when has_key */
    double kf_dist, kf_angle;
    double redeskew_rot, redeskew_trans;
    float xi[6];           /* the constant-velocity twist (zeros when deskewing is off) */
    float t_mid, inv_dt;
    int32_t min_points, has_prev, has_key, redeskew; /* redeskew: deskewing and re-deskewing both on */
} lw_frame_in;

typedef struct {
    lw_reg_result reg[3];  /* the frame's registrations in order (the first `registrations` are valid) */
    double T[16];          /* the frame's pose */
    double d[6];           /* lw_odom_frame_prior: the learned increment the guess used (when guess) */
    float xi[6];           /* the twist the frame ended deskewed with */
    int32_t m, skipped, ok, registrations, keyframe, map_voxels, map_overflow;
    int32_t guess;         /* lw_odom_frame_prior: the first registration started from the learned guess */
    int32_t fallback;      /* lw_odom_frame_prior: its first residual failed the gate; the frame went on from the
                              constant-velocity prediction */
    int32_t pad;
} lw_frame_out;

int lw_odom_frame(lw_ctx *ctx, const lw_reg_params *p, const lw_frame_in *in, lw_frame_out *out);

/* The same chain with the learned pose-increment prior in place of the constant-velocity guess (Odometry.prior,
 * LearnedPrior.guess / accept): in->T_pred and in->xi hold the constant-velocity prediction and twist. When the
 * frame holds at least min_points points, the first registration starts from prior_T exp(d), d = hist0 +
 * y sd_t + mu_t (y: the network's 6 outputs, on the device: queue the forward pass on the context's stream
 * first), with the points deskewed by d; if its first residual (rms0) exceeds rms_gate the frame is deskewed with
 * the constant-velocity twist and registered from the constant-velocity prediction instead; the re-deskew test,
 * the keyframe test and the insert follow as in lw_odom_frame. Up to three registrations (out->reg in order);
 * out->guess, out->fallback and out->d report the prior's part. A frame beyond the one-launch preparation
 * (more than 32768 points appended, or a thinning table under 4096 slots) returns LW_ERANGE with nothing queued.
 * No counterpart in the reference (the learned prior is this build's extension). */
typedef struct {
    const float *y;        /* device: the network's output for this frame (6 float32) */
    double T[16];          /* the previous frame's pose */
    double hist0[6];       /* the last registered increment (body twist) */
    double sd_t[6], mu_t[6]; /* the output's scale and offset */
    double rms_gate;       /* fallback x the running median of earlier first residuals (HUGE_VAL: none yet) */
} lw_frame_prior;

int lw_odom_frame_prior(lw_ctx *ctx, const lw_reg_params *p, const lw_frame_in *in, const lw_frame_prior *prior,
                        lw_frame_out *out);

/* ---- The batched odometry: one frame of each of S independent sequences per call. No counterpart in the
 * reference (it runs one Odometry per stream); each sequence's frame goes through lw_odom_frame's chain --
 * deskew, thinning, registration, the re-deskew test and second registration, the keyframe test, the insert into
 * that sequence's own map -- but every step is one launch for the whole batch (a sequence's registration runs in
 * one workgroup), with one upload and one readback. The S contexts keep their maps and buffers (any per-context
 * entry point can read them between calls); they must be bound and share the batch's stream. */
typedef struct lw_batch lw_batch;
/* A sequence's frame: whole (its m points already cut from the stream, m <= 32768 and <= frame_cap), device
 * pointers to its points (float32 x 3), attribute bits and times; warm: a warm-up frame, inserted at T_pred
 * unregistered (odom.py:669-670). */
typedef struct {
    const float *xyz;
    const uint32_t *attr;
    const float *t;
    int32_t m;
    int32_t warm;
} lw_batch_src;
int lw_batch_create(lw_batch **out, lw_ctx *const *ctxs, int32_t S, void *stream);
int lw_batch_destroy(lw_batch *b);
/* in[s], src[s] -> out[s] for every sequence, as lw_odom_frame (the warm-up frames' registrations count 0). */
int lw_batch_frame(lw_batch *b, const lw_reg_params *p, const lw_frame_in *in, const lw_batch_src *src,
                   lw_frame_out *out);
/* The batch's registration launches measured as lw_prof_* measures a context's (out: 8 doubles). */
int lw_batch_prof_enable(lw_batch *b, int32_t on);
int lw_batch_prof_reset(lw_batch *b);
int lw_batch_prof_read(lw_batch *b, double *out);

/* Odometry._register (odom.py:752-803): thin the m deskewed frame points, then run the Gauss-Newton
 * loop on the device -- accumulation, float64 host step (odom.py:387-455) and convergence test -- with
 * one readback at the end. T_pred: the prediction (float64 4x4). m < 0: the open frame's count on the
 * device, returned in res->m. */
int lw_register(lw_ctx *ctx, int32_t m, const lw_reg_params *p, const double *T_pred, lw_reg_result *res);
/* The same loop over the already-thinned points (lw_thin / lw_thin_load), n of them. */
int lw_register_thinned(lw_ctx *ctx, int32_t n, const lw_reg_params *p, const double *T_pred, lw_reg_result *res);
/* on = 1 (default): a registration is one persistent launch that runs every iteration, a plain launch when
 * its grid is at most half of what the device holds at once, else a cooperative one; on = 2: always the
 * cooperative launch; on = 0: one launch per iteration (the path taken automatically where cooperative
 * launches are unavailable). Same results up to the order of the float64 partial sums. No counterpart in the
 * reference. */
int lw_set_persistent(lw_ctx *ctx, int32_t on);

/* Batched global localisation parameters. */
typedef struct lw_loc_params {
    double damping;    /* 1e-3: fraction of the diagonal added (localize.py:538) */
    float reg_voxel;   /* the prior map's voxel edge, m */
    float noise;       /* range noise sigma for the voxel precision, m (odom.py:509) */
    float alpha;       /* weight of loose directions (odom.py:511) */
    int32_t iters;     /* Gauss-Newton iterations per hypothesis */
    int32_t min_count; /* matches below which a step is skipped (30, localize.py:535) */
    int32_t block;     /* tile: threads per workgroup, 64 / 128 / 256 / 512 */
    int32_t grid;      /* persistent workgroups (0: four per compute unit) */
    int32_t use_graph; /* capture the launch in a hipGraph and replay it while the arguments stay the same
                          (needs a context stream other than the null stream) */
    int32_t chunk;     /* tile: neighbour keys a lane resolves together, 7 or 14 (0: 14) */
} lw_loc_params;

/* GlobalLocalizer._icp (localize.py:517-553, kernels :126-179), batched: n_hyp hypotheses for each of n_queries
 * scans, registered against the context's map in one launch. Correspondence and residual are odom.py's
 * (closest voxel mean in the 27 neighbours, point-to-distribution; odom.py:275-375), the step localize.py's.
 * scans: float32 x 3, concatenated; scan_off: n_queries + 1 offsets; order: the queries in processing order
 * (longest first balances the work); hyp_T / out_T: n_queries * n_hyp row-major 4x4 float64 (query-major);
 * out_score: sum of robust weights of the last pass; out_count: its matched points; out_iters: steps taken.
 * Every array is a DEVICE pointer. Queued on the context's stream; it does not wait. The first call allocates
 * the map's neighbourhood bitmap (two bits per slot), and a call after any change of the map rebuilds it on the
 * stream before the launch (outside a captured graph). */
int lw_loc_batch(lw_ctx *ctx, const float *scans, const int32_t *scan_off, const int32_t *order, int32_t n_queries,
                 const double *hyp_T, int32_t n_hyp, const lw_loc_params *p, double *out_T, double *out_score,
                 int32_t *out_count, int32_t *out_iters);
/* The compiled kernel of a tile (block 64 / 128 / 256 / 512, chunk 7 / 14): info[0] scratch bytes per lane
 * (register spills), info[1] registers per lane, info[2] static LDS bytes, info[3] largest block. The tile
 * autotuner skips tiles that spill. Needs a device. No counterpart in the reference. */
int lw_loc_tile_info(int32_t block, int32_t chunk, int32_t *info);

/* ---- Learned pose-increment prior. No counterpart in the reference: it replaces the constant-velocity guess
 * of Odometry.predict (odom.py:591-592, used by _estimate at odom.py:666) when enabled, and plain ICP from that
 * guess stays the fallback. A 152 -> 256 -> 256 -> 256 -> 6 MLP with GELU, BF16 weights on the matrix cores
 * with FP32 accumulation, trained by the forward / loss / backward / Adam entry points below. Its parameter
 * vector is lw_mlp_param_count() float32s: W1 160x256, b1 256, W2 256x256, b2 256, W3 256x256, b3 256,
 * W4 256x16, b4 16 (row-major [in][out]; W1 rows 152..159 and W4 columns 6..15 stay zero). */
typedef struct lw_mlp lw_mlp;

int lw_mlp_param_count(void);
/* Device buffers for batches of up to batch_cap rows; work is queued on stream (NULL: the null stream). */
int lw_mlp_create(lw_mlp **out, int32_t batch_cap, void *stream);
int lw_mlp_destroy(lw_mlp *m);
int lw_mlp_set_stream(lw_mlp *m, void *stream);
/* Parameters from / to host memory (float32, the layout above); setting them clears the Adam moments. */
int lw_mlp_set_params(lw_mlp *m, const float *host);
int lw_mlp_get_params(lw_mlp *m, float *host);
/* The gradient of the last backward pass (host, the parameter layout). */
int lw_mlp_get_grads(lw_mlp *m, float *host);
/* y (n x 6 float32, device) = MLP(x (n x 152 float32, device, row stride ldx)); keeps the activations. */
int lw_mlp_forward(lw_mlp *m, const float *x, int64_t ldx, int32_t n, float *y);
/* *loss (device float) = mean of (y - target)^2 over the n x 6 entries; keeps d loss / d y for the backward. */
int lw_mlp_loss(lw_mlp *m, const float *y, const float *target, int32_t n, float *loss);
/* Gradients of the loss with respect to every parameter, through the last forward (same n). */
int lw_mlp_backward(lw_mlp *m, int32_t n);
/* One Adam step (bias-corrected; step counts from 1) on the float32 masters, refreshing the BF16 copy. */
int lw_mlp_adam(lw_mlp *m, float lr, float beta1, float beta2, float eps, int32_t step);
/* One training step -- forward, loss into *loss (device float), backward, Adam at `step` -- on x (n x 152
 * float32, device, row stride ldx) and target (n x 6 float32, device). use_graph = 1 (with a stream other than
 * the null stream): the step's launches are captured once into a hipGraph and replayed while x, ldx, target,
 * loss and n stay the same, Adam's scalars set before each replay; same results as the eager sequence. */
int lw_mlp_train_step(lw_mlp *m, const float *x, int64_t ldx, const float *target, int32_t n, float lr, float beta1,
                      float beta2, float eps, int32_t step, float *loss, int32_t use_graph);

/* The 152 inputs of the open frame's m points (the bound f_raw buffer, sensor frame) into out (device):
 * motion[24] as given, then 64 range bins over [0, r_max) and 64 elevation bins over [-e_max, e_max), each
 * count * 64 / m. m < 0: the open frame's count, read on the device (up to frame_cap). */
int lw_frame_features(lw_ctx *ctx, int32_t m, float r_max, float e_max, const float *motion, float *out);

/* Live measurement of the accumulation launches (for throughput and roofline reports): lw_prof_enable(ctx, 1)
 * turns on HIP events around each accumulation launch and device counters, lw_prof_enable(ctx, 2) also the
 * device clock stamps that lw_prof_phases reports (they slow the loop a little), 0 turns both off;
 * lw_prof_read returns, since
 * lw_prof_reset: out[0] launches timed, out[1] their total duration (ms), out[2] point-iterations, out[3]
 * neighbour lookups, out[4] occupied neighbours hit, out[5] matched points, out[6] launches that did work,
 * out[7] reserved. No counterpart in the reference (its timing is host wall time, odom.py:644, 705). */
int lw_prof_enable(lw_ctx *ctx, int32_t on);
int lw_prof_reset(lw_ctx *ctx);
int lw_prof_read(lw_ctx *ctx, double *out);
/* Where the registration passes spend their time, since lw_prof_reset, from device wall-clock stamps
 * (microseconds, summed over passes; lw_prof_enable level 2): out[0] from the pass's start to the moment the
 * workgroup that finishes the pass has every partial row (search, terms, rows), out[1] the rest of the
 * cross-workgroup reduction, out[2] the float64 step and the pose's publication; out[3] the passes timed.
 * The first searching workgroup's view of the persistent loop's passes: out[4] from the publication of the
 * pose to its wake-up, out[5] its search and terms, out[6] its row write and flag; out[7] the passes it
 * timed. out[8] steps taken on the whole wave, out[9] steps taken by the general one-lane routine
 * (a degenerate direction, an indefinite system). The wave steps' parts: out[10] set-up (matrices, basis,
 * prior), out[11] the two Cholesky factorisations (degeneracy test and damped system, side by side), out[12]
 * the substitution, out[13] the pose update; out[14] the wave steps timed. The first searching workgroup's
 * search and terms (out[5]) in parts: out[15] point to closest means (neighbour search), out[16] the terms
 * into its LDS rows, out[17] the sum of its rows; counted with out[7]. Every searching workgroup's passes
 * after the first: out[18] from the pose's publication to its wake-up, out[19] to its row flag; out[20] the
 * searcher-passes timed; out[21] the slowest searcher's row per pass, on average, from the publication;
 * out[22] the first searcher's neighbour phase summed over the narrow (27-voxel) passes alone, out[23] those
 * passes. out holds 24 doubles. No counterpart in the reference. */
int lw_prof_phases(lw_ctx *ctx, double *out);

/* ---- the integration map (the reference's Pipeline map, python/livox_warp/gpu.py) ----------------------
 * The consumer of the odometry's frames: voxel records keyed and probed like the registration map (key,
 * hits, carving misses, float32 sums of x, y, z and reflectivity over the first 4096 hits, first and last
 * time), the device pose table the inserts and rays read, and free-space carving. One context per map, on
 * its own stream (void *: a hipStream_t). Point arrays are device pointers. */
typedef struct lw_imap lw_imap;

/* The raw-return filters of an insert (the keep_raw fields of gpu.py:77-99 View, applied as gpu.py:156-167). */
typedef struct lw_map_view {
    float min_range, max_range;
    int32_t min_refl;
    int32_t noise_mask; /* bit k keeps points whose noise confidence (tag bits 0-1) is k */
    int32_t ret_mask;   /* bit k keeps return k */
} lw_map_view;

/* slots: a power of two in [64, 2^27] (Pipeline(map_capacity), gpu.py:1172-1219); the map starts cleared. */
int lw_imap_create(lw_imap **out, int64_t slots, void *stream);
int lw_imap_destroy(lw_imap *m);
/* Pipeline.clear_map (gpu.py:1304-1312). */
int lw_imap_clear(lw_imap *m);
/* The poses of frame ids idx[0..n) (host arrays; mats: n row-major 4x4 float32 sensor-to-world matrices),
 * written to slot id & 65535 of the device pose table: Pipeline.flush_poses (gpu.py:1346-1365). */
int lw_imap_set_poses(lw_imap *m, const int32_t *idx, const float *mats, int32_t n);
/* Pipeline.map_insert / k_map_insert (gpu.py:1421-1434, 373-413): n staged points (xyz, attribute bits,
 * time, frame id) through their frames' poses into voxels of size voxel. */
int lw_imap_insert(lw_imap *m, const float *sx, const uint32_t *sa, const float *st, const int32_t *sf, int32_t n,
                   const lw_map_view *v, float voxel);
/* Pipeline.carve / k_carve (gpu.py:1436-1452, 470-595): every `every`-th return's ray, up to max_steps voxels
 * and max_range metres, adds `every` misses to the occupied voxels whose surface patch it pierces. */
int lw_imap_carve(lw_imap *m, const float *sx, const uint32_t *sa, const int32_t *sf, int32_t n, float voxel,
                  int32_t every, float max_range, int32_t max_steps);
/* Pipeline.rebase_clock (gpu.py:1367-1372): first and last times of the occupied slots become old - shift. */
int lw_imap_rebase_clock(lw_imap *m, float shift);
/* out[0] voxels claimed, out[1] points that found no slot within 64 probes (Pipeline.map_stats). Syncs. */
int lw_imap_stats(lw_imap *m, int32_t *out);
/* The records as they lie in device memory (bytes = 64 x slots), to host memory / from it. Sync. 64-byte
 * record: int64 key (-1 empty), int32 hits, int32 misses, float sums[4] (x, y, z, reflectivity), float
 * first, float last, 24 bytes unused. */
int lw_imap_read(lw_imap *m, void *out, int64_t bytes);
int lw_imap_load(lw_imap *m, const void *in, int64_t bytes);

/* ---- LVXR recordings (the reference's firing-level reader, python/livox_warp/lvxr.py:76-260) ----------------
 * Decode n_packets point packets of a recording already in device memory (bytes: the whole file) into rows of
 * the firing arrays: packet p's datagram starts at byte pk_off[p] (its 18-byte data header first), holds
 * pk_nf[p] firings of data type pk_type[p] (0-5, 7, 8; dual = 1: the Mid-40 dual profile, two plain records
 * a firing), starts pk_t_ns[p] ns after the recording's time base, and fills rows pk_first[p] ..
 * pk_first[p] + pk_nf[p] - 1. Each row: t (float64 s, firing j at pk_t_ns + j interval_ns), `returns`
 * (1-3) return slots of xyz (float32 m), reflectivity, tag and valid (uint8), and the firing's unit direction
 * (float32, NaN where none is known). All arrays are device pointers; asynchronous on stream (a hipStream_t).
 * Replaces read_firings' decoding (lvxr.py:189-260); the host keeps the framing and the time base. */
/* Host functions (no device): the recording's framing, `host_ns:u64 len:u16 datagram[len]` records from byte
 * `start` of data[0, n) (record.rs:1-5; a truncated last record dropped): offs[k] is datagram k's first byte
 * and lens[k] its length, for the first cap; *count = the number of records (cap 0: count only; more than a
 * nonzero cap: LW_EINVAL). Replaces the framing loop of lvxr.py:77-95 (_packets). */
int lw_lvxr_index(const uint8_t *data, int64_t n, int64_t start, int64_t *offs, int32_t *lens, int64_t cap,
                  int64_t *count);
/* The live path's time base (net.rs Sink, lvxr.py:116-134 _time_base): out[i] = ns of packet i since the base,
 * from its header stamp ts[i] (ns) and its span spans[i] (firings x interval, ns); a stamp below the base
 * starts a new base, a packet later than expected by more than half its span and by at least jump_ns moves
 * the base by the gap. */
int lw_lvxr_clock(const uint64_t *ts, const int64_t *spans, int64_t n, int64_t jump_ns, int64_t *out);
int lw_lvxr_decode(const uint8_t *bytes, const int64_t *pk_off, const int32_t *pk_type, const int32_t *pk_nf,
                   const int64_t *pk_t_ns, const int64_t *pk_first, int32_t n_packets, int32_t dual, int64_t interval_ns,
                   int32_t returns, double *t, float *xyz, uint8_t *refl, uint8_t *tag, uint8_t *valid, float *dir,
                   void *stream);

/* ---- Bulk voxelisation of a prior map (python/livox_warp/prior_map.py:55-123, voxel_downsample :169-203) ----
 * A table of 2^slots_log2 voxels: each point's offset from its voxel's centre, colour and intensity summed
 * into the slot claimed for its key (gpu.py's key and hash, 128 probes: k_vox_insert), then the claimed
 * voxels compacted in slot order into dense means (k_vox_flag, array_scan, k_vox_emit). Work is queued on
 * the table's stream. */
typedef struct lw_vox lw_vox;

int lw_vox_create(lw_vox **out, int32_t slots_log2, void *stream); /* 10 <= slots_log2 <= 30; cleared */
int lw_vox_destroy(lw_vox *v);
int lw_vox_clear(lw_vox *v); /* every slot empty, the stats zero */
/* k_vox_insert over n points: xyz float32 x 3, intensity float32 and rgb float32 x 3 (either may be NULL:
 * zeros); DEVICE pointers. voxel: the edge in metres (the kernel gets 1 / voxel and voxel as float32). */
int lw_vox_insert(lw_vox *v, const float *xyz, const float *intensity, const float *rgb, int64_t n, double voxel);
/* out[0] voxels claimed, out[1] points dropped (no slot within 128 probes); waits for the table's stream. */
int lw_vox_stats(lw_vox *v, int32_t *out);
/* The claimed voxels in slot order into out_xyz (mean position, float32 x 3), out_col (mean r, g, b,
 * intensity) and out_cnt (points): DEVICE arrays of out_cap voxels; *count = the voxels (LW_ERANGE, nothing
 * written, when more than out_cap). */
int lw_vox_emit(lw_vox *v, double voxel, float *out_xyz, float *out_col, int32_t *out_cnt, int64_t out_cap,
                int64_t *count);
/* The table as plain DEVICE arrays of its 2^slots_log2 slots: keys (-1 empty), offset sums (x 3), colour and
 * intensity sums (x 4), counts. For checks; no counterpart in the reference. */
int lw_vox_read(lw_vox *v, int64_t *keys, float *rsum, float *csum, int32_t *cnt);

/* ---- The prior map's dense grid and coarse scoring (python/livox_warp/mapgrid.py, localize.py:55-123) -----
 * Geometry, host pointers read at the call: geo = {origin x, y, z, cell, 1 / cell} (float32), dims = {nx, ny,
 * nz} with nx ny nz < 2^31; cell c = (iz ny + iy) nx + ix. Map points (xyz, n < 2^23) are sorted by cell and
 * start (ncell + 1 int32) is the prefix of the per-cell counts (PriorGrid.prepare, mapgrid.py:215-241). Every
 * other array is a device pointer; each call is asynchronous on stream (a hipStream_t).
 *
 * lw_grid_build: the seeds nn (int32 per cell: the map point nearest the cell centre, -1 beyond the band) and
 * the distance field df (uint8 cm per cell, 255 beyond the band): seed, unpack, the n_steps jump-flood steps
 * (host array, 8 4 2 1 1 in the reference), each reading every seed before writing any, then df. work:
 * device scratch of at least lw_grid_work_bytes(dims) bytes (two 16-byte seed records per cell), 16-byte
 * aligned. Replaces PriorGrid.build's launches (mapgrid.py:275-290). */
int64_t lw_grid_work_bytes(const int32_t *dims);
int lw_grid_build(const float *xyz, int32_t n, const float *geo, const int32_t *dims, const int32_t *steps,
                  int32_t n_steps, int32_t *nn, void *work, int64_t work_bytes, uint8_t *df, void *stream);
/* Distance of each of n points, placed in the map by T (host, row-major 4x4 float32), to the nearest map point
 * (exact within one cell, else through the seeds), 1e3 beyond the band. Replaces PriorGrid.distances /
 * k_changes_T (mapgrid.py:170-193, 292-300). */
int lw_grid_distances(const float *pts, int32_t n, const float *T, const float *geo, const int32_t *dims,
                      const int32_t *start, const int32_t *nn, const float *xyz, float *out, void *stream);
/* The soft count (sum over the n_pts points within tau_cm of the map of 1 - d^2 / tau^2, d from df) of every
 * (position, heading): out[r n_yaw + k] for position p (x, y, z; gravity-aligned, list them x-fastest for
 * speed), r = out_row[p] (device int32 per position, or NULL for r = p), heading yaw0 + k dyaw; -1 where the
 * position is outside the grid or within free_cm of a mapped surface. Replaces k_score_grid
 * (localize.py:55-92). */
int lw_score_grid(const float *pts, int32_t n_pts, const float *pos, int32_t n_pos, int32_t n_yaw, float yaw0,
                  float dyaw, const float *geo, const int32_t *dims, const uint8_t *df, float tau_cm, float free_cm,
                  const int32_t *out_row, float *out, void *stream);
/* The same soft count for a list of (x, y, z, yaw) hypotheses (hyp 16-byte aligned). Replaces k_score_list
 * (localize.py:95-123). */
int lw_score_list(const float *pts, int32_t n_pts, const float *hyp, int32_t n_hyp, const float *geo,
                  const int32_t *dims, const uint8_t *df, float tau_cm, float *out, void *stream);

/* Point-to-plane ICP of k poses (T: device, k row-major float32 4x4, k <= 65535) at once against the grid's
 * map points, normals (nrm) and planarity: one Gauss-Newton accumulation each, into out (device, k x 64
 * doubles in the odometry's ACC layout: H upper triangle 0-20, b 21-26, matched count 27, weighted error 28,
 * weight sum 30). Correspondences within sqrt(max_d2) of the nearest map point; precision n n^T (1 - alpha) +
 * alpha I on planar points, 4 alpha I elsewhere; Geman-McClure weight with scale c2. work: device scratch of
 * lw_plane_work_bytes(k) bytes. Replaces k_icp_batch + k_reduce_batch (localize.py:126-180). */
int64_t lw_plane_work_bytes(int32_t k);
int lw_plane_icp(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo, const int32_t *dims,
                 const int32_t *start, const int32_t *nn, const float *xyz, const float *nrm, const float *planar,
                 float max_d2, float c2, float alpha, void *work, int64_t work_bytes, double *out, void *stream);
/* lw_plane_icp with no workspace argument and no allocation of any kind: the partial rows live in a fixed device
 * array of the library (one per GPU, 16 KiB a pose, LW_PLANE_NOALLOC_MAX poses: 4 MiB, held until the process
 * ends). No counterpart in the reference: it is for callers that cannot allocate -- code captured into a HIP
 * graph, tight serving loops, harnesses that forbid hipMalloc. k above LW_PLANE_NOALLOC_MAX returns LW_ERANGE.
 * Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two streams, two threads) race,
 * so a caller serialises them. The same results as lw_plane_icp, bit for bit. */
#define LW_PLANE_NOALLOC_MAX 256
int lw_plane_icp_noalloc(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo,
                         const int32_t *dims, const int32_t *start, const int32_t *nn, const float *xyz,
                         const float *nrm, const float *planar, float max_d2, float c2, float alpha, double *out,
                         void *stream);
/* Per pose, the scan points whose nearest map point lies within radii[0..2] (host array): counts (device, k x 3
 * int32, zeroed by the call). Replaces k_eval_batch (localize.py:183-206). */
int lw_plane_eval(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo, const int32_t *dims,
                  const int32_t *start, const int32_t *nn, const float *xyz, const float *radii, int32_t *counts,
                  void *stream);

/* Host builds of the device's float64 step and per-voxel precision, for tests on machines without a
 * GPU (same source as the device code). */
int lw_host_gn_step(const double *v, const double *T, const double *T_pred, double damping, double prior_info,
                    double degen_thr, int32_t min_corr, double *delta, double *rms, int32_t *n_degen);
int lw_host_voxel_precision(const float *rsum, const float *router6, int32_t c, float voxel, float noise2,
                            float tight2, float alpha, float *M6, float *Mg6, float *q3);
int lw_host_apply_twist(double *T, const double *delta);
int lw_host_twist_between(const double *T_from, const double *T_to, double *xi);
double lw_host_cond(const double *H);
/* The device's two forms of the float64 step on one reduced accumulator v (64 doubles, the ACC layout) at pose
 * T with prediction T_pred (host arrays): the wave form (the persistent registration's common case) and the
 * one-lane scalar form, each followed by its pose update, in one launch on the current device. out (48
 * doubles): [0, 16) the wave form's new pose, [16, 32) the scalar form's, 32 the wave form's result (1
 * handled, 0 left to the scalar form, -1 too few correspondences), 33 its rms, [34, 40) its step, 40 the scalar
 * form's return code (0 ok), 41 its rms, [42, 48) its step. No counterpart in the reference: a check that the
 * two device forms agree bit for bit. */
int lw_gn_step_check(const double *v, const double *T, const double *T_pred, double damping, double prior_info,
                     double degen_thr, int32_t min_corr, double *out);

#ifdef __cplusplus
}
#endif

#endif
