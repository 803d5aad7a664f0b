// The C ABI of include/livox_warp_amd.h: context lifecycle, argument checks, staging and readbacks.
// Kernels live in lw_map.hip, lw_frame.hip and lw_icp.hip.
#include <climits>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/livox_warp_amd.h"
#include "lw_gn_wave.h"
#include "lw_internal.h"
#include "lw_loc.h"
#include "lw_mlp.h"
#include "lw_prec.h"

#define LW_CK(x)                                        \
    do {                                                \
        hipError_t e_ = (x);                            \
        if (e_ != hipSuccess) return -(int)e_;          \
    } while (0)

namespace {

constexpr size_t PIN_BYTES = 1 << 20;
constexpr size_t PIN_UP = 0;          // uploads (poses)
constexpr size_t PIN_DOWN = 1 << 16;  // readbacks

__global__ void k_fill_i64(int64_t *p, int64_t n, int64_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_set_tf(lw::RegState *st, const float *T) {
    if (threadIdx.x < 16) st->Tf[threadIdx.x] = T[threadIdx.x];
    if (threadIdx.x == 0) {
        st->done = 0;
        st->ticket = 0;
    }
}

template <typename T>
hipError_t dalloc(T **p, size_t count) {
    return hipMalloc((void **)p, count * sizeof(T) + 64);
}

bool pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

void free_all(lw_ctx *c) {
    void *ptrs[] = {c->slots, c->occ, c->vox_key, c->vox_mom, c->vox_slot, c->vox_prec, c->vox_nin, c->vox_quota,
                    c->vox_dirty, c->dirty_list, c->touched, c->ctr, c->pt_slot, c->pt_vid, c->pt_take,
                    c->ds_keys, c->ds_idx, c->ds_xyz, c->blk_cnt, c->blk_off, c->chain, c->partials, c->best_slot,
                    c->dtmp, c->loop_epoch, c->loop_rows, c->prep_sync, c->nb_occ};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c->prof_dev) (void)hipFree(c->prof_dev);
    if (c->loc_queue) (void)hipFree(c->loc_queue);
    if (c->loc_exec) (void)hipGraphExecDestroy(c->loc_exec);
    if (c->loc_graph) (void)hipGraphDestroy(c->loc_graph);
    if (c->prof_ev) {
        for (int i = 0; i < 2 * c->prof_cap; ++i) (void)hipEventDestroy(c->prof_ev[i]);
        delete[] c->prof_ev;
    }
    if (c->pin) (void)hipHostFree(c->pin);
    if (c->chain_ev) (void)hipEventDestroy(c->chain_ev);
}

// Add the elapsed times of the recorded launch pairs (call after the stream has been synchronised).
hipError_t prof_harvest(lw_ctx *c) {
    for (int i = 0; i < c->prof_used; ++i) {
        float ms = 0.0f;
        hipError_t e = hipEventElapsedTime(&ms, c->prof_ev[2 * i], c->prof_ev[2 * i + 1]);
        if (e != hipSuccess) return e;
        c->prof_ms += ms;
        c->prof_launches++;
    }
    c->prof_used = 0;
    return hipSuccess;
}

// The chain block (lw_ctx::chain): [reg | reg1 | chain_tp | fctl | chain_counts], each part 256-byte aligned.
constexpr size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }
constexpr size_t CH_REG1 = align256(sizeof(lw::RegState));
constexpr size_t CH_REGFB = CH_REG1 + align256(sizeof(lw::RegState));
constexpr size_t CH_TP = CH_REGFB + align256(sizeof(lw::RegState));
constexpr size_t CH_TP_BYTES = 16 * sizeof(double) + 64 * sizeof(float);
constexpr size_t CH_FCTL = CH_TP + align256(CH_TP_BYTES);
constexpr size_t CH_COUNTS = CH_FCTL + align256(sizeof(lw::FrameCtl));
constexpr size_t CH_BYTES = CH_COUNTS + 256;
static_assert(CH_BYTES <= PIN_DOWN, "the chain block fits the pinned readback area");

char *pin_up(lw_ctx *c) { return (char *)c->pin + PIN_UP; }
char *pin_down(lw_ctx *c) { return (char *)c->pin + PIN_DOWN; }

// The per-voxel precision must match the query's parameters: refresh what the inserts queued, or every
// voxel when the parameters changed.
hipError_t ensure_precision(lw_ctx *c, float voxel, float noise2, float tight2, float alpha) {
    bool same = c->prec_valid && c->prec_voxel == voxel && c->prec_noise2 == noise2 && c->prec_tight2 == tight2 &&
                c->prec_alpha == alpha;
    if (same && !c->dirty_maybe) return hipSuccess;  // nothing queued since the last refresh
    hipError_t e = lw::launch_refresh(c, !same, lw::PrecParams{voxel, noise2, tight2, alpha});
    c->prec_valid = true;
    c->prec_voxel = voxel;
    c->prec_noise2 = noise2;
    c->prec_tight2 = tight2;
    c->prec_alpha = alpha;
    return e;
}

// Apply a deferred lw_frame_reset before anything reads the frame's count (ctr[6]).
hipError_t frame_count_ready(lw_ctx *c) {
    if (!c->count_reset) return hipSuccess;
    c->count_reset = false;
    return lw::launch_fill_i32(c, &c->ctr[6], 1, 0);
}

int read_ctr(lw_ctx *c, int idx, int32_t *out) {
    LW_CK(hipMemcpyAsync(pin_down(c), &c->ctr[idx], sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    std::memcpy(out, pin_down(c), sizeof(int32_t));
    return 0;
}

}  // namespace

extern "C" {

int lw_version(void) { return 100; }

int lw_create(lw_ctx **out, int64_t reg_slots, int64_t frame_cap, void *stream) {
    if (!out || !pow2(reg_slots) || reg_slots < 64 || reg_slots > (1ll << 27) || frame_cap <= 0 ||
        frame_cap > (1ll << 26))
        return LW_EINVAL;
    lw_ctx *c = new (std::nothrow) lw_ctx();
    if (!c) return LW_EINVAL;
    int rc = 0;
    auto fail = [&](hipError_t e) {
        free_all(c);
        delete c;
        return -(int)e;
    };
    hipError_t e;
    if ((e = hipGetDevice(&c->device)) != hipSuccess) return fail(e);
    c->stream = (hipStream_t)stream;
    c->cap = reg_slots;
    c->mask = (int32_t)(reg_slots - 1);
    c->vox_cap = reg_slots;
    c->occ_words = (reg_slots + 31) / 32;
    c->frame_cap = frame_cap;
    int64_t ds = 1;
    while (ds < frame_cap) ds <<= 1;  // odom.py:485
    c->ds_cap = ds;
    c->max_blocks = (frame_cap + lw::BLOCK - 1) / lw::BLOCK;
    if ((e = dalloc(&c->slots, c->cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->occ, c->occ_words)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_key, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_mom, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_slot, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_prec, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_nin, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_quota, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->vox_dirty, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->dirty_list, c->vox_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->touched, frame_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->ctr, 64)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->pt_slot, frame_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->pt_vid, frame_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->pt_take, frame_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->ds_keys, c->ds_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->ds_idx, c->ds_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->ds_xyz, 3 * frame_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->blk_cnt, c->max_blocks)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->blk_off, c->max_blocks)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->chain, CH_BYTES)) != hipSuccess) return fail(e);
    c->reg = reinterpret_cast<lw::RegState *>(c->chain);
    c->reg1 = reinterpret_cast<lw::RegState *>(c->chain + CH_REG1);
    c->reg_fb = reinterpret_cast<lw::RegState *>(c->chain + CH_REGFB);
    c->chain_tp = reinterpret_cast<double *>(c->chain + CH_TP);
    c->fctl = reinterpret_cast<lw::FrameCtl *>(c->chain + CH_FCTL);
    c->chain_counts = reinterpret_cast<int32_t *>(c->chain + CH_COUNTS);
    if ((e = dalloc(&c->partials, (size_t)lw::icp_max_groups(frame_cap) * 128)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->loop_epoch, 1)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->prep_sync, 1)) != hipSuccess) return fail(e);
    if ((e = lw::launch_fill_i32(c, reinterpret_cast<int32_t *>(c->prep_sync), sizeof(lw::PrepSync) / 4, 0)) !=
        hipSuccess)
        return fail(e);
    if ((e = dalloc(&c->loop_rows, (size_t)lw::LOOP_TAGGED_ROWS * lw::LOOP_ROW_WORDS)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->best_slot, frame_cap)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->dtmp, frame_cap + 64)) != hipSuccess) return fail(e);
    if ((e = dalloc(&c->loc_queue, 16)) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc(&c->pin, PIN_BYTES, hipHostMallocDefault)) != hipSuccess) return fail(e);
    c->pin_bytes = PIN_BYTES;
    if ((e = hipEventCreateWithFlags(&c->chain_ev, hipEventDisableTiming)) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_fill_i64, dim3(1024), dim3(256), 0, c->stream, c->ds_keys, c->ds_cap, (int64_t)-1);
    if ((e = lw::launch_fill_i32(c, c->ds_idx, c->ds_cap, INT_MAX)) != hipSuccess) return fail(e);
    if ((e = lw::launch_fill_i32(c, c->ctr, 64, 0)) != hipSuccess) return fail(e);
    c->reg_starts = lw::LOOP_EPOCH_LIMIT;  // the first registration start clears the rows and the epoch
    if ((e = lw::launch_map_reset(c)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(e);
    (void)rc;
    *out = c;
    return 0;
}

int lw_destroy(lw_ctx *c) {
    if (!c) return LW_EINVAL;
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return 0;
}

int lw_set_stream(lw_ctx *c, void *stream) {
    if (!c) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(c->stream));
    c->stream = (hipStream_t)stream;
    return 0;
}

int lw_synchronize(lw_ctx *c) {
    if (!c) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(c->stream));
    return 0;
}

int lw_bind_frame(lw_ctx *c, float *f_raw, float *f_xyz, uint32_t *f_attr, float *f_t, int32_t *f_frame) {
    if (!c || !f_raw || !f_xyz || !f_attr || !f_t || !f_frame) return LW_EINVAL;
    c->f_raw = f_raw;
    c->f_xyz = f_xyz;
    c->f_attr = f_attr;
    c->f_t = f_t;
    c->f_frame = f_frame;
    return 0;
}

int lw_map_reset(lw_ctx *c) {
    if (!c) return LW_EINVAL;
    LW_CK(lw::launch_map_reset(c));
    c->prec_valid = false;
    return 0;
}

int lw_frame_reset(lw_ctx *c) {
    if (!c) return LW_EINVAL;
    c->count_reset = true;  // applied by the next append (which then starts its count from 0) or reader
    c->frame_bound = 0;
    return 0;
}

int lw_frame_append(lw_ctx *c, const float *sx, const uint32_t *sa, const float *st, int64_t n, float inv_dt,
                    int32_t k_min, int32_t k_open) {
    if (!c || n < 0 || (n > 0 && (!sx || !sa || !st))) return LW_EINVAL;
    if (!c->f_raw) return LW_ENOTBOUND;
    const int64_t chunk = c->max_blocks * (int64_t)lw::BLOCK;
    for (int64_t a = 0; a < n; a += chunk) {  // chunks in order keep the points in input order
        int32_t m = (int32_t)((n - a) < chunk ? (n - a) : chunk);
        LW_CK(lw::launch_frame_append(c, sx + 3 * a, sa + a, st + a, m, inv_dt, k_min, k_open, c->count_reset));
        c->count_reset = false;
    }
    c->frame_bound = c->frame_bound + n < c->frame_cap ? c->frame_bound + n : c->frame_cap;
    return 0;
}

int lw_frame_count(lw_ctx *c, int32_t *out) {
    if (!c || !out) return LW_EINVAL;
    LW_CK(frame_count_ready(c));
    return read_ctr(c, 6, out);
}

int lw_deskew(lw_ctx *c, int32_t m, float t_mid, float inv_dt, const float *xi) {
    if (!c || !xi || m > c->frame_cap) return LW_EINVAL;
    if (!c->f_raw) return LW_ENOTBOUND;
    LW_CK(frame_count_ready(c));
    LW_CK(lw::launch_deskew(c, m, t_mid, inv_dt, xi));
    return 0;
}

int lw_thin(lw_ctx *c, int32_t m, float inv_v, float min_range, float max_range, int32_t ret_mask, int32_t *n_ds) {
    if (!c || !n_ds || m < 0 || m > c->frame_cap) return LW_EINVAL;
    if (!c->f_raw) return LW_ENOTBOUND;
    LW_CK(lw::launch_thin(c, m, inv_v, min_range, max_range, ret_mask));
    return read_ctr(c, 5, n_ds);
}

int lw_thin_read(lw_ctx *c, float *out, int32_t n) {
    if (!c || !out || n < 0 || n > c->frame_cap) return LW_EINVAL;
    LW_CK(hipMemcpyAsync(out, c->ds_xyz, 12 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    return 0;
}

int lw_thin_load(lw_ctx *c, const float *xyz, int32_t n) {
    if (!c || n < 0 || n > c->frame_cap || (n > 0 && !xyz)) return LW_EINVAL;
    LW_CK(hipMemcpyAsync(c->ds_xyz, xyz, 12 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    return 0;
}

int lw_map_insert(lw_ctx *c, const float *xyz, const uint32_t *attr, int64_t n, const float *T, float inv_v,
                  float voxel, float max_range, int32_t ret_mask) {
    if (!c || !T) return LW_EINVAL;
    LW_CK(frame_count_ready(c));
    const int32_t *n_dev = nullptr;
    if (!xyz && !attr) {
        if (!c->f_xyz) return LW_ENOTBOUND;
        if (n > c->frame_cap) return LW_EINVAL;
        if (n < 0) {  // the open frame's count, on the device
            n = c->frame_cap;
            n_dev = &c->ctr[6];
        }
        xyz = c->f_xyz;
        attr = c->f_attr;
    } else if (!xyz || !attr || n < 0) {
        return LW_EINVAL;
    }
    lw::Mat4f M;
    std::memcpy(M.m, T, sizeof(M.m));
    for (int64_t a = 0; a < n; a += c->frame_cap) {  // in index order, so the cap keeps the first points
        int32_t m = (int32_t)((n - a) < c->frame_cap ? (n - a) : c->frame_cap);
        LW_CK(lw::launch_map_insert(c, xyz + 3 * a, attr + a, m, n_dev, M, inv_v, voxel, max_range, ret_mask));
    }
    return 0;
}

int lw_map_stats(lw_ctx *c, int32_t *out) {
    if (!c || !out) return LW_EINVAL;
    LW_CK(hipMemcpyAsync(pin_down(c), c->ctr, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    std::memcpy(out, pin_down(c), 2 * sizeof(int32_t));
    return 0;
}

int lw_map_read(lw_ctx *c, int64_t *keys, float *rsum, float *router, int32_t *rcnt) {
    if (!c) return LW_EINVAL;
    int32_t nvox = 0;
    int rc = read_ctr(c, 2, &nvox);
    if (rc) return rc;
    if (nvox > c->vox_cap) nvox = (int32_t)c->vox_cap;
    std::vector<lw::Slot> slots(c->cap);
    std::vector<lw::VoxMom> mom(nvox > 0 ? nvox : 1);
    LW_CK(hipMemcpy(slots.data(), c->slots, c->cap * sizeof(lw::Slot), hipMemcpyDeviceToHost));
    if (nvox > 0) LW_CK(hipMemcpy(mom.data(), c->vox_mom, nvox * sizeof(lw::VoxMom), hipMemcpyDeviceToHost));
    const int map3[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int64_t s = 0; s < c->cap; ++s) {
        const lw::Slot &sl = slots[s];
        bool have = sl.key != -1 && sl.vid >= 0 && sl.vid < nvox;
        if (keys) keys[s] = sl.key;
        const lw::VoxMom *m = have ? &mom[sl.vid] : nullptr;
        if (rsum)
            for (int a = 0; a < 3; ++a) rsum[3 * s + a] = m ? m->s[a] : 0.0f;
        if (router)
            for (int a = 0; a < 9; ++a) router[9 * s + a] = m ? m->o[map3[a]] : 0.0f;
        if (rcnt) rcnt[s] = m ? m->cnt : 0;
    }
    return 0;
}

int lw_map_read_voxels(lw_ctx *c, int32_t cap, int64_t *keys, float *rsum, float *router, int32_t *rcnt,
                       int32_t *count) {
    if (!c || cap < 0 || !count) return LW_EINVAL;
    int32_t nvox = 0;
    if (int rc = read_ctr(c, 2, &nvox)) return rc;  // after everything queued on the context's stream
    if (nvox > c->vox_cap) nvox = (int32_t)c->vox_cap;
    *count = nvox;
    const int32_t n = nvox < cap ? nvox : cap;
    if (n == 0) return 0;
    std::vector<int64_t> k(n);
    std::vector<lw::VoxMom> mom(n);
    LW_CK(hipMemcpyAsync(k.data(), c->vox_key, n * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipMemcpyAsync(mom.data(), c->vox_mom, n * sizeof(lw::VoxMom), hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    const int map3[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int32_t v = 0; v < n; ++v) {
        if (keys) keys[v] = k[v];
        if (rsum)
            for (int a = 0; a < 3; ++a) rsum[3 * (int64_t)v + a] = mom[v].s[a];
        if (router)
            for (int a = 0; a < 9; ++a) router[9 * (int64_t)v + a] = mom[v].o[map3[a]];
        if (rcnt) rcnt[v] = mom[v].cnt;
    }
    return 0;
}

int lw_map_find(lw_ctx *c, const int64_t *keys_host, int32_t n, int32_t *slots_host) {
    if (!c || n < 0 || n > c->frame_cap || (n > 0 && (!keys_host || !slots_host))) return LW_EINVAL;
    if (n == 0) return 0;
    LW_CK(hipMemcpyAsync(c->dtmp, keys_host, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
    LW_CK(lw::launch_map_find(c, c->dtmp, n, c->pt_slot));
    LW_CK(hipMemcpyAsync(slots_host, c->pt_slot, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    return 0;
}

int lw_icp_accum(lw_ctx *c, int32_t n, const float *T, float inv_v, float voxel, int32_t reach, float max_d2,
                 float noise2, float tight2, float alpha, float kappa, int32_t acc_mode, double *out,
                 int32_t *best_slot) {
    if (!c || !T || !out || n < 0 || n > c->frame_cap || reach < 0 || reach > 8 || acc_mode < 0 || acc_mode > 2)
        return LW_EINVAL;
    LW_CK(ensure_precision(c, voxel, noise2, tight2, alpha));
    std::memcpy(pin_up(c), T, 16 * sizeof(float));
    LW_CK(hipMemcpyAsync(c->dtmp, pin_up(c), 16 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_set_tf, dim3(1), dim3(64), 0, c->stream, c->reg, (const float *)c->dtmp);
    LW_CK(hipGetLastError());
    lw::IcpPass p{};
    p.n = n;
    p.reach = reach;
    p.inv_v = inv_v;
    p.voxel = voxel;
    p.max_d2 = max_d2;
    p.kappa = kappa;
    p.it = 0;
    p.gn = 0;
    p.want_slots = best_slot != nullptr;
    LW_CK(lw::launch_icp_pass(c, p, acc_mode));
    LW_CK(hipMemcpyAsync(pin_down(c), c->reg->acc, 64 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (best_slot && n > 0)
        LW_CK(hipMemcpyAsync(best_slot, c->best_slot, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    LW_CK(prof_harvest(c));
    std::memcpy(out, pin_down(c), 64 * sizeof(double));
    return 0;
}

}  // extern "C"

namespace {

// The Gauss-Newton loop's schedule (odom.py:761-767), computed in float64 and passed as float32 like the
// reference's launch, and the per-voxel precision it needs refreshed.
int reg_setup(lw_ctx *c, const lw_reg_params *p, lw::RegLoop &L) {
    const double voxel = p->reg_voxel;
    const float noise2 = (float)((double)p->noise * (double)p->noise);
    const double tf = 0.6 * voxel / std::sqrt(12.0);  // TIGHT_FRACTION (odom.py:89, 761)
    const float tight2 = (float)(tf * tf);
    LW_CK(ensure_precision(c, (float)voxel, noise2, tight2, p->alpha));
    L = lw::RegLoop{};
    L.max_iter = p->max_iter;
    L.wide_iters = p->wide_iters;
    L.min_corr = p->min_corr;
    L.inv_v = (float)(1.0 / voxel);
    L.voxel = (float)voxel;
    L.max_d2_r1 = (float)((1.5 * voxel) * (1.5 * voxel));
    L.max_d2_r2 = (float)((2.5 * voxel) * (2.5 * voxel));
    for (int it = 0; it < 64; ++it) {
        const double kap = 6.0 * std::pow(0.7, it);
        L.kappa[it] = (float)(kap > 2.0 ? kap : 2.0);
    }
    L.sp.damping = p->damping;
    L.sp.prior_info = p->prior_info;
    L.sp.degen_thr = p->degen_thr;
    L.sp.min_corr = p->min_corr;
    L.sp.pad = 0;
    return 0;
}

// The result of a registration from its final state (odom.py:756-757, 797-803): below min_corr thinned points
// the prediction, untrusted; else the pose, the residual, the condition number of H (a 6x6 eigen problem on
// the host: microseconds) and the weak flag.
void fill_result(const lw::RegState &st, const lw_reg_params *p, lw_reg_result *res, bool cond_done = false) {
    const int32_t m = res->m;
    std::memset(res, 0, sizeof(*res));
    res->m = m;
    res->n_ds = st.n;
    res->cond = 1.0;
    if (st.n < p->min_corr) {
        std::memcpy(res->T, st.T_pred, 16 * sizeof(double));
        return;
    }
    const double cond = cond_done ? st.cond : (st.corr >= p->min_corr ? lw::gn::cond_of(st.H) : 1.0);
    std::memcpy(res->T, st.T, 16 * sizeof(double));
    res->rms = st.rms;
    res->rms0 = st.rms0;
    res->cond = cond;
    res->corr = st.corr;
    res->iters = st.iters;
    res->ok = st.corr >= p->min_corr ? 1 : 0;
    res->degen = st.n_degen;
    res->weak = (cond < 1e-3 || st.n_degen > 0) ? 1 : 0;  // WEAK_COND (odom.py:90, 801)
}

// Odometry._register from the thinned points on (odom.py:756-803). The thinned count is on the device
// (n_dev) and n_bound bounds it on the host; n_known >= 0 when the host knows it exactly. m_frame: the frame
// points for res->m, or < 0 to read the open frame's count back with the result.
int run_registration(lw_ctx *c, const int32_t *n_dev, int32_t n_bound, int32_t n_known, int32_t m_frame,
                     const lw_reg_params *p, const double *T_pred, lw_reg_result *res) {
    std::memset(res, 0, sizeof(*res));
    std::memcpy(res->T, T_pred, 16 * sizeof(double));
    res->cond = 1.0;
    res->m = m_frame;
    if (n_known >= 0 && n_known < p->min_corr) {  // odom.py:756-757
        res->n_ds = n_known;
        return 0;
    }
    lw::RegLoop L;
    if (int rc = reg_setup(c, p, L)) return rc;
    std::memcpy(pin_up(c), T_pred, 16 * sizeof(double));
    std::memcpy(pin_up(c) + 16 * sizeof(double), L.kappa, 64 * sizeof(float));
    LW_CK(hipMemcpyAsync(c->dtmp, pin_up(c), 16 * sizeof(double) + 64 * sizeof(float), hipMemcpyHostToDevice,
                         c->stream));
    LW_CK(lw::loop_epoch_guard(c, 1));
    LW_CK(lw::launch_reg_begin(c, (const double *)c->dtmp));
    hipError_t e = hipErrorNotSupported;
    if (c->loop_ok) {
        e = lw::launch_register_loop(c, n_dev, n_bound, L, p->acc_mode);
        if (e == hipErrorNotSupported) {
            c->loop_ok = false;     // no persistent launch on this device: one launch per pass from now on
            (void)hipGetLastError();
        } else if (e != hipSuccess) {
            return -(int)e;
        }
    }
    int32_t n = n_known;
    if (e != hipSuccess) {
        if (n < 0) {
            int rc = read_ctr(c, 5, &n);
            if (rc) return rc;
        }
        if (n < p->min_corr) {
            res->n_ds = n;
            return 0;
        }
        lw::IcpPass q{};
        q.n = n;
        q.inv_v = L.inv_v;
        q.voxel = L.voxel;
        q.gn = 1;
        q.sp = L.sp;
        for (int it = 0; it < p->max_iter; ++it) {  // odom.py:764-796
            q.it = it;
            q.reach = it < p->wide_iters ? 2 : 1;
            q.max_d2 = q.reach == 2 ? L.max_d2_r2 : L.max_d2_r1;
            q.kappa = L.kappa[it < 64 ? it : 63];
            LW_CK(lw::launch_icp_pass(c, q, p->acc_mode));
        }
    }
    constexpr size_t M_AT = (sizeof(lw::RegState) + 63) & ~(size_t)63;  // the frame count beside the state
    LW_CK(hipMemcpyAsync(pin_down(c), c->reg, sizeof(lw::RegState), hipMemcpyDeviceToHost, c->stream));
    if (m_frame < 0)
        LW_CK(hipMemcpyAsync(pin_down(c) + M_AT, &c->ctr[6], sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipStreamSynchronize(c->stream));
    LW_CK(prof_harvest(c));
    lw::RegState st;
    std::memcpy(&st, pin_down(c), sizeof(st));
    if (m_frame < 0) {
        int32_t mf;
        std::memcpy(&mf, pin_down(c) + M_AT, sizeof(mf));
        res->m = mf < c->frame_cap ? mf : (int32_t)c->frame_cap;
    }
    if (st.status < 0) return LW_ELOOP;
    if (n < 0) n = st.n;
    res->n_ds = n;
    if (n < p->min_corr) return 0;
    // odom.py:797-800, on the host copy of the last accepted H (a 6x6 eigen problem: microseconds here)
    st.cond = st.corr >= p->min_corr ? lw::gn::cond_of(st.H) : 1.0;
    std::memcpy(res->T, st.T, 16 * sizeof(double));
    res->rms = st.rms;
    res->rms0 = st.rms0;
    res->cond = st.cond;
    res->corr = st.corr;
    res->iters = st.iters;
    res->ok = st.corr >= p->min_corr ? 1 : 0;
    res->degen = st.n_degen;
    res->weak = (st.cond < 1e-3 || st.n_degen > 0) ? 1 : 0;  // WEAK_COND (odom.py:90, 801)
    return 0;
}

// The registration state keeps 64 robust-kernel scales, so max_iter is at most 64.
bool bad_params(const lw_reg_params *p) {
    return !p || p->max_iter < 0 || p->max_iter > 64 || p->acc_mode < 0 || p->acc_mode > 2 ||
           !(p->reg_voxel > 0.0f);
}

}  // namespace

extern "C" {

int lw_register_thinned(lw_ctx *c, int32_t n, const lw_reg_params *p, const double *T_pred, lw_reg_result *res) {
    if (!c || bad_params(p) || !T_pred || !res || n < 0 || n > c->frame_cap) return LW_EINVAL;
    LW_CK(lw::launch_fill_i32(c, &c->ctr[8], 1, n));
    return run_registration(c, &c->ctr[8], n, n, n, p, T_pred, res);
}

int lw_loc_batch(lw_ctx *c, const float *scans, const int32_t *scan_off, const int32_t *order, int32_t n_queries,
                 const double *hyp_T, int32_t n_hyp, const lw_loc_params *p, double *out_T, double *out_score,
                 int32_t *out_count, int32_t *out_iters) {
    if (!c || !p || !scans || !scan_off || !order || !hyp_T || !out_T || !out_score || !out_count || !out_iters ||
        n_queries <= 0 || n_hyp <= 0 || (int64_t)n_queries * n_hyp > (1ll << 30) || p->iters < 0 || p->iters > 64 ||
        !(p->reg_voxel > 0.0f) || (p->block != 64 && p->block != 128 && p->block != 256 && p->block != 512) ||
        (p->chunk != 0 && p->chunk != 7 && p->chunk != 14))
        return LW_EINVAL;
    const double voxel = p->reg_voxel;
    const float noise2 = (float)((double)p->noise * (double)p->noise);
    const double tf = 0.6 * voxel / std::sqrt(12.0);
    LW_CK(ensure_precision(c, (float)voxel, noise2, (float)(tf * tf), p->alpha));
    lw::LocLaunch L{};
    L.scans = scans;
    L.scan_off = scan_off;
    L.order = order;
    L.hyp_T = hyp_T;
    L.out_T = out_T;
    L.out_score = out_score;
    L.out_count = out_count;
    L.out_iters = out_iters;
    L.prof = c->prof_on ? c->prof_dev : nullptr;
    L.Q = n_queries;
    L.H = n_hyp;
    L.iters = p->iters;
    L.min_count = p->min_count;
    L.block = p->block;
    L.chunk = p->chunk ? p->chunk : 14;
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device);
    L.grid = p->grid > 0 ? p->grid : 4 * cus;
    L.inv_v = (float)(1.0 / voxel);
    L.max_d2 = (float)((1.5 * voxel) * (1.5 * voxel));  // reach 1 (odom.py:766)
    L.damping = p->damping;
    for (int it = 0; it < 16; ++it) {  // the odometry's robust-kernel schedule (odom.py:767)
        const double kap = 6.0 * std::pow(0.7, it);
        L.kappa[it] = (float)(kap > 2.0 ? kap : 2.0);
    }
    // the neighbourhood bitmap the kernel tests before each point's search: built on first use and again after
    // any change of the map, before the launch (and outside a captured graph)
    if (!c->nb_occ) {
        LW_CK(dalloc(&c->nb_occ, lw::NB_SCALE * c->occ_words));
        c->nb_stale = true;
    }
    if (c->nb_stale) {
        LW_CK(lw::launch_nb_build(c));
        c->nb_stale = false;
    }
    const bool timed = c->prof_on && c->prof_used < c->prof_cap;
    if (timed) LW_CK(hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream));
    bool launched = false;
    if (p->use_graph && c->stream != nullptr) {  // a captured launch needs a stream other than the null stream
        unsigned char key[512] = {0};
        static_assert(sizeof(lw::LocLaunch) <= sizeof(key), "key");
        std::memcpy(key, &L, sizeof(L));
        if (!c->loc_exec || std::memcmp(key, c->loc_key, sizeof(key)) != 0) {
            if (c->loc_exec) (void)hipGraphExecDestroy(c->loc_exec);
            if (c->loc_graph) (void)hipGraphDestroy(c->loc_graph);
            c->loc_exec = nullptr;
            c->loc_graph = nullptr;
            LW_CK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = lw::launch_loc(c, L);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(c->stream, &g);
            if (e != hipSuccess) return -(int)e;
            if (e2 != hipSuccess) return -(int)e2;
            c->loc_graph = g;
            LW_CK(hipGraphInstantiate(&c->loc_exec, g, nullptr, nullptr, 0));
            std::memcpy(c->loc_key, key, sizeof(key));
        }
        LW_CK(hipGraphLaunch(c->loc_exec, c->stream));
        launched = true;
    }
    if (!launched) LW_CK(lw::launch_loc(c, L));
    if (timed) {
        LW_CK(hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream));
        c->prof_used++;
    }
    return 0;
}

int lw_loc_tile_info(int32_t block, int32_t chunk, int32_t *info) {
    if (!info || (block != 64 && block != 128 && block != 256 && block != 512) || (chunk != 7 && chunk != 14))
        return LW_EINVAL;
    hipFuncAttributes fa{};
    LW_CK(lw::loc_tile_attributes(block, chunk, &fa));
    info[0] = (int32_t)fa.localSizeBytes;
    info[1] = (int32_t)fa.numRegs;
    info[2] = (int32_t)fa.sharedSizeBytes;
    info[3] = (int32_t)fa.maxThreadsPerBlock;
    return 0;
}

int lw_frame_features(lw_ctx *c, int32_t m, float r_max, float e_max, const float *motion, float *out) {
    if (!c || !motion || !out || m > c->frame_cap || !(r_max > 0.0f) || !(e_max > 0.0f)) return LW_EINVAL;
    if (!c->f_raw) return LW_ENOTBOUND;
    if (m < 0) LW_CK(frame_count_ready(c));  // the open frame's count, read on the device
    lw::FeatArgs f{};
    f.r_max = r_max;
    f.e_max = e_max;
    std::memcpy(f.motion, motion, sizeof(f.motion));
    LW_CK(lw::launch_features(c->stream, c->f_raw, m, m < 0 ? &c->ctr[6] : nullptr, (int32_t)c->frame_cap, f, out));
    return 0;
}

int lw_set_persistent(lw_ctx *c, int32_t on) {
    if (!c || on < 0 || on > 2) return LW_EINVAL;
    c->loop_ok = on != 0;
    c->loop_plain = on == 1;
    return 0;
}

int lw_register(lw_ctx *c, int32_t m, const lw_reg_params *p, const double *T_pred, lw_reg_result *res) {
    if (!c || bad_params(p) || !T_pred || !res || m > c->frame_cap) return LW_EINVAL;
    if (!c->f_raw) return LW_ENOTBOUND;
    LW_CK(frame_count_ready(c));
    // m < 0: the open frame's count on the device (thinning reads it there; the result carries it back)
    LW_CK(lw::launch_thin(c, m < 0 ? -1 : m, (float)(2.0 / (double)p->reg_voxel), p->min_range, p->max_range,
                          p->ret_mask));
    const int32_t bound = m < 0 ? (int32_t)c->frame_cap : m;
    return run_registration(c, &c->ctr[5], bound, -1, m < 0 ? -1 : m, p, T_pred, res);  // thinned count on the device
}

namespace {

// lw_odom_frame's chain, with the learned prior's steps when pr is given (lw_odom_frame_prior).
int odom_frame(lw_ctx *c, const lw_reg_params *p, const lw_frame_in *in, const lw_frame_prior *pr,
               lw_frame_out *out) {
    if (!c || bad_params(p) || !in || !out) return LW_EINVAL;
    if (pr && !pr->y) return LW_EINVAL;
    // the prior's chain prepares the frame in one launch: a larger frame is refused before anything is queued
    if (pr && !(c->frame_bound <= lw::frame_prep_max() && c->ds_cap >= 4096)) return LW_ERANGE;
    if (!c->f_raw) return LW_ENOTBOUND;
    if (!c->loop_ok) return LW_ENOLOOP;
    LW_CK(frame_count_ready(c));
    std::memset(out, 0, sizeof(*out));
    lw::RegLoop L;
    if (int rc = reg_setup(c, p, L)) return rc;
    // one upload: the prediction and the kappa schedule (as run_registration), then the frame's control inputs,
    // contiguous in the chain block
    lw::FrameCtl f{};
    std::memcpy(f.T_pred, in->T_pred, sizeof(f.T_pred));
    std::memcpy(f.T_prev_inv, in->T_prev_inv, sizeof(f.T_prev_inv));
    std::memcpy(f.T_key_inv, in->T_key_inv, sizeof(f.T_key_inv));
    f.kf_dist = in->kf_dist;
    f.kf_angle = in->kf_angle;
    f.redeskew_rot = in->redeskew_rot;
    f.redeskew_trans = in->redeskew_trans;
    std::memcpy(f.xi0, in->xi, sizeof(f.xi0));
    f.min_points = in->min_points;
    f.min_corr = p->min_corr;
    f.has_prev = in->has_prev;
    f.has_key = in->has_key;
    f.redeskew = in->redeskew;
    f.frame_cap = (int32_t)c->frame_cap;
    if (pr) {
        f.prior_on = 1;
        std::memcpy(f.prior_T, pr->T, sizeof(f.prior_T));
        std::memcpy(f.hist0, pr->hist0, sizeof(f.hist0));
        std::memcpy(f.sd_t, pr->sd_t, sizeof(f.sd_t));
        std::memcpy(f.mu_t, pr->mu_t, sizeof(f.mu_t));
        f.rms_gate = pr->rms_gate;
    }
    constexpr size_t UP_CTL = CH_FCTL - CH_TP;
    std::memcpy(pin_up(c), in->T_pred, 16 * sizeof(double));
    std::memcpy(pin_up(c) + 16 * sizeof(double), L.kappa, 64 * sizeof(float));
    std::memcpy(pin_up(c) + UP_CTL, &f, lw::FRAME_CTL_INPUT);
    LW_CK(hipMemcpyAsync(c->chain_tp, pin_up(c), UP_CTL + lw::FRAME_CTL_INPUT, hipMemcpyHostToDevice, c->stream));
    const float inv_v_thin = (float)(2.0 / (double)p->reg_voxel);
    const int32_t cap = (int32_t)c->frame_cap;
    // a frame one workgroup can prepare (up to frame_prep_max() points, with a thinning table of at least 4096
    // slots): deskew, thinning and the registration's start in one launch each time
    const bool one = c->frame_bound <= lw::frame_prep_max() && c->ds_cap >= 4096;
    LW_CK(lw::loop_epoch_guard(c, pr ? 3 : 2));  // the registrations' starts below
    if (pr) LW_CK(lw::launch_prior_guess(c, pr->y));  // the learned guess into the prediction and the twist
    // odom.py:671-681: deskew with the model's twist, register from the prediction
    if (one) {
        LW_CK(lw::launch_frame_prep(c, c->chain_tp, in->t_mid, in->inv_dt, in->xi, inv_v_thin, p->min_range,
                                    p->max_range, p->ret_mask, false, pr ? c->fctl->xi0 : nullptr));
    } else {
        LW_CK(lw::launch_deskew(c, -1, in->t_mid, in->inv_dt, in->xi));
        LW_CK(lw::launch_thin(c, -1, inv_v_thin, p->min_range, p->max_range, p->ret_mask));
        LW_CK(lw::launch_reg_begin(c, c->chain_tp));
    }
    hipError_t e = lw::launch_register_loop(c, &c->ctr[5], cap, L, p->acc_mode);
    if (e == hipErrorNotSupported) {
        c->loop_ok = false;  // no persistent launch on this device: the caller takes the step-by-step path
        (void)hipGetLastError();
        LW_CK(hipStreamSynchronize(c->stream));
        return LW_ENOLOOP;
    }
    if (e != hipSuccess) return -(int)e;
    if (pr) {  // LearnedPrior.accept: a first residual above the gate hands the frame to the fallback
        LW_CK(lw::launch_prior_fallback(c));
        LW_CK(lw::launch_frame_prep_fallback(c, in->t_mid, in->inv_dt, inv_v_thin, p->min_range, p->max_range,
                                             p->ret_mask));
        LW_CK(lw::launch_register_loop(c, &c->ctr[5], cap, L, p->acc_mode, &c->fctl->fb));
    }
    // odom.py:683-690: the re-deskew test, then (gated on it) deskew and register again from the first pose
    const int32_t *redo = &c->fctl->redo;
    if (one) {
        LW_CK(lw::launch_frame_prep(c, c->chain_tp, in->t_mid, in->inv_dt, nullptr, inv_v_thin, p->min_range,
                                    p->max_range, p->ret_mask, true));
    } else {
        LW_CK(lw::launch_frame_check(c));
        LW_CK(lw::launch_deskew(c, -1, in->t_mid, in->inv_dt, nullptr, c->fctl->xi1, redo));
        LW_CK(lw::launch_thin(c, -1, inv_v_thin, p->min_range, p->max_range, p->ret_mask, redo));
        LW_CK(lw::launch_reg_begin(c, c->chain_tp, redo, c->fctl->T1));
    }
    LW_CK(lw::launch_register_loop(c, &c->ctr[5], cap, L, p->acc_mode, redo));
    // odom.py:650-651, 712-727: the keyframe test, then (gated on it) the insert at the frame's pose, whose
    // touched voxels get their precision at once (the parameters are this frame's)
    LW_CK(lw::launch_keyframe_check(c));
    const lw::PrecParams q{c->prec_voxel, c->prec_noise2, c->prec_tight2, c->prec_alpha};
    const lw::InsertArgs ins{c->f_xyz, c->f_attr, cap, &c->ctr[6], lw::Mat4f{}, (float)(1.0 / (double)p->reg_voxel),
                             p->reg_voxel, p->max_range, p->ret_mask, &c->fctl->keyframe, c->fctl->T32, &q,
                             c->chain_counts};
    LW_CK(lw::launch_insert_head(c, ins));
    // one readback, once the insert's claims have fixed the map's counts: both registrations' states, the
    // decisions and the counts; the host waits for it alone, so the rest of the insert runs on while the host
    // goes on (anything it queues next runs after it)
    char *h = pin_down(c);
    LW_CK(hipMemcpyAsync(h, c->chain, CH_BYTES, hipMemcpyDeviceToHost, c->stream));
    LW_CK(hipEventRecord(c->chain_ev, c->stream));
    LW_CK(lw::launch_insert_tail(c, ins));
    LW_CK(hipEventSynchronize(c->chain_ev));
    LW_CK(prof_harvest(c));
    lw::RegState st2, st1, stf;
    lw::FrameCtl fc;
    std::memcpy(&st2, h, sizeof(st2));
    std::memcpy(&st1, h + CH_REG1, sizeof(st1));
    std::memcpy(&stf, h + CH_REGFB, sizeof(stf));
    std::memcpy(&fc, h + CH_FCTL, sizeof(fc));
    if (one && fc.err == 2) {  // a grid barrier's wait ran out: its arrival count is reset for the next frame
        LW_CK(lw::launch_fill_i32(c, reinterpret_cast<int32_t *>(c->prep_sync), sizeof(lw::PrepSync) / 4, 0));
        LW_CK(hipStreamSynchronize(c->stream));
        return LW_ELOOP;
    }
    if (one && fc.err) return LW_EINVAL;
    const bool fb = pr && fc.fb;
    if (st1.status < 0 || (fb && stf.status < 0) || (fc.redo && st2.status < 0)) return LW_ELOOP;
    out->m = fc.m;
    out->skipped = fc.skipped;
    out->ok = fc.ok;
    out->keyframe = fc.keyframe;
    out->registrations = fc.skipped ? 0 : 1 + (fb ? 1 : 0) + fc.redo;
    out->guess = pr ? fc.guess : 0;
    out->fallback = fb ? 1 : 0;
    if (pr && fc.guess) std::memcpy(out->d, fc.d, sizeof(out->d));
    std::memcpy(out->T, fc.T, sizeof(out->T));
    std::memcpy(out->xi, fc.xi1, sizeof(out->xi));
    std::memcpy(&out->map_voxels, h + CH_COUNTS, sizeof(int32_t));
    std::memcpy(&out->map_overflow, h + CH_COUNTS + sizeof(int32_t), sizeof(int32_t));
    // the registrations in order: the guess's (kept aside when the fallback ran), the fallback's, the second
    const lw::RegState *order[3] = {fb ? &stf : &st1, fb ? &st1 : &st2, &st2};
    for (int k = 0; k < out->registrations && k < 3; ++k) {
        out->reg[k].m = fc.m;
        fill_result(*order[k], p, &out->reg[k]);
    }
    return 0;
}

}  // namespace

int lw_odom_frame(lw_ctx *c, const lw_reg_params *p, const lw_frame_in *in, lw_frame_out *out) {
    return odom_frame(c, p, in, nullptr, out);
}

int lw_odom_frame_prior(lw_ctx *c, const lw_reg_params *p, const lw_frame_in *in, const lw_frame_prior *prior,
                        lw_frame_out *out) {
    if (!prior) return LW_EINVAL;
    return odom_frame(c, p, in, prior, out);
}

int lw_prof_enable(lw_ctx *c, int32_t on) {
    if (!c) return LW_EINVAL;
    if (on && !c->prof_ev) {
        const int cap = 512;
        c->prof_ev = new (std::nothrow) hipEvent_t[2 * cap];
        if (!c->prof_ev) return LW_EINVAL;
        for (int i = 0; i < 2 * cap; ++i) {
            hipError_t e = hipEventCreate(&c->prof_ev[i]);
            if (e != hipSuccess) {
                for (int j = 0; j < i; ++j) (void)hipEventDestroy(c->prof_ev[j]);
                delete[] c->prof_ev;
                c->prof_ev = nullptr;
                return -(int)e;
            }
        }
        c->prof_cap = cap;
        LW_CK(hipMalloc((void **)&c->prof_dev, lw::LW_PROF_COUNTERS * sizeof(unsigned long long)));
        LW_CK(hipMemsetAsync(c->prof_dev, 0, lw::LW_PROF_COUNTERS * sizeof(unsigned long long), c->stream));
        LW_CK(hipStreamSynchronize(c->stream));
    }
    c->prof_on = on != 0;
    c->prof_phases = on >= 2;
    return 0;
}

int lw_prof_reset(lw_ctx *c) {
    if (!c) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(c->stream));
    c->prof_used = 0;
    c->prof_ms = 0.0;
    c->prof_launches = 0;
    if (c->prof_dev) {
        LW_CK(hipMemsetAsync(c->prof_dev, 0, lw::LW_PROF_COUNTERS * sizeof(unsigned long long), c->stream));
        LW_CK(hipStreamSynchronize(c->stream));
    }
    return 0;
}

int lw_prof_read(lw_ctx *c, double *out) {
    if (!c || !out) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(c->stream));
    LW_CK(prof_harvest(c));
    unsigned long long v[lw::LW_PROF_COUNTERS] = {0};
    if (c->prof_dev) LW_CK(hipMemcpy(v, c->prof_dev, sizeof(v), hipMemcpyDeviceToHost));
    out[0] = (double)c->prof_launches;
    out[1] = c->prof_ms;
    for (int i = 0; i < 5; ++i) out[2 + i] = (double)v[i];
    out[7] = 0.0;
    return 0;
}

int lw_prof_phases(lw_ctx *c, double *out) {
    if (!c || !out) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(c->stream));
    unsigned long long v[lw::LW_PROF_COUNTERS] = {0};
    if (c->prof_dev) LW_CK(hipMemcpy(v, c->prof_dev, sizeof(v), hipMemcpyDeviceToHost));
    int khz = 0;
    LW_CK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    const double us_per_tick = khz > 0 ? 1000.0 / (double)khz : 0.0;
    for (int i = 0; i < 3; ++i) out[i] = (double)v[8 + i] * us_per_tick;
    out[3] = (double)v[11];
    for (int i = 0; i < 3; ++i) out[4 + i] = (double)v[12 + i] * us_per_tick;
    out[7] = (double)v[15];
    out[8] = (double)v[5];
    out[9] = (double)v[6];
    for (int i = 0; i < 4; ++i) out[10 + i] = (double)v[16 + i] * us_per_tick;
    out[14] = (double)v[20];
    for (int i = 0; i < 3; ++i) out[15 + i] = (double)v[21 + i] * us_per_tick;
    for (int i = 0; i < 2; ++i) out[18 + i] = (double)v[24 + i] * us_per_tick;
    out[20] = (double)v[26];
    out[21] = v[29] ? (double)v[28] * us_per_tick / (double)v[29] : 0.0;
    out[22] = (double)v[30] * us_per_tick;
    out[23] = (double)v[31];
    return 0;
}

// ---- host builds of the device math (tests) ---------------------------------------------------------

int lw_host_gn_step(const double *v, const double *T, const double *T_pred, double damping, double prior_info,
                    double degen_thr, int32_t min_corr, double *delta, double *rms, int32_t *n_degen) {
    if (!v || !T || !T_pred || !delta || !rms || !n_degen) return LW_EINVAL;
    lw::gn::StepParams P{damping, prior_info, degen_thr, min_corr, 0};
    lw::gn::StepOut o;
    int rc = lw::gn::step(v, T, T_pred, P, o);
    std::memcpy(delta, o.delta, sizeof(o.delta));
    *rms = o.rms;
    *n_degen = o.n_degen;
    return rc;
}

int lw_host_voxel_precision(const float *rsum, const float *router6, int32_t c, float voxel, float noise2,
                            float tight2, float alpha, float *M6, float *Mg6, float *q3) {
    if (!rsum || !router6 || !M6 || !Mg6 || !q3 || c <= 0) return LW_EINVAL;
    return lw::voxel_precision(rsum, router6, c, voxel, noise2, tight2, alpha, M6, Mg6, q3);
}

int lw_host_apply_twist(double *T, const double *delta) {
    if (!T || !delta) return LW_EINVAL;
    lw::gn::apply_twist(T, delta);
    return 0;
}

int lw_host_twist_between(const double *T_from, const double *T_to, double *xi) {
    if (!T_from || !T_to || !xi) return LW_EINVAL;
    lw::gn::twist_between(T_from, T_to, xi);
    return 0;
}

double lw_host_cond(const double *H) { return H ? lw::gn::cond_of(H) : -1.0; }

namespace {

// One wave: the wave form of the step and its pose update, then the scalar form and its update in lane 0, on
// the same accumulator (lw_gn_step_check's out layout).
__global__ __launch_bounds__(64) void k_gn_step_check(const double *v, const double *T, const double *Tp,
                                                      lw::gn::StepParams P, double *out) {
    __shared__ double sh[lw::gnw::W_DOUBLES], sT[16], xi[6], H[36];
    const int lane = threadIdx.x;
    if (lane < 16) sT[lane] = T[lane];
    if (lane == 0) lw::gn::twist_between(T, Tp, xi);
    __syncthreads();
    lw::gnw::WaveOut o;
    unsigned long long t1 = 0ull, t2 = 0ull, t3 = 0ull;
    const int r = lw::gnw::wave_step(v, sT, xi, P, sh, o, H, false, t1, t2, t3);
    if (r > 0 && lane == 0) lw::gn::apply_twist(sT, o.delta);  // as the loop's accept_step moves the pose
    __syncthreads();
    if (lane < 16) out[lane] = sT[lane];
    if (lane == 0) {
        out[32] = (double)r;
        out[33] = r > 0 ? o.rms : 0.0;
        for (int k = 0; k < 6; ++k) out[34 + k] = r > 0 ? o.delta[k] : 0.0;
        lw::gn::StepOut so;
        double Ts[16];
        for (int k = 0; k < 16; ++k) Ts[k] = T[k];
        const int rc = lw::gn::step(v, T, Tp, P, so);
        if (rc == lw::gn::STEP_OK) lw::gn::apply_twist(Ts, so.delta);
        for (int k = 0; k < 16; ++k) out[16 + k] = Ts[k];
        out[40] = (double)rc;
        out[41] = so.rms;
        for (int k = 0; k < 6; ++k) out[42 + k] = so.delta[k];
    }
}

}  // namespace

int lw_gn_step_check(const double *v, const double *T, const double *T_pred, double damping, double prior_info,
                     double degen_thr, int32_t min_corr, double *out) {
    if (!v || !T || !T_pred || !out) return LW_EINVAL;
    double *d = nullptr;
    LW_CK(hipMalloc((void **)&d, (64 + 16 + 16 + 48) * sizeof(double)));
    hipError_t e = hipMemcpy(d, v, 64 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 64, T, 16 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d + 80, T_pred, 16 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        lw::gn::StepParams P{damping, prior_info, degen_thr, min_corr, 0};
        hipLaunchKernelGGL(k_gn_step_check, dim3(1), dim3(64), 0, nullptr, d, d + 64, d + 80, P, d + 96);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + 96, 48 * sizeof(double), hipMemcpyDeviceToHost);
    const hipError_t ef = hipFree(d);
    LW_CK(e);
    LW_CK(ef);
    return 0;
}

}  // extern "C"

// ---- the batched odometry: one frame of each of S sequences per call ----------------------------------------

namespace {
constexpr size_t al256(size_t n) { return (n + 255) & ~(size_t)255; }
constexpr size_t TP_BYTES = 16 * sizeof(double) + 64 * sizeof(float);  // a prediction and its kappa schedule
constexpr int BATCH_EVENTS = 1024;                                      // timed registration launches in flight
}  // namespace

struct lw_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int32_t S = 0;
    std::vector<lw_ctx *> ctx;
    int64_t frame_cap = 0;  // the smallest of the contexts'
    // one device block: [prep1 | prep2 | ins | tp | fc] uploaded with every frame, [fc | st1 | st | counts] read
    // back, then the descriptors that never change [reg1 | reg2 | key]
    char *dev = nullptr;
    size_t o_prep2 = 0, o_ins = 0, o_tp = 0, o_fc = 0, o_st1 = 0, o_st = 0, o_counts = 0, o_reg1 = 0, o_reg2 = 0,
           o_key = 0, o_end = 0;
    char *pin = nullptr;  // pinned mirror of [prep1 .. counts]
    unsigned long long *prof_dev = nullptr;
    bool prof_on = false;
    hipEvent_t *ev = nullptr;
    int ev_used = 0;
    double prof_ms = 0.0;
    int64_t prof_launches = 0;
};

namespace {

template <class T>
T *at(lw_batch *b, size_t off) {
    return reinterpret_cast<T *>(b->dev + off);
}
template <class T>
T *pin_at(lw_batch *b, size_t off) {
    return reinterpret_cast<T *>(b->pin + off);
}

void batch_free(lw_batch *b) {
    if (b->dev) (void)hipFree(b->dev);
    if (b->pin) (void)hipHostFree(b->pin);
    if (b->prof_dev) (void)hipFree(b->prof_dev);
    if (b->ev) {
        for (int i = 0; i < 2 * BATCH_EVENTS; ++i) (void)hipEventDestroy(b->ev[i]);
        delete[] b->ev;
    }
}

hipError_t batch_harvest(lw_batch *b) {
    for (int i = 0; i < b->ev_used; ++i) {
        float ms = 0.0f;
        hipError_t e = hipEventElapsedTime(&ms, b->ev[2 * i], b->ev[2 * i + 1]);
        if (e != hipSuccess) return e;
        b->prof_ms += ms;
        b->prof_launches++;
    }
    b->ev_used = 0;
    return hipSuccess;
}

hipError_t batch_register(lw_batch *b, size_t o_reg, const lw::RegLoop &L, int acc_mode) {
    const bool timed = b->prof_on && b->ev_used < BATCH_EVENTS;
    hipError_t e;
    if (timed && (e = hipEventRecord(b->ev[2 * b->ev_used], b->stream)) != hipSuccess) return e;
    if ((e = lw::launch_register_seq(b->stream, at<lw::SeqReg>(b, o_reg), b->S, L, acc_mode,
                                     b->prof_on ? b->prof_dev : nullptr)) != hipSuccess)
        return e;
    if (timed) {
        if ((e = hipEventRecord(b->ev[2 * b->ev_used + 1], b->stream)) != hipSuccess) return e;
        b->ev_used++;
    }
    return hipSuccess;
}

}  // namespace

extern "C" {

int lw_batch_create(lw_batch **out, lw_ctx *const *ctxs, int32_t S, void *stream) {
    if (!out || !ctxs || S <= 0) return LW_EINVAL;
    *out = nullptr;
    auto *b = new (std::nothrow) lw_batch();
    if (!b) return LW_EINVAL;
    b->stream = (hipStream_t)stream;
    b->S = S;
    b->ctx.assign(ctxs, ctxs + S);
    b->frame_cap = INT64_MAX;
    hipError_t e = hipGetDevice(&b->device);
    for (int32_t s = 0; s < S && e == hipSuccess; ++s) {
        lw_ctx *c = ctxs[s];
        if (!c || c->stream != b->stream || c->device != b->device || !c->f_xyz || c->ds_cap < 4096) {
            delete b;
            return c && !c->f_xyz ? LW_ENOTBOUND : LW_EINVAL;
        }
        b->frame_cap = c->frame_cap < b->frame_cap ? c->frame_cap : b->frame_cap;
    }
    if (e != hipSuccess) {
        delete b;
        return -(int)e;
    }
    const size_t n = (size_t)S;
    b->o_prep2 = al256(n * sizeof(lw::PrepArgs));
    b->o_ins = b->o_prep2 + al256(n * sizeof(lw::PrepArgs));
    b->o_tp = b->o_ins + al256(n * sizeof(lw::InsSeq));
    b->o_fc = b->o_tp + al256(n * TP_BYTES);
    b->o_st1 = b->o_fc + al256(n * sizeof(lw::FrameCtl));
    b->o_st = b->o_st1 + al256(n * sizeof(lw::RegState));
    b->o_counts = b->o_st + al256(n * sizeof(lw::RegState));
    b->o_reg1 = b->o_counts + al256(n * 2 * sizeof(int32_t));
    b->o_reg2 = b->o_reg1 + al256(n * sizeof(lw::SeqReg));
    b->o_key = b->o_reg2 + al256(n * sizeof(lw::SeqReg));
    b->o_end = b->o_key + al256(n * sizeof(lw::KeySeq));
    auto fail = [&](hipError_t err) {
        batch_free(b);
        delete b;
        return -(int)err;
    };
    if ((e = hipMalloc((void **)&b->dev, b->o_end)) != hipSuccess) return fail(e);
    if ((e = hipHostMalloc((void **)&b->pin, b->o_reg1, hipHostMallocDefault)) != hipSuccess) return fail(e);
    if ((e = hipMemsetAsync(b->dev, 0, b->o_end, b->stream)) != hipSuccess) return fail(e);
    // the descriptors that never change: each sequence's two registrations and its keyframe decision
    std::vector<char> fixed(b->o_end - b->o_reg1);
    auto *reg1 = reinterpret_cast<lw::SeqReg *>(fixed.data());
    auto *reg2 = reinterpret_cast<lw::SeqReg *>(fixed.data() + (b->o_reg2 - b->o_reg1));
    auto *key = reinterpret_cast<lw::KeySeq *>(fixed.data() + (b->o_key - b->o_reg1));
    for (int32_t s = 0; s < S; ++s) {
        lw_ctx *c = ctxs[s];
        lw::FrameCtl *fc = at<lw::FrameCtl>(b, b->o_fc) + s;
        lw::RegState *st = at<lw::RegState>(b, b->o_st) + s;
        reg1[s] = lw::SeqReg{c->ds_xyz, &c->ctr[5], c->slots, c->occ, c->vox_prec, st, &fc->run1, c->mask, 0};
        reg2[s] = lw::SeqReg{c->ds_xyz, &c->ctr[5], c->slots, c->occ, c->vox_prec, st, &fc->redo, c->mask, 0};
        key[s] = lw::KeySeq{st, fc, c->ctr, at<int32_t>(b, b->o_counts) + 2 * s};
    }
    if ((e = hipMemcpyAsync(b->dev + b->o_reg1, fixed.data(), fixed.size(), hipMemcpyHostToDevice, b->stream)) !=
        hipSuccess)
        return fail(e);
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return fail(e);
    *out = b;
    return 0;
}

int lw_batch_destroy(lw_batch *b) {
    if (!b) return LW_EINVAL;
    (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
    return 0;
}

int lw_batch_frame(lw_batch *b, const lw_reg_params *p, const lw_frame_in *in, const lw_batch_src *src,
                   lw_frame_out *out) {
    if (!b || bad_params(p) || !in || !src || !out) return LW_EINVAL;
    const int32_t S = b->S;
    for (int32_t s = 0; s < S; ++s)
        if (src[s].m < 0 || src[s].m > b->frame_cap || src[s].m > lw::frame_prep_max() ||
            (src[s].m > 0 && (!src[s].xyz || !src[s].attr || !src[s].t)))
            return LW_EINVAL;
    lw::RegLoop L;
    for (int32_t s = 0; s < S; ++s)  // the same parameters for every map; refreshes only when they change
        if (int rc = reg_setup(b->ctx[s], p, L)) return rc;
    const float inv_v_thin = (float)(2.0 / (double)p->reg_voxel);
    const lw::PrecParams q{b->ctx[0]->prec_voxel, b->ctx[0]->prec_noise2, b->ctx[0]->prec_tight2, b->ctx[0]->prec_alpha};
    auto *prep1 = pin_at<lw::PrepArgs>(b, 0);
    auto *prep2 = pin_at<lw::PrepArgs>(b, b->o_prep2);
    auto *ins = pin_at<lw::InsSeq>(b, b->o_ins);
    auto *fcs = pin_at<lw::FrameCtl>(b, b->o_fc);
    for (int32_t s = 0; s < S; ++s) {
        lw_ctx *c = b->ctx[s];
        const lw_frame_in &fi = in[s];
        const bool warm = src[s].warm != 0;
        lw::FrameCtl *fc_dev = at<lw::FrameCtl>(b, b->o_fc) + s;
        const double *tp_dev = reinterpret_cast<const double *>(b->dev + b->o_tp + (size_t)s * TP_BYTES);
        char *tp = b->pin + b->o_tp + (size_t)s * TP_BYTES;
        std::memcpy(tp, fi.T_pred, 16 * sizeof(double));
        std::memcpy(tp + 16 * sizeof(double), L.kappa, 64 * sizeof(float));
        lw::FrameCtl f{};
        std::memcpy(f.T_pred, fi.T_pred, sizeof(f.T_pred));
        std::memcpy(f.T_prev_inv, fi.T_prev_inv, sizeof(f.T_prev_inv));
        std::memcpy(f.T_key_inv, fi.T_key_inv, sizeof(f.T_key_inv));
        f.kf_dist = fi.kf_dist;
        f.kf_angle = fi.kf_angle;
        f.redeskew_rot = fi.redeskew_rot;
        f.redeskew_trans = fi.redeskew_trans;
        std::memcpy(f.xi0, fi.xi, sizeof(f.xi0));
        f.min_points = fi.min_points;
        f.min_corr = p->min_corr;
        f.has_prev = warm ? 0 : fi.has_prev;
        f.has_key = fi.has_key;
        f.redeskew = warm ? 0 : fi.redeskew;
        f.frame_cap = (int32_t)c->frame_cap;
        f.warm = warm ? 1 : 0;
        f.run1 = warm ? 0 : 1;
        f.m_in = src[s].m;
        fcs[s] = f;
        lw::PrepArgs A{};
        A.raw = src[s].xyz;
        A.t = src[s].t;
        A.attr = src[s].attr;
        A.xyz = c->f_xyz;
        A.m_dev = &fc_dev->m_in;
        A.frame_cap = (int32_t)c->frame_cap;
        A.t_mid = fi.t_mid;
        A.inv_dt = fi.inv_dt;
        for (int k = 0; k < 6; ++k) A.xi[k] = fi.xi[k];
        A.inv_v = inv_v_thin;
        A.min_range = p->min_range;
        A.max_range = p->max_range;
        A.ret_mask = p->ret_mask;
        A.ds_mask = (int32_t)(c->ds_cap - 1);
        A.ds_keys = c->ds_keys;
        A.ds_idx = c->ds_idx;
        A.ds_xyz = c->ds_xyz;
        A.n_out = &c->ctr[5];
        A.st = at<lw::RegState>(b, b->o_st) + s;
        A.Tp = tp_dev;
        A.epoch = nullptr;  // the batched registration has no persistent loop
        A.st1 = nullptr;
        A.fc = fc_dev;
        prep1[s] = A;
        A.st1 = at<lw::RegState>(b, b->o_st1) + s;  // the check: keep the first registration, then maybe redo
        prep2[s] = A;
        ins[s] = lw::InsSeq{lw::map_refs(c), c->f_xyz, src[s].attr, (int32_t)c->frame_cap, &fc_dev->m_in, fc_dev->T32,
                            &fc_dev->keyframe, at<int32_t>(b, b->o_counts) + 2 * s};
    }
    LW_CK(hipMemcpyAsync(b->dev, b->pin, b->o_st1, hipMemcpyHostToDevice, b->stream));
    // deskew and thin every frame, register, the re-deskew test and (gated) the second preparation and
    // registration, the keyframe decisions, then every keyframe's insert with its voxels refreshed
    LW_CK(lw::launch_frame_prep_seq(b->stream, at<lw::PrepArgs>(b, 0), S));
    LW_CK(batch_register(b, b->o_reg1, L, p->acc_mode));
    LW_CK(lw::launch_frame_prep_seq(b->stream, at<lw::PrepArgs>(b, b->o_prep2), S));
    LW_CK(batch_register(b, b->o_reg2, L, p->acc_mode));
    LW_CK(lw::launch_keyframe_check_seq(b->stream, at<lw::KeySeq>(b, b->o_key), S));
    LW_CK(lw::launch_insert_seq(b->stream, at<lw::InsSeq>(b, b->o_ins), S, (int32_t)b->frame_cap,
                                (float)(1.0 / (double)p->reg_voxel), p->reg_voxel, p->max_range, p->ret_mask, q));
    LW_CK(hipMemcpyAsync(b->pin + b->o_fc, b->dev + b->o_fc, b->o_reg1 - b->o_fc, hipMemcpyDeviceToHost, b->stream));
    LW_CK(hipStreamSynchronize(b->stream));
    LW_CK(batch_harvest(b));
    const auto *st1s = pin_at<lw::RegState>(b, b->o_st1);
    const auto *sts = pin_at<lw::RegState>(b, b->o_st);
    const auto *counts = pin_at<int32_t>(b, b->o_counts);
    int rc = 0;
    for (int32_t s = 0; s < S; ++s) {
        const lw::FrameCtl &fc = fcs[s];
        lw_frame_out &o = out[s];
        std::memset(&o, 0, sizeof(o));
        if (fc.err) rc = LW_EINVAL;
        o.m = fc.m;
        o.skipped = fc.skipped;
        o.ok = fc.ok;
        o.keyframe = fc.keyframe;
        o.registrations = (fc.skipped || fc.warm) ? 0 : 1 + fc.redo;
        std::memcpy(o.T, fc.T, sizeof(o.T));
        std::memcpy(o.xi, fc.xi1, sizeof(o.xi));
        o.map_voxels = counts[2 * s];
        o.map_overflow = counts[2 * s + 1];
        if (o.registrations >= 1) {  // the condition numbers come from the registration kernel
            o.reg[0].m = fc.m;
            fill_result(st1s[s], p, &o.reg[0], true);
        }
        if (o.registrations >= 2) {
            o.reg[1].m = fc.m;
            fill_result(sts[s], p, &o.reg[1], true);
        }
    }
    return rc;
}

int lw_batch_prof_enable(lw_batch *b, int32_t on) {
    if (!b) return LW_EINVAL;
    if (on && !b->ev) {
        b->ev = new (std::nothrow) hipEvent_t[2 * BATCH_EVENTS];
        if (!b->ev) return LW_EINVAL;
        for (int i = 0; i < 2 * BATCH_EVENTS; ++i) LW_CK(hipEventCreate(&b->ev[i]));
        LW_CK(hipMalloc((void **)&b->prof_dev, lw::LW_PROF_COUNTERS * sizeof(unsigned long long)));
        LW_CK(hipMemsetAsync(b->prof_dev, 0, lw::LW_PROF_COUNTERS * sizeof(unsigned long long), b->stream));
        LW_CK(hipStreamSynchronize(b->stream));
    }
    b->prof_on = on != 0;
    return 0;
}

int lw_batch_prof_reset(lw_batch *b) {
    if (!b) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(b->stream));
    b->ev_used = 0;
    b->prof_ms = 0.0;
    b->prof_launches = 0;
    if (b->prof_dev) {
        LW_CK(hipMemsetAsync(b->prof_dev, 0, lw::LW_PROF_COUNTERS * sizeof(unsigned long long), b->stream));
        LW_CK(hipStreamSynchronize(b->stream));
    }
    return 0;
}

int lw_batch_prof_read(lw_batch *b, double *out) {
    if (!b || !out) return LW_EINVAL;
    LW_CK(hipStreamSynchronize(b->stream));
    LW_CK(batch_harvest(b));
    unsigned long long v[lw::LW_PROF_COUNTERS] = {0};
    if (b->prof_dev) LW_CK(hipMemcpy(v, b->prof_dev, sizeof(v), hipMemcpyDeviceToHost));
    out[0] = (double)b->prof_launches;
    out[1] = b->prof_ms;
    for (int i = 0; i < 5; ++i) out[2 + i] = (double)v[i];
    out[7] = 0.0;
    return 0;
}

}  // extern "C"

