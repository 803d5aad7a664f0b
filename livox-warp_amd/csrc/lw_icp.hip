// Scan-to-map accumulation and the Gauss-Newton loop on the device: k_icp + k_reduce
// (python/livox_warp/odom.py:275-384) and the host step of Odometry._register (odom.py:752-803).
//
// One pass = one launch of ceil(n / 64) one-wave workgroups, one thinned point per lane. A lane poses its
// point with the float32 pose and searches the (2 reach + 1)^3 voxels around it for the closest voxel mean,
// in the reference's order (dx outer, dz inner, strict <; odom.py:315-332). The search runs in chunks: the
// bitmap words of a whole chunk are loaded before any is tested, then the chunk's slots (32 B records that
// carry key, voxel id, count and mean), so a chunk costs two dependent memory round trips. The lane then
// forms the 53 float32 terms of the ACC layout exactly as the reference rounds them (odom.py:246-259,
// 269-271, 372-375); the wave sums them with cross-lane shuffles and writes one partial row. The last
// workgroup to finish (atomic ticket, no waiting) sums the rows in a fixed order into the reduced
// accumulator. In a registration a one-wave kernel then runs the float64 step (lw_gn.h) and updates the
// pose in device memory; passes after convergence return at once, so a registration is queued in full
// with one readback at the end.
//
// Accumulation forms (acc_mode): 0 float64 (the reference), 1 compensated float32 (Neumaier sums, two-sum
// merges), 2 naive float32.
#include "lw_internal.h"

namespace lw {

namespace {

template <int MODE>
struct Acc;

template <>
struct Acc<0> {
    double s[ACC_USED];
    __device__ void zero() {
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) s[k] = 0.0;
    }
    __device__ void add(int k, float x) { s[k] += (double)x; }
    __device__ void wave_reduce() {
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) {
            double v = s[k];
            LW_UNROLL for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            s[k] = v;
        }
    }
    __device__ void store(double *row, int k) const { row[k] = s[k]; row[64 + k] = 0.0; }
};

template <>
struct Acc<2> {
    float s[ACC_USED];
    __device__ void zero() {
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) s[k] = 0.0f;
    }
    __device__ void add(int k, float x) { s[k] += x; }
    __device__ void wave_reduce() {
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) {
            float v = s[k];
            LW_UNROLL for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            s[k] = v;
        }
    }
    __device__ void store(double *row, int k) const { row[k] = (double)s[k]; row[64 + k] = 0.0; }
};

__device__ inline void two_sum(float a, float b, float &s, float &e) {
    s = a + b;
    float bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

template <>
struct Acc<1> {
    float s[ACC_USED], c[ACC_USED];
    __device__ void zero() {
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) s[k] = c[k] = 0.0f;
    }
    __device__ void add(int k, float x) {
        float t, e;
        two_sum(s[k], x, t, e);
        s[k] = t;
        c[k] += e;
    }
    __device__ void wave_reduce() {
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k) {
            float v = s[k], w = c[k];
            LW_UNROLL for (int off = 32; off > 0; off >>= 1) {
                float ov = __shfl_xor(v, off, 64), ow = __shfl_xor(w, off, 64);
                float t, e;
                two_sum(v, ov, t, e);
                v = t;
                w = w + ow + e;
            }
            s[k] = v;
            c[k] = w;
        }
    }
    __device__ void store(double *row, int k) const { row[k] = (double)s[k]; row[64 + k] = (double)c[k]; }
};

struct IcpArgs {
    const float *ds;
    const Slot *slots;
    const uint32_t *occ;
    const VoxPrec *vox_prec;
    RegState *st;
    double *partials;  // [G][128]: sums, compensations
    int32_t *best_slot;
    unsigned long long *prof;  // measurement counters (see lw_ctx), or null
    int32_t n;
    int32_t mask;
    float inv_v, voxel, max_d2, kappa;
};

#define LW_H_PAIRS(X) \
    X(0, 0, 0) X(0, 1, 1) X(0, 2, 2) X(0, 3, 3) X(0, 4, 4) X(0, 5, 5) X(1, 1, 6) X(1, 2, 7) X(1, 3, 8) X(1, 4, 9) \
    X(1, 5, 10) X(2, 2, 11) X(2, 3, 12) X(2, 4, 13) X(2, 5, 14) X(3, 3, 15) X(3, 4, 16) X(3, 5, 17) X(4, 4, 18)   \
    X(4, 5, 19) X(5, 5, 20)

// A slot as two 16-byte loads: lo = {key lo, key hi, voxel id, count bits}, hi = {mean x, y, z, pad}.
__device__ inline void load_slot(const Slot *slots, int32_t s, int4 &lo, float4 &hi) {
    const int4 *p = reinterpret_cast<const int4 *>(slots + s);
    lo = p[0];
    hi = reinterpret_cast<const float4 *>(p)[1];
}

__device__ inline int64_t key_of(const int4 &lo) {
    return (int64_t)(((uint64_t)(uint32_t)lo.y << 32) | (uint64_t)(uint32_t)lo.x);
}

// The rest of the linear probe after a first slot that holds another key (rare at the loads the map runs
// at): gpu.py:254-265 from probe 1 on. Returns the slot or -1.
__device__ __noinline__ int32_t probe_on(const Slot *slots, const uint32_t *occ, int32_t mask, int32_t h,
                                         int64_t key) {
    for (int probe = 1; probe < PROBES; ++probe) {
        int32_t t = (h + probe) & mask;
        if (!((occ[t >> 5] >> (t & 31)) & 1u)) return -1;
        if (slots[t].key == key) return t;
    }
    return -1;
}

// Closest voxel mean within reach R (odom.py:308-334): chunks of C neighbours in the reference's order.
template <int R, int C>
__device__ inline int32_t closest_mean(const IcpArgs &a, float wx, float wy, float wz, float &mx, float &my,
                                       float &mz, int32_t &hits) {
    constexpr int D = 2 * R + 1, K = D * D * D;
    const int32_t bx = ifloor(wx * a.inv_v), by = ifloor(wy * a.inv_v), bz = ifloor(wz * a.inv_v);
    float best_d2 = a.max_d2;
    int32_t best_s = -1;
    for (int j0 = 0; j0 < K; j0 += C) {
        int64_t key[C];
        int32_t h[C];
        uint32_t w[C];
        LW_UNROLL for (int c = 0; c < C; ++c) {
            const int j = j0 + c;
            const int dx = j / (D * D) - R, dy = (j / D) % D - R, dz = j % D - R;
            key[c] = voxel_key_ijk(bx + dx, by + dy, bz + dz);
            h[c] = slot_hash(key[c]) & a.mask;
        }
        LW_UNROLL for (int c = 0; c < C; ++c) w[c] = (j0 + c < K) ? a.occ[h[c] >> 5] : 0u;
        int4 lo[C];
        float4 hi[C];
        LW_UNROLL for (int c = 0; c < C; ++c) {
            lo[c] = make_int4(-1, -1, -1, 0);
            hi[c] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if ((w[c] >> (h[c] & 31)) & 1u) load_slot(a.slots, h[c], lo[c], hi[c]);
        }
        LW_UNROLL for (int c = 0; c < C; ++c) {
            if (!((w[c] >> (h[c] & 31)) & 1u)) continue;
            int32_t s = h[c];
            if (key_of(lo[c]) != key[c]) {
                s = probe_on(a.slots, a.occ, a.mask, h[c], key[c]);
                if (s < 0) continue;
                load_slot(a.slots, s, lo[c], hi[c]);
            }
            ++hits;
            if (!(__int_as_float(lo[c].w) > 0.0f)) continue;  // count 0 (odom.py:324)
            float d0 = wx - hi[c].x, d1 = wy - hi[c].y, d2 = wz - hi[c].z;
            float dd = d0 * d0;
            dd = dd + d1 * d1;
            dd = dd + d2 * d2;
            if (dd < best_d2) {
                best_d2 = dd;
                best_s = s;
                mx = hi[c].x;
                my = hi[c].y;
                mz = hi[c].z;
            }
        }
    }
    return best_s;
}

template <int MODE, int R, int C>
__device__ inline void point_terms(const IcpArgs &a, const float *Tf, int32_t i, Acc<MODE> &acc, int32_t &hits,
                                   int32_t &matched) {
    const float px = a.ds[3 * (int64_t)i], py = a.ds[3 * (int64_t)i + 1], pz = a.ds[3 * (int64_t)i + 2];
    float wx, wy, wz;
    transform_point(Tf, px, py, pz, wx, wy, wz);
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    const int32_t s = closest_mean<R, C>(a, wx, wy, wz, mx, my, mz, hits);
    if (a.best_slot) a.best_slot[i] = s;
    if (s < 0) return;
    ++matched;
    const int32_t vid = a.slots[s].vid;
    const float r0 = wx - mx, r1 = wy - my, r2 = wz - mz;
    const VoxPrec P = a.vox_prec[vid];
    const float M[3][3] = {{P.M[0], P.M[1], P.M[2]}, {P.M[1], P.M[3], P.M[4]}, {P.M[2], P.M[4], P.M[5]}};
    const float G[3][3] = {{P.Mg[0], P.Mg[1], P.Mg[2]}, {P.Mg[1], P.Mg[3], P.Mg[4]}, {P.Mg[2], P.Mg[4], P.Mg[5]}};
    float rn;
    if (P.shape > 0.5f) {
        rn = r0 * P.q[0];
        rn = rn + r1 * P.q[1];
        rn = rn + r2 * P.q[2];
    } else {
        rn = length3(r0, r1, r2);
    }
    float Mr[3];
    LW_UNROLL for (int k = 0; k < 3; ++k) {
        float t = M[k][0] * r0;
        t = t + M[k][1] * r1;
        Mr[k] = t + M[k][2] * r2;
    }
    float d2 = r0 * Mr[0];
    d2 = d2 + r1 * Mr[1];
    d2 = d2 + r2 * Mr[2];
    float wgt = 1.0f / (1.0f + d2 / (a.kappa * a.kappa));
    wgt = wgt * wgt;  // Geman-McClure (odom.py:370-371)
    acc.add(ACC_RMS, wgt * rn * rn);
#define LW_ADD_H(p, q, k) acc.add(ACC_H + k, wgt * hval<p, q>(M, wx, wy, wz));
    LW_H_PAIRS(LW_ADD_H)
#undef LW_ADD_H
    acc.add(ACC_B + 0, wgt * bval<0>(Mr, wx, wy, wz));
    acc.add(ACC_B + 1, wgt * bval<1>(Mr, wx, wy, wz));
    acc.add(ACC_B + 2, wgt * bval<2>(Mr, wx, wy, wz));
    acc.add(ACC_B + 3, wgt * bval<3>(Mr, wx, wy, wz));
    acc.add(ACC_B + 4, wgt * bval<4>(Mr, wx, wy, wz));
    acc.add(ACC_B + 5, wgt * bval<5>(Mr, wx, wy, wz));
    acc.add(ACC_COUNT, 1.0f);
    acc.add(ACC_ERR, wgt * d2);
    acc.add(ACC_WEIGHT, wgt);
#define LW_ADD_G(p, q, k) acc.add(ACC_HG + k, wgt * hval<p, q>(G, wx, wy, wz));
    LW_H_PAIRS(LW_ADD_G)
#undef LW_ADD_G
    acc.add(ACC_RANGE, length3(px, py, pz));
}

// Chunk sizes: 27 neighbours in two chunks, 125 in eight.
template <int R>
struct Chunk;
template <>
struct Chunk<1> {
    static constexpr int C = 14;
};
template <>
struct Chunk<2> {
    static constexpr int C = 16;
};

template <int MODE, int R>
__global__ __launch_bounds__(ICP_BLOCK) void k_icp(IcpArgs a) {
    RegState *st = a.st;
    if (st->done) return;
    __shared__ float Tf[16];
    __shared__ int32_t s_last;
    if (threadIdx.x < 16) Tf[threadIdx.x] = st->Tf[threadIdx.x];
    __syncthreads();
    const int32_t i = blockIdx.x * ICP_BLOCK + threadIdx.x;
    int32_t hits = 0, matched = 0;
    double *row = a.partials + (size_t)blockIdx.x * 128;
    {
        Acc<MODE> acc;
        acc.zero();
        if (i < a.n) point_terms<MODE, R, Chunk<R>::C>(a, Tf, i, acc, hits, matched);
        acc.wave_reduce();
        // lane k stores entry k (lanes 53..63 idle); the reduced values are uniform across the wave
        LW_UNROLL for (int k = 0; k < ACC_USED; ++k)
            if (threadIdx.x == k) acc.store(row, k);
    }
    if (a.prof) {
        LW_UNROLL for (int off = 32; off > 0; off >>= 1) {
            hits += __shfl_xor(hits, off, 64);
            matched += __shfl_xor(matched, off, 64);
        }
        if (threadIdx.x == 0) {
            atomicAdd(&a.prof[2], (unsigned long long)hits);
            atomicAdd(&a.prof[3], (unsigned long long)matched);
        }
    }
    __threadfence();
    if (threadIdx.x == 0) s_last = (atomicAdd(&st->ticket, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    const int k = threadIdx.x;
    double v = 0.0;
    if (k < ACC_USED) {
        if (MODE == 0) {
            double sum = 0.0;
            for (int b = 0; b < (int)gridDim.x; ++b)
                sum += __hip_atomic_load(&a.partials[(size_t)b * 128 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v = sum;
        } else if (MODE == 2) {
            float sum = 0.0f;
            for (int b = 0; b < (int)gridDim.x; ++b)
                sum += (float)__hip_atomic_load(&a.partials[(size_t)b * 128 + k], __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
            v = (double)sum;
        } else {
            float s = 0.0f, c = 0.0f;
            for (int b = 0; b < (int)gridDim.x; ++b) {
                float t, e;
                two_sum(s, (float)__hip_atomic_load(&a.partials[(size_t)b * 128 + k], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT), t, e);
                s = t;
                c = c + (float)__hip_atomic_load(&a.partials[(size_t)b * 128 + 64 + k], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) + e;
            }
            v = (double)s + (double)c;
        }
    }
    st->acc[k] = v;
    if (k == 0) {
        st->ticket = 0;
        if (a.prof) {
            constexpr unsigned long long K = (2 * R + 1) * (2 * R + 1) * (2 * R + 1);
            atomicAdd(&a.prof[0], (unsigned long long)a.n);
            atomicAdd(&a.prof[1], (unsigned long long)a.n * K);
            atomicAdd(&a.prof[4], 1ull);
        }
    }
}

// The float64 step after a pass (odom.py:769-796), in one lane.
__global__ __launch_bounds__(64) void k_gn(RegState *st, gn::StepParams sp, int32_t it) {
    if (threadIdx.x != 0 || st->done) return;
    gn::StepOut o;
    int rc = gn::step(st->acc, st->T, st->T_pred, sp, o);
    st->corr = o.corr;
    if (rc == gn::STEP_FEW) {  // corr < min_corr: stop (odom.py:770-771)
        st->done = 1;
        st->status = rc;
        return;
    }
    st->rms = o.rms;
    for (int k = 0; k < 36; ++k) st->H[k] = o.H[k];
    if (rc == gn::STEP_SINGULAR) {  // LinAlgError: stop (odom.py:791-792)
        st->done = 1;
        st->status = rc;
        return;
    }
    st->n_degen = o.n_degen;
    gn::apply_twist(st->T, o.delta);
    for (int k = 0; k < 16; ++k) st->Tf[k] = (float)st->T[k];
    st->iters = it + 1;
    double nr = sqrt(o.delta[0] * o.delta[0] + o.delta[1] * o.delta[1] + o.delta[2] * o.delta[2]);
    double nt = sqrt(o.delta[3] * o.delta[3] + o.delta[4] * o.delta[4] + o.delta[5] * o.delta[5]);
    if (nr < 1e-4 && nt < 1e-4) {  // STEP_TOL (odom.py:88, 795-796)
        st->done = 1;
        st->status = 3;
    }
}

__global__ void k_reg_begin(RegState *st, const double *Tp) {
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 16; ++k) {
        st->T[k] = Tp[k];
        st->T_pred[k] = Tp[k];
        st->Tf[k] = (float)Tp[k];
    }
    for (int k = 0; k < 36; ++k) st->H[k] = (k % 7 == 0) ? 1.0 : 0.0;  // np.eye(6) (odom.py:763)
    st->rms = 0.0;
    st->cond = 1.0;
    st->done = 0;
    st->iters = 0;
    st->corr = 0;
    st->n_degen = 0;
    st->status = 0;
    st->ticket = 0;
}

__global__ void k_reg_finish(RegState *st, int32_t min_corr) {
    if (threadIdx.x != 0) return;
    st->cond = st->corr >= min_corr ? gn::cond_of(st->H) : 1.0;  // odom.py:797-800
}

template <int MODE>
hipError_t launch_mode(lw_ctx *c, const IcpArgs &a, int reach, int g) {
    if (reach == 1)
        hipLaunchKernelGGL((k_icp<MODE, 1>), dim3(g), dim3(ICP_BLOCK), 0, c->stream, a);
    else if (reach == 2)
        hipLaunchKernelGGL((k_icp<MODE, 2>), dim3(g), dim3(ICP_BLOCK), 0, c->stream, a);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

hipError_t launch_reg_begin(lw_ctx *c, const double *T_pred_dev) {
    hipLaunchKernelGGL(k_reg_begin, dim3(1), dim3(64), 0, c->stream, c->reg, T_pred_dev);
    return hipGetLastError();
}

hipError_t launch_reg_finish(lw_ctx *c, int32_t min_corr) {
    hipLaunchKernelGGL(k_reg_finish, dim3(1), dim3(64), 0, c->stream, c->reg, min_corr);
    return hipGetLastError();
}

hipError_t launch_icp_pass(lw_ctx *c, const IcpPass &p, int acc_mode) {
    if (p.n > c->frame_cap) return hipErrorInvalidValue;
    IcpArgs a;
    a.ds = c->ds_xyz;
    a.slots = c->slots;
    a.occ = c->occ;
    a.vox_prec = c->vox_prec;
    a.st = c->reg;
    a.partials = c->partials;
    a.best_slot = p.want_slots ? c->best_slot : nullptr;
    a.prof = c->prof_on ? c->prof_dev : nullptr;
    a.n = p.n;
    a.mask = c->mask;
    a.inv_v = p.inv_v;
    a.voxel = p.voxel;
    a.max_d2 = p.max_d2;
    a.kappa = p.kappa;
    int g = (p.n + ICP_BLOCK - 1) / ICP_BLOCK;
    if (g < 1) g = 1;
    const bool timed = c->prof_on && c->prof_used < c->prof_cap;
    if (timed) {
        hipError_t e = hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream);
        if (e != hipSuccess) return e;
    }
    hipError_t e;
    switch (acc_mode) {
    case 0: e = launch_mode<0>(c, a, p.reach, g); break;
    case 1: e = launch_mode<1>(c, a, p.reach, g); break;
    case 2: e = launch_mode<2>(c, a, p.reach, g); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (timed) {
        e = hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream);
        c->prof_used++;
    }
    if (e == hipSuccess && p.gn) {
        hipLaunchKernelGGL(k_gn, dim3(1), dim3(64), 0, c->stream, c->reg, p.sp, p.it);
        e = hipGetLastError();
    }
    return e;
}

}  // namespace lw
