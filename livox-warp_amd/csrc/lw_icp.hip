// Scan-to-map accumulation and the Gauss-Newton loop on the device: k_icp + k_reduce
// (python/livox_warp/odom.py:275-384) and the host step of Odometry._register (odom.py:752-803).
//
// One pass = one kernel launch of G one-wave workgroups. Each lane takes thinned points (grid-stride),
// poses them with the float32 pose, searches the (2 reach + 1)^3 voxels around it for the closest voxel
// mean (dx outer, dz inner, strict <, as odom.py:315-332), and forms the 53 float32 terms of the ACC layout
// exactly as the reference rounds them (odom.py:246-259, 269-271, 372-375). The terms are summed per lane,
// reduced across the wave with cross-lane shuffles, and each workgroup writes one partial row. The last
// workgroup to finish (atomic ticket, no waiting) sums the rows in a fixed order and, in a registration,
// runs the float64 step (lw_gn.h) and updates the pose in device memory. Later passes of a converged
// registration return at once, so a registration is max_iter launches and one readback.
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
    const float4 *vox_mu;
    const VoxPrec *vox_prec;
    RegState *st;
    double *partials;  // [G][128]: sums, compensations
    int32_t *best_slot;
    unsigned long long *prof;  // measurement counters (see lw_ctx), or null
    int32_t n;
    int32_t mask;
    float inv_v, voxel, max_d2, kappa;
    int32_t reach;
    int32_t it;
    int32_t gn;
    gn::StepParams sp;
};

#define LW_H_PAIRS(X) \
    X(0, 0, 0) X(0, 1, 1) X(0, 2, 2) X(0, 3, 3) X(0, 4, 4) X(0, 5, 5) X(1, 1, 6) X(1, 2, 7) X(1, 3, 8) X(1, 4, 9) \
    X(1, 5, 10) X(2, 2, 11) X(2, 3, 12) X(2, 4, 13) X(2, 5, 14) X(3, 3, 15) X(3, 4, 16) X(3, 5, 17) X(4, 4, 18)   \
    X(4, 5, 19) X(5, 5, 20)

// Closest voxel mean within reach (odom.py:308-334); returns the voxel id or -1, and its slot.
__device__ inline int32_t closest_mean(const IcpArgs &a, float wx, float wy, float wz, float &mx, float &my, float &mz,
                                       int32_t &best_s, int32_t &hits) {
    const int32_t bx = ifloor(wx * a.inv_v), by = ifloor(wy * a.inv_v), bz = ifloor(wz * a.inv_v);
    float best_d2 = a.max_d2;
    int32_t best_vid = -1;
    best_s = -1;
    const int R = a.reach;
    for (int dx = -R; dx <= R; ++dx)
        for (int dy = -R; dy <= R; ++dy)
            for (int dz = -R; dz <= R; ++dz) {
                int64_t key = voxel_key_ijk(bx + dx, by + dy, bz + dz);
                int32_t h = slot_hash(key) & a.mask;
                int32_t vid = -1, s = -1;
                for (int probe = 0; probe < PROBES; ++probe) {
                    int32_t t = (h + probe) & a.mask;
                    uint32_t w = a.occ[t >> 5];
                    if (!((w >> (t & 31)) & 1u)) break;
                    Slot sl = a.slots[t];
                    if (sl.key == key) {
                        vid = sl.vid;
                        s = t;
                        break;
                    }
                }
                if (vid < 0) continue;
                ++hits;
                float4 mu = a.vox_mu[vid];
                if (!(mu.w > 0.0f)) continue;  // count 0 (odom.py:324)
                float d0 = wx - mu.x, d1 = wy - mu.y, d2 = wz - mu.z;
                float dd = d0 * d0;
                dd = dd + d1 * d1;
                dd = dd + d2 * d2;
                if (dd < best_d2) {
                    best_d2 = dd;
                    best_vid = vid;
                    best_s = s;
                    mx = mu.x;
                    my = mu.y;
                    mz = mu.z;
                }
            }
    return best_vid;
}

template <int MODE>
__device__ inline void point_terms(const IcpArgs &a, const float *Tf, int32_t i, Acc<MODE> &acc, int32_t &hits,
                                   int32_t &matched) {
    const float px = a.ds[3 * (int64_t)i], py = a.ds[3 * (int64_t)i + 1], pz = a.ds[3 * (int64_t)i + 2];
    float wx, wy, wz;
    transform_point(Tf, px, py, pz, wx, wy, wz);
    float mx = 0.0f, my = 0.0f, mz = 0.0f;
    int32_t bs;
    int32_t vid = closest_mean(a, wx, wy, wz, mx, my, mz, bs, hits);
    if (a.best_slot) a.best_slot[i] = bs;
    if (vid < 0) return;
    ++matched;
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
    for (int k = 0; k < 3; ++k) {
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

// The float64 step at the end of a pass (odom.py:769-796), in one lane.
__device__ void gauss_newton(const IcpArgs &a, const double *v) {
    RegState *st = a.st;
    gn::StepOut o;
    int rc = gn::step(v, st->T, st->T_pred, a.sp, o);
    st->corr = o.corr;
    if (rc == gn::STEP_FEW) {
        st->done = 1;
        st->status = rc;
        return;
    }
    st->rms = o.rms;
    for (int k = 0; k < 36; ++k) st->H[k] = o.H[k];
    if (rc == gn::STEP_SINGULAR) {
        st->done = 1;
        st->status = rc;
        return;
    }
    st->n_degen = o.n_degen;
    gn::apply_twist(st->T, o.delta);
    for (int k = 0; k < 16; ++k) st->Tf[k] = (float)st->T[k];
    st->iters = a.it + 1;
    double nr = sqrt(o.delta[0] * o.delta[0] + o.delta[1] * o.delta[1] + o.delta[2] * o.delta[2]);
    double nt = sqrt(o.delta[3] * o.delta[3] + o.delta[4] * o.delta[4] + o.delta[5] * o.delta[5]);
    if (nr < 1e-4 && nt < 1e-4) {  // STEP_TOL (odom.py:88, 795)
        st->done = 1;
        st->status = 3;
    }
}

template <int MODE>
__global__ __launch_bounds__(ICP_BLOCK) void k_icp(IcpArgs a) {
    RegState *st = a.st;
    if (st->done) return;
    __shared__ float Tf[16];
    __shared__ int32_t s_last;
    if (threadIdx.x < 16) Tf[threadIdx.x] = st->Tf[threadIdx.x];
    __syncthreads();
    Acc<MODE> acc;
    acc.zero();
    int32_t hits = 0, matched = 0;
    for (int32_t i = blockIdx.x * ICP_BLOCK + threadIdx.x; i < a.n; i += gridDim.x * ICP_BLOCK)
        point_terms<MODE>(a, Tf, i, acc, hits, matched);
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
    acc.wave_reduce();
    double *row = a.partials + (size_t)blockIdx.x * 128;
    // lane k stores entry k (lanes 53..63 idle); the reduced values are uniform across the wave
    LW_UNROLL for (int k = 0; k < ACC_USED; ++k)
        if (threadIdx.x == k) acc.store(row, k);
    __threadfence();
    if (threadIdx.x == 0) s_last = (atomicAdd(&st->ticket, 1u) == gridDim.x - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    __shared__ double v[64];
    const int k = threadIdx.x;
    if (k < ACC_USED) {
        if (MODE == 0) {
            double sum = 0.0;
            for (int b = 0; b < (int)gridDim.x; ++b) sum += a.partials[(size_t)b * 128 + k];
            v[k] = sum;
        } else if (MODE == 2) {
            float sum = 0.0f;
            for (int b = 0; b < (int)gridDim.x; ++b) sum += (float)a.partials[(size_t)b * 128 + k];
            v[k] = (double)sum;
        } else {
            float s = 0.0f, c = 0.0f;
            for (int b = 0; b < (int)gridDim.x; ++b) {
                float t, e;
                two_sum(s, (float)a.partials[(size_t)b * 128 + k], t, e);
                s = t;
                c = c + (float)a.partials[(size_t)b * 128 + 64 + k] + e;
            }
            v[k] = (double)s + (double)c;
        }
    } else {
        v[k] = 0.0;
    }
    __syncthreads();
    if (k < 64) st->acc[k] = v[k];
    if (k == 0) {
        st->ticket = 0;
        if (a.prof) {
            const unsigned long long K = (unsigned long long)((2 * a.reach + 1) * (2 * a.reach + 1) * (2 * a.reach + 1));
            atomicAdd(&a.prof[0], (unsigned long long)a.n);
            atomicAdd(&a.prof[1], (unsigned long long)a.n * K);
            atomicAdd(&a.prof[4], 1ull);
        }
        if (a.gn) gauss_newton(a, v);
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
    IcpArgs a;
    a.ds = c->ds_xyz;
    a.slots = c->slots;
    a.occ = c->occ;
    a.vox_mu = c->vox_mu;
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
    a.reach = p.reach;
    a.it = p.it;
    a.gn = p.gn;
    a.sp = p.sp;
    int g = (p.n + ICP_BLOCK - 1) / ICP_BLOCK;
    if (g < 1) g = 1;
    if (g > ICP_MAX_GROUPS) g = ICP_MAX_GROUPS;
    const bool timed = c->prof_on && c->prof_used < c->prof_cap;
    if (timed) {
        hipError_t e = hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream);
        if (e != hipSuccess) return e;
    }
    switch (acc_mode) {
    case 0: hipLaunchKernelGGL(k_icp<0>, dim3(g), dim3(ICP_BLOCK), 0, c->stream, a); break;
    case 1: hipLaunchKernelGGL(k_icp<1>, dim3(g), dim3(ICP_BLOCK), 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL(k_icp<2>, dim3(g), dim3(ICP_BLOCK), 0, c->stream, a); break;
    default: return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (timed) {
        e = hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream);
        c->prof_used++;
    }
    return e;
}

}  // namespace lw
