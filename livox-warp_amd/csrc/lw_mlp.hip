// The learned pose-increment prior: a 152 -> 256 -> 256 -> 256 -> 6 MLP with GELU, BF16 weights and activations
// on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate), trained by these kernels:
// forward, backward, MSE loss and Adam, plus the per-frame input features. No counterpart in the reference;
// it replaces the constant-velocity guess of Odometry.predict (odom.py:591-592, used at odom.py:666) when
// enabled.
//
// Parameter vector (float32 masters, one BF16 copy for the matrix cores), row-major [in][out]:
//   W1 160 x 256 (input rows 152..159 zero), b1 256, W2 256 x 256, b2 256, W3 256 x 256, b3 256,
//   W4 256 x 16 (output columns 6..15 zero), b4 16.
// Activations for n rows: X (bf16, 160 wide), Z1..Z3 (pre-activations, fp32), H1..H3 (GELU, bf16), Y (fp32).
// Backward keeps dZ in bf16 (the matrix cores' operand) and accumulates every gradient in fp32; a bias
// gradient is the product of a row of ones with dZ, so it follows the same arithmetic as the weights'.
//
// One GEMM kernel serves every product: C (M x N) = A (M x K) B (K x N), operands addressed by two strides
// each (a transpose is a stride swap), a 64 x 64 tile per workgroup of four waves, each wave 2 x 2 blocks of
// 16 x 16, K in steps of 32 staged through LDS as bf16 with k contiguous, and a fused epilogue. Each thread
// moves one 16-byte vector of each operand a step (8 bf16 along the operand's contiguous axis, written to LDS
// as one vector when that axis is k, else transposed element by element), the next step's vectors in flight
// while the matrix cores work on the current one, with two LDS buffers and one barrier a step. The weight
// gradients (K = the batch, an output of only 160-256 x 256) split K over the grid's z axis: each split writes
// its partial tile, and a second kernel adds the splits in split order (the same sum on every run).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "../../include/livox_warp_amd.h"
#include "lw_internal.h"
#include "lw_mlp.h"

namespace lw {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int GT = 64;        // output tile edge
constexpr int GK = 32;        // K per step (one MFMA)
constexpr int LROW = GK + 8;  // LDS row: 32 bf16 + 16 B of padding (80 B rows: the 16 B fragment reads of a
                              // wave's 16 rows fall in different banks)

__device__ inline uint16_t to_bf16(float f) {  // round to nearest even (finite values)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_STORE = 0, EPI_BIAS = 1, EPI_BIAS_GELU = 2, EPI_GELU_BWD = 3 };

struct Gemm {
    const uint16_t *A;
    const uint16_t *B;
    int64_t sa0, sa1;  // A(i, k) = A[i * sa0 + k * sa1]
    int64_t sb0, sb1;  // B(k, j) = B[k * sb0 + j * sb1]
    int32_t M, N, K;
    const float *bias;  // [N] (EPI_BIAS, EPI_BIAS_GELU)
    const float *zin;   // Z for the GELU derivative (EPI_GELU_BWD), row stride ldz
    float *cf;          // fp32 output, row stride ldc
    uint16_t *cb;       // bf16 output, row stride ldh
    int64_t ldz, ldc, ldh;
    int32_t kc;         // K of one split (a multiple of GK); gridDim.z splits
    float *ws;          // the splits' partial tiles, [split][M][N] (EPI_STORE with more than one split)
    int32_t a_vec, b_vec;  // 16-byte operand loads allowed (aligned base, strides and K multiples of 8)
};

// One operand's 8 bf16 of a step, for a tile of GT rows (i or j) x GK k. Thread t takes, along k (k_contig:
// the operand's k stride is 1), row t / 4 and k (t % 4) * 8 .. + 7, or across rows, k t / 8 and rows
// (t % 8) * 8 .. + 7. Outside [0, rows) x [0, kend) the values are zero.
struct Frag {
    uint4 v;
};

__device__ inline uint16_t half_of(const uint4 &v, int q) {
    const uint32_t w = q < 2 ? v.x : (q < 4 ? v.y : (q < 6 ? v.z : v.w));
    return (uint16_t)((q & 1) ? (w >> 16) : (w & 0xFFFFu));
}

// rows r0 .. (the tile's rows), p(row, k) = base[row * s_row + k * s_k]
__device__ inline Frag fetch_frag(const uint16_t *base, int64_t s_row, int64_t s_k, int32_t r0, int32_t rows,
                                  int32_t k0, int32_t kend, bool vec) {
    const int t = threadIdx.x;
    Frag f;
    f.v = make_uint4(0u, 0u, 0u, 0u);
    uint16_t h[8];
    if (s_k == 1) {  // along k
        const int32_t r = r0 + (t >> 2), k = k0 + (t & 3) * 8;
        if (r < rows && k < kend) {
            const uint16_t *p = base + (int64_t)r * s_row + k;
            if (vec) {
                f.v = *reinterpret_cast<const uint4 *>(p);
                return f;
            }
            LW_UNROLL for (int q = 0; q < 8; ++q) h[q] = k + q < kend ? p[q] : (uint16_t)0;
        } else {
            return f;
        }
    } else {  // across rows, one k
        const int32_t k = k0 + (t >> 3), r = r0 + (t & 7) * 8;
        if (k < kend && r < rows) {
            const uint16_t *p = base + (int64_t)k * s_k + (int64_t)r * s_row;
            if (vec && r + 8 <= rows) {
                f.v = *reinterpret_cast<const uint4 *>(p);
                return f;
            }
            LW_UNROLL for (int q = 0; q < 8; ++q) h[q] = r + q < rows ? p[(int64_t)q * s_row] : (uint16_t)0;
        } else {
            return f;
        }
    }
    f.v = make_uint4((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16),
                     (uint32_t)h[4] | ((uint32_t)h[5] << 16), (uint32_t)h[6] | ((uint32_t)h[7] << 16));
    return f;
}

// The fragment into the LDS tile [row][k] (row stride LROW).
__device__ inline void stash_frag(uint16_t *tile, bool k_contig, const Frag &f) {
    const int t = threadIdx.x;
    if (k_contig) {
        *reinterpret_cast<uint4 *>(&tile[(t >> 2) * LROW + (t & 3) * 8]) = f.v;
    } else {
        const int k = t >> 3, r = (t & 7) * 8;
        LW_UNROLL for (int q = 0; q < 8; ++q) tile[(r + q) * LROW + k] = half_of(f.v, q);
    }
}

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(Gemm g) {
    __shared__ __align__(16) uint16_t As[2][GT * LROW];  // [i][k]
    __shared__ __align__(16) uint16_t Bs[2][GT * LROW];  // [j][k]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = blockIdx.y * GT, j0 = blockIdx.x * GT;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    const int32_t kbeg = blockIdx.z * g.kc, kend = min(g.K, kbeg + g.kc);
    const bool a_k = g.sa1 == 1, b_k = g.sb0 == 1;  // the operand's k axis is its contiguous one
    // B(k, j): rows are j with stride sb1, k with stride sb0
    f32x4 acc[2][2];
    LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    Frag fa = fetch_frag(g.A, g.sa0, g.sa1, i0, g.M, kbeg, kend, g.a_vec != 0);
    Frag fb = fetch_frag(g.B, g.sb1, g.sb0, j0, g.N, kbeg, kend, g.b_vec != 0);
    stash_frag(As[0], a_k, fa);
    stash_frag(Bs[0], b_k, fb);
    __syncthreads();
    int buf = 0;
    const int q = 8 * (lane >> 4), rl = lane & 15;
    for (int32_t k0 = kbeg; k0 < kend; k0 += GK) {
        const bool more = k0 + GK < kend;
        if (more) {  // the next step's operands, in flight during this step's products
            fa = fetch_frag(g.A, g.sa0, g.sa1, i0, g.M, k0 + GK, kend, g.a_vec != 0);
            fb = fetch_frag(g.B, g.sb1, g.sb0, j0, g.N, k0 + GK, kend, g.b_vec != 0);
        }
        bf16x8 af[2], bfr[2];
        LW_UNROLL for (int a = 0; a < 2; ++a)
            af[a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&As[buf][(wi + 16 * a + rl) * LROW + q]));
        LW_UNROLL for (int b = 0; b < 2; ++b)
            bfr[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&Bs[buf][(wj + 16 * b + rl) * LROW + q]));
        LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int b = 0; b < 2; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
        if (more) {
            stash_frag(As[buf ^ 1], a_k, fa);
            stash_frag(Bs[buf ^ 1], b_k, fb);
        }
        __syncthreads();
        buf ^= 1;
    }
    // lane holds rows 4 * (lane / 16) + r, column lane % 16 of each 16 x 16 block
    const bool split = gridDim.z > 1;
    LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int b = 0; b < 2; ++b) LW_UNROLL for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi + 16 * a + 4 * (lane >> 4) + r;
        const int j = j0 + wj + 16 * b + (lane & 15);
        if (i >= g.M || j >= g.N) continue;
        const float v = acc[a][b][r];
        if constexpr (EPI == EPI_STORE) {
            if (split)
                g.ws[((int64_t)blockIdx.z * g.M + i) * g.N + j] = v;
            else
                g.cf[(int64_t)i * g.ldc + j] = v;
        } else if constexpr (EPI == EPI_BIAS) {
            g.cf[(int64_t)i * g.ldc + j] = v + g.bias[j];
        } else if constexpr (EPI == EPI_BIAS_GELU) {
            const float z = v + g.bias[j];
            g.cf[(int64_t)i * g.ldc + j] = z;
            g.cb[(int64_t)i * g.ldh + j] = to_bf16(gelu(z));
        } else {
            g.cb[(int64_t)i * g.ldh + j] = to_bf16(v * gelu_grad(g.zin[(int64_t)i * g.ldz + j]));
        }
    }
}

// The splits' partial tiles added in split order: C(i, j) = sum_z ws[z][i][j].
__global__ __launch_bounds__(256) void k_split_sum(const float *ws, int32_t splits, int32_t M, int32_t N, float *cf,
                                                   int64_t ldc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t mn = (int64_t)M * N;
    if (e >= mn) return;
    float s = 0.0f;
    for (int32_t z = 0; z < splits; ++z) s += ws[z * mn + e];
    cf[(e / N) * ldc + e % N] = s;
}

// x (n x 152 fp32, row stride ldx) -> X (n x 160 bf16, columns 152..159 zero)
__global__ void k_to_bf16_rows(const float *x, int64_t ldx, int32_t n, uint16_t *X) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)n * MLP_IN_PAD) return;
    const int64_t i = e / MLP_IN_PAD;
    const int c = (int)(e % MLP_IN_PAD);
    X[e] = c < MLP_IN ? to_bf16(x[i * ldx + c]) : (uint16_t)0;
}

// loss = mean over rows and the 6 outputs of (y - t)^2; dY (bf16, 16 wide) = d loss / d y, zero beyond column 6.
// One workgroup: the sum runs in a fixed order.
__global__ __launch_bounds__(256) void k_mse(const float *y, const float *t, int32_t n, uint16_t *dY, float *loss) {
    __shared__ float part[256];
    const float scale = 2.0f / (float)(6 * (int64_t)n);
    float s = 0.0f;
    for (int32_t i = threadIdx.x; i < n; i += 256) {
        LW_UNROLL for (int j = 0; j < MLP_OUT_PAD; ++j) {
            float g = 0.0f;
            if (j < MLP_OUT) {
                const float d = y[(int64_t)i * MLP_OUT + j] - t[(int64_t)i * MLP_OUT + j];
                s += d * d;
                g = scale * d;
            }
            dY[(int64_t)i * MLP_OUT_PAD + j] = to_bf16(g);
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float tot = 0.0f;
        for (int k = 0; k < 256; ++k) tot += part[k];
        *loss = tot / (float)(6 * (int64_t)n);
    }
}

// Adam (bias-corrected), then the bf16 copy of the updated weight.
__global__ void k_adam(float *w, float *m, float *v, const float *g, uint16_t *wb, int32_t P, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    const float wp = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    w[p] = wp;
    wb[p] = to_bf16(wp);
}

__global__ void k_params_bf16(const float *w, uint16_t *wb, int32_t P) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) wb[p] = to_bf16(w[p]);
}

__global__ void k_fill_u16(uint16_t *p, int32_t n, uint16_t v) {
    const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) p[e] = v;
}

// The 128 histogram inputs of one frame (points in the sensor frame): 64 range bins over [0, r_max) and 64
// elevation bins over [-e_max, e_max), each as count * 64 / m (a flat histogram reads 1); out-of-range values
// go to the end bins. The 24 motion inputs come with the arguments.
__global__ __launch_bounds__(1024) void k_features(const float *xyz, int32_t m, FeatArgs f, float *out) {
    __shared__ int32_t hr[64], he[64];
    if (threadIdx.x < 64) hr[threadIdx.x] = he[threadIdx.x] = 0;
    __syncthreads();
    const float rs = 64.0f / f.r_max, es = 32.0f / f.e_max;
    for (int32_t i = threadIdx.x; i < m; i += blockDim.x) {
        const float x = xyz[3 * (int64_t)i], y = xyz[3 * (int64_t)i + 1], z = xyz[3 * (int64_t)i + 2];
        float h = x * x;
        h = h + y * y;
        const float r = sqrtf(h + z * z);
        const float e = atan2f(z, sqrtf(h));
        const int br = min(max((int)floorf(r * rs), 0), 63);
        const int be = min(max((int)floorf((e + f.e_max) * es), 0), 63);
        atomicAdd(&hr[br], 1);
        atomicAdd(&he[be], 1);
    }
    __syncthreads();
    const float inv = m > 0 ? 64.0f / (float)m : 0.0f;
    if (threadIdx.x < MLP_MOTION) out[threadIdx.x] = f.motion[threadIdx.x];
    if (threadIdx.x < 64) {
        out[MLP_MOTION + threadIdx.x] = (float)hr[threadIdx.x] * inv;
        out[MLP_MOTION + 64 + threadIdx.x] = (float)he[threadIdx.x] * inv;
    }
}

constexpr int SPLIT_MAX = 32;                            // splits of one product
constexpr int64_t SPLIT_WS = (int64_t)SPLIT_MAX * MLP_HID * MLP_HID;  // floats of the partial tiles

// 16-byte loads of an operand: its base 16-byte aligned, its row stride a multiple of 8 (its k stride is
// then 1, or its k stride a multiple of 8 and its row stride 1), K a multiple of 8.
inline bool vec_ok(const uint16_t *p, int64_t s_row, int64_t s_k, int32_t K) {
    if (((uintptr_t)p & 15) != 0 || (K & 7) != 0) return false;
    if (s_k == 1) return (s_row & 7) == 0;
    return s_row == 1 && (s_k & 7) == 0;
}

// ws: room for SPLIT_WS floats, or null (no split)
template <int EPI>
hipError_t gemm(hipStream_t s, Gemm g, float *ws = nullptr) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipSuccess;
    g.a_vec = vec_ok(g.A, g.sa0, g.sa1, g.K) ? 1 : 0;
    g.b_vec = vec_ok(g.B, g.sb1, g.sb0, g.K) ? 1 : 0;
    const int32_t tiles = ((g.N + GT - 1) / GT) * ((g.M + GT - 1) / GT), steps = (g.K + GK - 1) / GK;
    int32_t splits = 1;
    if (EPI == EPI_STORE && ws && (int64_t)g.M * g.N * SPLIT_MAX <= SPLIT_WS)  // fill the device: >= 256 tiles
        splits = std::min(std::min(steps, SPLIT_MAX), std::max(1, (256 + tiles - 1) / tiles));
    const int32_t per = (steps + splits - 1) / splits;
    g.kc = per * GK;
    splits = (steps + per - 1) / per;
    g.ws = ws;
    const dim3 grid((unsigned)((g.N + GT - 1) / GT), (unsigned)((g.M + GT - 1) / GT), (unsigned)splits);
    hipLaunchKernelGGL(k_gemm<EPI>, grid, dim3(256), 0, s, g);
    if (splits > 1) {
        const int64_t mn = (int64_t)g.M * g.N;
        hipLaunchKernelGGL(k_split_sum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, s, (const float *)ws, splits,
                           g.M, g.N, g.cf, g.ldc);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_features(hipStream_t s, const float *xyz, int32_t m, const FeatArgs &f, float *out) {
    hipLaunchKernelGGL(k_features, dim3(1), dim3(1024), 0, s, xyz, m, f, out);
    return hipGetLastError();
}

}  // namespace lw

// ---- the C ABI (include/livox_warp_amd.h) --------------------------------------------------------------

using namespace lw;

struct lw_mlp {
    hipStream_t stream = nullptr;
    int32_t cap = 0;  // rows
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *wb = nullptr;
    uint16_t *X = nullptr, *H[3] = {nullptr, nullptr, nullptr}, *dZ[2] = {nullptr, nullptr}, *dY = nullptr;
    uint16_t *ones = nullptr;
    float *Z[3] = {nullptr, nullptr, nullptr};
    float *ws = nullptr;  // split-K partial tiles of the weight gradients
    float *ystep = nullptr;  // the training step's outputs (n x 6)
    // the captured training step (lw_mlp_train_step): replayed while its pointers and row count stay the same,
    // with the Adam node's scalars updated before each replay
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    hipGraphNode_t adam_node = nullptr;
    hipKernelNodeParams adam_params{};
    const void *step_key[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t step_ldx = 0;
    int32_t step_n = 0;
    int32_t rows = 0;  // rows of the last forward
};

namespace {

constexpr int32_t OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IN_PAD * MLP_HID, OFF_W2 = OFF_B1 + MLP_HID,
                  OFF_B2 = OFF_W2 + MLP_HID * MLP_HID, OFF_W3 = OFF_B2 + MLP_HID, OFF_B3 = OFF_W3 + MLP_HID * MLP_HID,
                  OFF_W4 = OFF_B3 + MLP_HID, OFF_B4 = OFF_W4 + MLP_HID * MLP_OUT_PAD, PARAMS = OFF_B4 + MLP_OUT_PAD;

#define MLP_CK(x)                              \
    do {                                       \
        hipError_t e_ = (x);                   \
        if (e_ != hipSuccess) return -(int)e_; \
    } while (0)

void drop_step_graph(lw_mlp *p) {
    if (p->step_exec) (void)hipGraphExecDestroy(p->step_exec);
    if (p->step_graph) (void)hipGraphDestroy(p->step_graph);
    p->step_exec = nullptr;
    p->step_graph = nullptr;
    p->adam_node = nullptr;
}

void mlp_free(lw_mlp *p) {
    drop_step_graph(p);
    void *bufs[] = {p->w, p->g, p->m, p->v, p->wb, p->X, p->H[0], p->H[1], p->H[2], p->dZ[0], p->dZ[1], p->dY,
                    p->ones, p->Z[0], p->Z[1], p->Z[2], p->ws, p->ystep};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
}

template <class T>
hipError_t alloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
}

// Linear layer `l` (0..3) of the forward pass over n rows: input `in` (bf16, row stride k_pad).
Gemm layer(const lw_mlp *p, int l, const uint16_t *in, int32_t n) {
    static const int32_t w_off[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, b_off[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    Gemm g{};
    const int32_t k = l == 0 ? MLP_IN_PAD : MLP_HID, ncols = l == 3 ? MLP_OUT_PAD : MLP_HID;
    g.A = in;
    g.sa0 = k;
    g.sa1 = 1;
    g.B = p->wb + w_off[l];
    g.sb0 = ncols;
    g.sb1 = 1;
    g.M = n;
    g.N = l == 3 ? MLP_OUT : MLP_HID;
    g.K = k;
    g.bias = p->w + b_off[l];
    return g;
}

}  // namespace

extern "C" {

int lw_mlp_param_count(void) { return PARAMS; }

int lw_mlp_create(lw_mlp **out, int32_t batch_cap, void *stream) {
    if (!out || batch_cap <= 0 || batch_cap > (1 << 22)) return LW_EINVAL;
    lw_mlp *p = new (std::nothrow) lw_mlp();
    if (!p) return LW_EINVAL;
    p->stream = (hipStream_t)stream;
    p->cap = batch_cap;
    const size_t n = (size_t)batch_cap;
    hipError_t e = hipSuccess;
    if (e == hipSuccess) e = alloc(&p->w, PARAMS);
    if (e == hipSuccess) e = alloc(&p->g, PARAMS);
    if (e == hipSuccess) e = alloc(&p->m, PARAMS);
    if (e == hipSuccess) e = alloc(&p->v, PARAMS);
    if (e == hipSuccess) e = alloc(&p->wb, PARAMS);
    if (e == hipSuccess) e = alloc(&p->X, n * MLP_IN_PAD);
    for (int l = 0; l < 3 && e == hipSuccess; ++l) e = alloc(&p->H[l], n * MLP_HID);
    for (int l = 0; l < 3 && e == hipSuccess; ++l) e = alloc(&p->Z[l], n * MLP_HID);
    for (int l = 0; l < 2 && e == hipSuccess; ++l) e = alloc(&p->dZ[l], n * MLP_HID);
    if (e == hipSuccess) e = alloc(&p->dY, n * MLP_OUT_PAD);
    if (e == hipSuccess) e = alloc(&p->ones, n);
    if (e == hipSuccess) e = alloc(&p->ws, (size_t)SPLIT_WS);
    if (e == hipSuccess) e = alloc(&p->ystep, n * MLP_OUT);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_fill_u16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->stream, p->ones,
                           (int32_t)n, (uint16_t)0x3F80u);  // bf16 1.0
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) {
        mlp_free(p);
        delete p;
        return -(int)e;
    }
    *out = p;
    return 0;
}

int lw_mlp_destroy(lw_mlp *p) {
    if (!p) return LW_EINVAL;
    (void)hipStreamSynchronize(p->stream);
    mlp_free(p);
    delete p;
    return 0;
}

int lw_mlp_set_params(lw_mlp *p, const float *host) {
    if (!p || !host) return LW_EINVAL;
    MLP_CK(hipMemcpyAsync(p->w, host, PARAMS * sizeof(float), hipMemcpyHostToDevice, p->stream));
    MLP_CK(hipMemsetAsync(p->m, 0, PARAMS * sizeof(float), p->stream));
    MLP_CK(hipMemsetAsync(p->v, 0, PARAMS * sizeof(float), p->stream));
    hipLaunchKernelGGL(k_params_bf16, dim3((PARAMS + 255) / 256), dim3(256), 0, p->stream, p->w, p->wb, PARAMS);
    MLP_CK(hipGetLastError());
    MLP_CK(hipStreamSynchronize(p->stream));
    return 0;
}

int lw_mlp_get_params(lw_mlp *p, float *host) {
    if (!p || !host) return LW_EINVAL;
    MLP_CK(hipMemcpyAsync(host, p->w, PARAMS * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    MLP_CK(hipStreamSynchronize(p->stream));
    return 0;
}

int lw_mlp_get_grads(lw_mlp *p, float *host) {
    if (!p || !host) return LW_EINVAL;
    MLP_CK(hipMemcpyAsync(host, p->g, PARAMS * sizeof(float), hipMemcpyDeviceToHost, p->stream));
    MLP_CK(hipStreamSynchronize(p->stream));
    return 0;
}

int lw_mlp_forward(lw_mlp *p, const float *x, int64_t ldx, int32_t n, float *y) {
    if (!p || !x || !y || n <= 0 || n > p->cap || ldx < MLP_IN) return LW_EINVAL;
    const int64_t e = (int64_t)n * MLP_IN_PAD;
    hipLaunchKernelGGL(k_to_bf16_rows, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, p->stream, x, ldx, n, p->X);
    MLP_CK(hipGetLastError());
    const uint16_t *in = p->X;
    for (int l = 0; l < 3; ++l) {
        Gemm g = layer(p, l, in, n);
        g.cf = p->Z[l];
        g.ldc = MLP_HID;
        g.cb = p->H[l];
        g.ldh = MLP_HID;
        MLP_CK(gemm<EPI_BIAS_GELU>(p->stream, g));
        in = p->H[l];
    }
    Gemm g = layer(p, 3, in, n);
    g.cf = y;
    g.ldc = MLP_OUT;
    MLP_CK(gemm<EPI_BIAS>(p->stream, g));
    p->rows = n;
    return 0;
}

int lw_mlp_loss(lw_mlp *p, const float *y, const float *target, int32_t n, float *loss) {
    if (!p || !y || !target || !loss || n <= 0 || n > p->cap) return LW_EINVAL;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, p->stream, y, target, n, p->dY, loss);
    MLP_CK(hipGetLastError());
    return 0;
}

int lw_mlp_backward(lw_mlp *p, int32_t n) {
    if (!p || n <= 0 || n != p->rows) return LW_EINVAL;
    static const int32_t w_off[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4}, b_off[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    const uint16_t *dout = p->dY;  // gradient at layer l's output (bf16), row stride dstride
    int32_t dstride = MLP_OUT_PAD, dcols = MLP_OUT_PAD;
    for (int l = 3; l >= 0; --l) {
        const uint16_t *in = l == 0 ? p->X : p->H[l - 1];
        const int32_t kin = l == 0 ? MLP_IN_PAD : MLP_HID;
        const int32_t wcols = l == 3 ? MLP_OUT_PAD : MLP_HID;
        Gemm g{};  // dW_l = in^T dout: (kin x dcols), reduction over the n rows
        g.A = in;
        g.sa0 = 1;
        g.sa1 = kin;
        g.B = dout;
        g.sb0 = dstride;
        g.sb1 = 1;
        g.M = kin;
        g.N = dcols;
        g.K = n;
        g.cf = p->g + w_off[l];
        g.ldc = wcols;
        MLP_CK(gemm<EPI_STORE>(p->stream, g, p->ws));
        Gemm gb{};  // db_l = 1^T dout
        gb.A = p->ones;
        gb.sa0 = 0;
        gb.sa1 = 1;
        gb.B = dout;
        gb.sb0 = dstride;
        gb.sb1 = 1;
        gb.M = 1;
        gb.N = dcols;
        gb.K = n;
        gb.cf = p->g + b_off[l];
        gb.ldc = wcols;
        MLP_CK(gemm<EPI_STORE>(p->stream, gb, p->ws));
        if (l == 0) break;
        Gemm gd{};  // dZ_{l-1} = (dout W_l^T) * gelu'(Z_{l-1})
        gd.A = dout;
        gd.sa0 = dstride;
        gd.sa1 = 1;
        gd.B = p->wb + w_off[l];
        gd.sb0 = 1;
        gd.sb1 = wcols;
        gd.M = n;
        gd.N = MLP_HID;
        gd.K = dcols;
        gd.zin = p->Z[l - 1];
        gd.ldz = MLP_HID;
        uint16_t *dz = p->dZ[l & 1];
        gd.cb = dz;
        gd.ldh = MLP_HID;
        MLP_CK(gemm<EPI_GELU_BWD>(p->stream, gd));
        dout = dz;
        dstride = MLP_HID;
        dcols = MLP_HID;
    }
    return 0;
}

int lw_mlp_adam(lw_mlp *p, float lr, float beta1, float beta2, float eps, int32_t step) {
    if (!p || step < 1 || !(lr >= 0.0f) || !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f))
        return LW_EINVAL;
    const float c1 = 1.0f - (float)std::pow((double)beta1, (double)step);
    const float c2 = 1.0f - (float)std::pow((double)beta2, (double)step);
    hipLaunchKernelGGL(k_adam, dim3((PARAMS + 255) / 256), dim3(256), 0, p->stream, p->w, p->m, p->v, p->g, p->wb,
                       PARAMS, lr, beta1, beta2, eps, c1, c2);
    MLP_CK(hipGetLastError());
    return 0;
}

int lw_mlp_set_stream(lw_mlp *p, void *stream) {
    if (!p) return LW_EINVAL;
    if (p->stream != (hipStream_t)stream) drop_step_graph(p);
    p->stream = (hipStream_t)stream;
    return 0;
}

int lw_mlp_train_step(lw_mlp *p, const float *x, int64_t ldx, const float *target, int32_t n, float lr, float beta1,
                      float beta2, float eps, int32_t step, float *loss, int32_t use_graph) {
    if (!p || !x || !target || !loss || n <= 0 || n > p->cap || ldx < MLP_IN || step < 1 || !(lr >= 0.0f) ||
        !(beta1 >= 0.0f && beta1 < 1.0f) || !(beta2 >= 0.0f && beta2 < 1.0f))
        return LW_EINVAL;
    float *y = p->ystep;
    if (!use_graph || p->stream == nullptr) {  // a captured graph needs a stream other than the null stream
        int rc = lw_mlp_forward(p, x, ldx, n, y);
        if (!rc) rc = lw_mlp_loss(p, y, target, n, loss);
        if (!rc) rc = lw_mlp_backward(p, n);
        if (!rc) rc = lw_mlp_adam(p, lr, beta1, beta2, eps, step);
        return rc;
    }
    const void *key[4] = {x, target, loss, y};
    if (!p->step_exec || std::memcmp(key, p->step_key, sizeof(key)) != 0 || p->step_ldx != ldx || p->step_n != n) {
        drop_step_graph(p);
        MLP_CK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        int rc = lw_mlp_forward(p, x, ldx, n, y);
        if (!rc) rc = lw_mlp_loss(p, y, target, n, loss);
        if (!rc) rc = lw_mlp_backward(p, n);
        if (!rc) rc = lw_mlp_adam(p, lr, beta1, beta2, eps, step);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(p->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        MLP_CK(e);
        p->step_graph = g;
        // the Adam node: the one kernel node whose arguments change from step to step
        size_t count = 0;
        MLP_CK(hipGraphGetNodes(g, nullptr, &count));
        hipGraphNode_t nodes[64];
        if (count > 64) return LW_EINVAL;
        MLP_CK(hipGraphGetNodes(g, nodes, &count));
        for (size_t k = 0; k < count; ++k) {
            hipGraphNodeType t;
            MLP_CK(hipGraphNodeGetType(nodes[k], &t));
            if (t != hipGraphNodeTypeKernel) continue;
            hipKernelNodeParams kp{};
            MLP_CK(hipGraphKernelNodeGetParams(nodes[k], &kp));
            if (kp.func == reinterpret_cast<const void *>(&k_adam)) {
                p->adam_node = nodes[k];
                p->adam_params = kp;
            }
        }
        if (!p->adam_node) {
            drop_step_graph(p);
            return LW_EINVAL;
        }
        MLP_CK(hipGraphInstantiate(&p->step_exec, g, nullptr, nullptr, 0));
        std::memcpy(p->step_key, key, sizeof(key));
        p->step_ldx = ldx;
        p->step_n = n;
    }
    // this step's Adam scalars (the node keeps the ones of the step it was captured at)
    float *w = p->w, *m = p->m, *v = p->v, *gr = p->g;
    uint16_t *wb = p->wb;
    int32_t P = PARAMS;
    float c1 = 1.0f - (float)std::pow((double)beta1, (double)step);
    float c2 = 1.0f - (float)std::pow((double)beta2, (double)step);
    void *args[] = {&w, &m, &v, &gr, &wb, &P, &lr, &beta1, &beta2, &eps, &c1, &c2};
    hipKernelNodeParams kp = p->adam_params;
    kp.kernelParams = args;
    kp.extra = nullptr;
    MLP_CK(hipGraphExecKernelNodeSetParams(p->step_exec, p->adam_node, &kp));
    MLP_CK(hipGraphLaunch(p->step_exec, p->stream));
    return 0;
}

}  // extern "C"
