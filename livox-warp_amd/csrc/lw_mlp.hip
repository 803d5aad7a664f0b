// The learned pose-increment prior: a 152 -> 256 -> 256 -> 256 -> 6 MLP with GELU, BF16 weights and activations
// on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate), trained by these kernels:
// forward, backward, MSE loss and Adam, plus the per-frame input features. No counterpart in the reference;
// it replaces the constant-velocity guess of Odometry.predict (odom.py:591-592, used at odom.py:666) when
// enabled.
//
// Parameter vector (float32 masters, one BF16 copy for the matrix cores, then the BF16 weights transposed),
// row-major [in][out]:
//   W1 160 x 256 (input rows 152..159 zero), b1 256, W2 256 x 256, b2 256, W3 256 x 256, b3 256,
//   W4 256 x 16 (output columns 6..15 zero), b4 16.
// Activations for n rows: X (bf16, 160 wide), Z1..Z3 (pre-activations, fp32), H1..H3 (GELU, bf16), Y (fp32).
// Backward keeps dZ in bf16 (the matrix cores' operand) and accumulates every gradient in fp32; a bias
// gradient is the product of a row of ones with dZ, so it follows the same arithmetic as the weights'.
//
// The forward is one kernel (k_mlp_fwd): 32 rows a workgroup through all four layers, the activations in LDS.
// The backward's products go through one GEMM kernel: C (M x N) = A (M x K) B (K x N), operands addressed by
// two strides each (a transpose is a stride swap), a 64 x 64 tile per workgroup of four waves, each wave
// 2 x 2 blocks of 16 x 16 on the matrix cores. Each thread moves 16-byte vectors of each operand (8 bf16 along
// the operand's contiguous axis, written to LDS as one vector when that axis is k, else transposed element by
// element), every fragment of up to 256 of K requested at once, so a product over K <= 256 waits on memory
// once (the latency of a step, not its bandwidth, bounds these small products). A layer's weight gradient
// (K = the batch, an output of only 161-257 x 256) splits K over many workgroups and shares its launch with
// the gradient flowing into the layer below; the last kernel adds every layer's splits in split order (the
// same sum on every run) and, in a training step, takes the Adam step and the loss. A step is six launches,
// which a caller can replay from a captured hipGraph.
#include <algorithm>
#include <climits>
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
constexpr int GK = 32;        // K of one MFMA and of one operand fragment
constexpr int KSUB = 8;       // fragments a stage: a stage stages GS = 256 of K through LDS
constexpr int GS = GK * KSUB;
constexpr int SPLIT_MAX = 32;  // splits of one product over K
constexpr int LROW = GS + 8;  // LDS row: 256 bf16 + 16 B of padding (528 B rows: the 16 B fragment reads of a
                              // wave's 16 rows fall in different banks)

constexpr int32_t OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_IN_PAD * MLP_HID, OFF_W2 = OFF_B1 + MLP_HID,
                  OFF_B2 = OFF_W2 + MLP_HID * MLP_HID, OFF_W3 = OFF_B2 + MLP_HID, OFF_B3 = OFF_W3 + MLP_HID * MLP_HID,
                  OFF_W4 = OFF_B3 + MLP_HID, OFF_B4 = OFF_W4 + MLP_HID * MLP_OUT_PAD, PARAMS = OFF_B4 + MLP_OUT_PAD;
// The weights again as bf16, transposed ([out][in]), after the bf16 copy of the parameter vector: the fused
// forward's matrix-core operands, 16 bytes of one output's inputs per load.
constexpr int32_t T_W1 = 0, T_W2 = T_W1 + MLP_HID * MLP_IN_PAD, T_W3 = T_W2 + MLP_HID * MLP_HID,
                  T_W4 = T_W3 + MLP_HID * MLP_HID, T_END = T_W4 + MLP_OUT_PAD * MLP_HID;

__device__ inline uint16_t to_bf16(float f) {  // round to nearest even (finite values)
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Parameter p's bf16 value into the parameter-vector copy wb and, for a weight, into the transposed copy that
// follows it (wb + PARAMS).
__device__ inline void put_bf16(uint16_t *wb, int32_t p, uint16_t h) {
    wb[p] = h;
    uint16_t *wt = wb + PARAMS;
    if (p < OFF_B1) {
        const int32_t e = p - OFF_W1;
        wt[T_W1 + (e % MLP_HID) * MLP_IN_PAD + e / MLP_HID] = h;
    } else if (p >= OFF_W2 && p < OFF_B2) {
        const int32_t e = p - OFF_W2;
        wt[T_W2 + (e % MLP_HID) * MLP_HID + e / MLP_HID] = h;
    } else if (p >= OFF_W3 && p < OFF_B3) {
        const int32_t e = p - OFF_W3;
        wt[T_W3 + (e % MLP_HID) * MLP_HID + e / MLP_HID] = h;
    } else if (p >= OFF_W4 && p < OFF_B4) {
        const int32_t e = p - OFF_W4;
        wt[T_W4 + (e % MLP_OUT_PAD) * MLP_HID + e / MLP_OUT_PAD] = h;
    }
}

__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum { EPI_STORE = 0, EPI_GELU_BWD = 1 };

struct Gemm {
    const uint16_t *A;
    const uint16_t *B;
    int64_t sa0, sa1;  // A(i, k) = A[i * sa0 + k * sa1]
    int64_t sb0, sb1;  // B(k, j) = B[k * sb0 + j * sb1]
    int32_t M, N, K;
    const float *zin;   // Z for the GELU derivative (EPI_GELU_BWD), row stride ldz
    float *cf;          // fp32 output, row stride ldc
    uint16_t *cb;       // bf16 output, row stride ldh
    int64_t ldz, ldc, ldh;
    int32_t kc;         // K of one split (a multiple of GK); gridDim.z splits
    float *ws;          // the splits' partial tiles, [split][M][N] (EPI_STORE with more than one split)
    int32_t a_ones;        // A's rows from here on read 1.0 (a bias gradient as one more output row)
    int32_t gx, gy, gz;    // the tile grid: column tiles, row tiles, splits of K
};

// One operand's 8 bf16 of a step, for a tile of GT rows (i or j) x GK k. ALONG (the operand's k stride is 1,
// its row stride a multiple of 8, K a multiple of 8): thread t takes row t / 4 and k (t % 4) * 8 .. + 7.
// ACROSS (its row stride is 1, its k stride a multiple of 8): thread t takes k t / 8 and rows (t % 8) * 8 ..
// + 7. Outside [0, rows) x [0, kend) the values are zero; ACROSS rows from `ones` on read 1.0. Each kernel
// instance carries only the modes of its operands: a small kernel's time goes largely to instruction fetch.
enum { ALONG = 0, ACROSS = 1 };

__device__ inline uint16_t half_of(const uint4 &v, int q) {
    const uint32_t w = q < 2 ? v.x : (q < 4 ? v.y : (q < 6 ? v.z : v.w));
    return (uint16_t)((q & 1) ? (w >> 16) : (w & 0xFFFFu));
}

// p(row, k) = base[row * s_row + k * s_k], rows r0 ..
template <int MODE>
__device__ inline uint4 fetch_frag(const uint16_t *base, int64_t s_row, int64_t s_k, int32_t r0, int32_t rows,
                                   int32_t k0, int32_t kend, int32_t ones) {
    const int t = threadIdx.x;
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    if constexpr (MODE == ALONG) {
        const int32_t r = r0 + (t >> 2), k = k0 + (t & 3) * 8;
        const bool ok = r < rows && k < kend;
        const uint4 v = *reinterpret_cast<const uint4 *>(base + (ok ? (int64_t)r * s_row + k : 0));
        return ok ? v : zero;
    } else {
        const int32_t k = k0 + (t >> 3), r = r0 + (t & 7) * 8;
        const bool full = k < kend && r + 8 <= rows && r + 8 <= ones;
        const uint4 v = *reinterpret_cast<const uint4 *>(base + (full ? (int64_t)k * s_k + r : 0));
        if (full) return v;
        if (k >= kend || r >= rows) return zero;
        const uint16_t *p = base + (int64_t)k * s_k + r;  // a tile edge, element by element
        uint16_t h[8];
        LW_UNROLL for (int q = 0; q < 8; ++q)
            h[q] = r + q >= rows ? (uint16_t)0 : (r + q >= ones ? (uint16_t)0x3F80u : p[q]);
        return make_uint4((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16),
                          (uint32_t)h[4] | ((uint32_t)h[5] << 16), (uint32_t)h[6] | ((uint32_t)h[7] << 16));
    }
}

// The fragment into the LDS tile [row][k] (row stride LROW), at k offset koff.
template <int MODE>
__device__ inline void stash_frag(uint16_t *tile, const uint4 &v, int koff) {
    const int t = threadIdx.x;
    if constexpr (MODE == ALONG) {
        *reinterpret_cast<uint4 *>(&tile[(t >> 2) * LROW + koff + (t & 3) * 8]) = v;
    } else {
        const int k = koff + (t >> 3), r = (t & 7) * 8;
        LW_UNROLL for (int q = 0; q < 8; ++q) tile[(r + q) * LROW + k] = half_of(v, q);
    }
}

// Tile (bx, by) of split bz of product g. As, Bs: the LDS tiles [row][k].
template <int EPI, int AM, int BM>
__device__ inline void gemm_tile(const Gemm &g, int bx, int by, int bz, uint16_t *As, uint16_t *Bs) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = by * GT, j0 = bx * GT;
    const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
    const int32_t kbeg = bz * g.kc, kend = min(g.K, kbeg + g.kc);
    // B(k, j): rows are j with stride sb1, k with stride sb0
    f32x4 acc[2][2];
    LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int q = 8 * (lane >> 4), rl = lane & 15;
    // a stage: every fragment of up to GS of K requested at once, one wait, one barrier, then the products
    for (int32_t k0 = kbeg; k0 < kend; k0 += GS) {
        const int nf = min(KSUB, (kend - k0 + GK - 1) / GK);  // the same for the whole workgroup
        uint4 fa[KSUB], fb[KSUB];
        LW_UNROLL for (int u = 0; u < KSUB; ++u) {
            if (u < nf) {
                fa[u] = fetch_frag<AM>(g.A, g.sa0, g.sa1, i0, g.M, k0 + u * GK, kend, g.a_ones);
                fb[u] = fetch_frag<BM>(g.B, g.sb1, g.sb0, j0, g.N, k0 + u * GK, kend, INT_MAX);
            }
        }
        if (k0 != kbeg) __syncthreads();  // the last stage's products are done with the tiles
        LW_UNROLL for (int u = 0; u < KSUB; ++u) {
            if (u < nf) {
                stash_frag<AM>(As, fa[u], u * GK);
                stash_frag<BM>(Bs, fb[u], u * GK);
            }
        }
        __syncthreads();
        LW_UNROLL for (int u = 0; u < KSUB; ++u) {
            if (u < nf) {
                bf16x8 af[2], bfr[2];
                LW_UNROLL for (int a = 0; a < 2; ++a)
                    af[a] = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4 *>(&As[(wi + 16 * a + rl) * LROW + u * GK + q]));
                LW_UNROLL for (int b = 0; b < 2; ++b)
                    bfr[b] = __builtin_bit_cast(
                        bf16x8, *reinterpret_cast<const uint4 *>(&Bs[(wj + 16 * b + rl) * LROW + u * GK + q]));
                LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
            }
        }
    }
    // lane holds rows 4 * (lane / 16) + r, column lane % 16 of each 16 x 16 block
    LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int b = 0; b < 2; ++b) LW_UNROLL for (int r = 0; r < 4; ++r) {
        const int i = i0 + wi + 16 * a + 4 * (lane >> 4) + r;
        const int j = j0 + wj + 16 * b + (lane & 15);
        if (i >= g.M || j >= g.N) continue;
        const float v = acc[a][b][r];
        if constexpr (EPI == EPI_STORE) {
            if (g.ws)  // a split's partial tile
                g.ws[((int64_t)bz * g.M + i) * g.N + j] = v;
            else
                g.cf[(int64_t)i * g.ldc + j] = v;
        } else {
            g.cb[(int64_t)i * g.ldh + j] = to_bf16(v * gelu_grad(g.zin[(int64_t)i * g.ldz + j]));
        }
    }
}

template <int EPI, int AM, int BM>
__global__ __launch_bounds__(256) void k_gemm(Gemm g) {
    __shared__ __align__(16) uint16_t As[GT * LROW];
    __shared__ __align__(16) uint16_t Bs[GT * LROW];
    gemm_tile<EPI, AM, BM>(g, blockIdx.x, blockIdx.y, blockIdx.z, As, Bs);
}

// Two independent products in one launch (the backward's weight gradient of one layer beside the gradient
// flowing into the layer below): blocks [0, tiles of g1) take g1, the rest g2.
template <int E1, int A1, int B1, int E2, int A2, int B2>
__global__ __launch_bounds__(256) void k_gemm_pair(Gemm g1, Gemm g2) {
    __shared__ __align__(16) uint16_t As[GT * LROW];
    __shared__ __align__(16) uint16_t Bs[GT * LROW];
    int b = blockIdx.x;
    const int n1 = g1.gx * g1.gy * g1.gz;
    if (b < n1) {
        gemm_tile<E1, A1, B1>(g1, b % g1.gx, (b / g1.gx) % g1.gy, b / (g1.gx * g1.gy), As, Bs);
    } else {
        b -= n1;
        gemm_tile<E2, A2, B2>(g2, b % g2.gx, (b / g2.gx) % g2.gy, b / (g2.gx * g2.gy), As, Bs);
    }
}

// The whole forward for FR rows a workgroup, the activations staying in LDS between layers: the input rows
// converted to bf16, then each layer's products on the matrix cores with the same fragments, k order and
// epilogue arithmetic as k_gemm (so the same values), each wave 64 of the 256 outputs (the output layer's 16:
// wave 0). The weight operands come straight from the transposed bf16 copy (L2-resident: every workgroup reads
// all of it). Z, H and X are kept for the backward where given; with a target, the output layer's epilogue
// forms the loss terms and d loss / d y (k_mse's arithmetic), and the workgroup's squared errors go to
// part[blockIdx.x]. One launch for any n; a single frame's inference is one workgroup.
constexpr int FR = 32;             // rows a workgroup: two 16-row blocks
constexpr int LACT = MLP_HID + 8;  // LDS activation row: 256 bf16 + 16 B of padding

struct FwdArgs {
    const float *x;
    int64_t ldx;
    int32_t n;
    const uint16_t *wt;  // transposed bf16 weights
    const float *w;      // float32 parameters (the biases)
    uint16_t *X;         // n x 160 bf16 inputs, or null
    float *Z[3];         // pre-activations, or null
    uint16_t *H[3];      // activations, or null
    float *y;            // n x 6
    const float *target; // n x 6, or null
    uint16_t *dy;
    float *part;
    float scale;
};

// One layer of the fused forward: K inputs, NB column blocks of 16 a wave. Every weight fragment of the layer
// is requested before the first product, so the layer waits on memory once.
template <int K, int NB>
__device__ inline void fwd_layer(const uint16_t *act_in, const uint16_t *wl, int col0, f32x4 (&acc)[2][4]) {
    const int lane = lane_id(), q = 8 * (lane >> 4), rl = lane & 15;
    constexpr int S = K / GK;
    uint4 wf[S][NB];
    LW_UNROLL for (int k = 0; k < S; ++k) LW_UNROLL for (int c = 0; c < NB; ++c)
        wf[k][c] = *reinterpret_cast<const uint4 *>(&wl[(col0 + 16 * c + rl) * K + k * GK + q]);
    // the fragments are needed here: keeps the compiler from sinking each load to its product, which would
    // wait on memory once per step instead of once per layer
    LW_UNROLL for (int k = 0; k < S; ++k) LW_UNROLL for (int c = 0; c < NB; ++c)
        asm volatile("" ::"v"(wf[k][c].x), "v"(wf[k][c].y), "v"(wf[k][c].z), "v"(wf[k][c].w));
    LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int c = 0; c < 4; ++c) acc[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    LW_UNROLL for (int k = 0; k < S; ++k) {
        bf16x8 af[2];
        LW_UNROLL for (int a = 0; a < 2; ++a)
            af[a] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(&act_in[(16 * a + rl) * LACT + k * GK + q]));
        LW_UNROLL for (int a = 0; a < 2; ++a) LW_UNROLL for (int c = 0; c < NB; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], __builtin_bit_cast(bf16x8, wf[k][c]), acc[a][c], 0, 0, 0);
    }
}

__global__ __launch_bounds__(256) void k_mlp_fwd(FwdArgs a) {
    __shared__ __align__(16) uint16_t act[2][FR * LACT];
    __shared__ float zt[FR * MLP_HID];  // a layer's pre-activations, before the epilogue
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t r0 = blockIdx.x * FR;
    // the input rows: every load issued before any store (the stores to X could alias x for all the compiler
    // knows, which would otherwise order each load after the previous store)
    constexpr int IN_PER = FR * MLP_IN_PAD / 256;
    static_assert(FR * MLP_IN_PAD % 256 == 0, "whole rounds");
    float xv[IN_PER];
    LW_UNROLL for (int u = 0; u < IN_PER; ++u) {
        const int e = tid + 256 * u, r = e / MLP_IN_PAD, c = e % MLP_IN_PAD;
        const int32_t i = r0 + r;
        const bool in = i < a.n && c < MLP_IN;
        const float t = a.x[in ? (int64_t)i * a.ldx + c : 0];  // a row inside the input, without a branch
        xv[u] = in ? t : 0.0f;
    }
#pragma unroll 4
    for (int u = 0; u < IN_PER; ++u) {
        const int e = tid + 256 * u, r = e / MLP_IN_PAD, c = e % MLP_IN_PAD;
        const int32_t i = r0 + r;
        const uint16_t v = to_bf16(xv[u]);
        act[0][r * LACT + c] = v;
        if (a.X && i < a.n) a.X[(int64_t)i * MLP_IN_PAD + c] = v;
    }
    __syncthreads();
    static constexpr int32_t t_off[4] = {T_W1, T_W2, T_W3, T_W4}, b_off[4] = {OFF_B1, OFF_B2, OFF_B3, OFF_B4};
    f32x4 acc[2][4];
    // The products of a layer go to LDS first and the epilogue runs as a loop over the tile (row-major, so the
    // stores of Z and H are coalesced): the whole kernel's code stays small, and a small kernel's time is
    // mostly its instruction fetches.
    for (int l = 0; l < 3; ++l) {
        const uint16_t *in = act[l & 1];
        uint16_t *out = act[(l + 1) & 1];
        const int col0 = wave * 64;
        if (l == 0)
            fwd_layer<MLP_IN_PAD, 4>(in, a.wt + t_off[0], col0, acc);
        else
            fwd_layer<MLP_HID, 4>(in, a.wt + t_off[l], col0, acc);
        LW_UNROLL for (int b = 0; b < 2; ++b) LW_UNROLL for (int c = 0; c < 4; ++c) LW_UNROLL for (int r = 0; r < 4; ++r)
            zt[(16 * b + 4 * (lane >> 4) + r) * MLP_HID + col0 + 16 * c + (lane & 15)] = acc[b][c][r];
        __syncthreads();
        const float *bias = a.w + b_off[l];
        float *Z = a.Z[l];
        uint16_t *H = a.H[l];
#pragma unroll 2
        for (int e = tid; e < FR * MLP_HID; e += 256) {
            const int row = e / MLP_HID, j = e % MLP_HID;
            const int32_t i = r0 + row;
            const float z = zt[e] + bias[j];
            const uint16_t h = to_bf16(gelu(z));
            out[row * LACT + j] = h;
            if (i < a.n) {
                if (Z) Z[(int64_t)i * MLP_HID + j] = z;
                if (H) H[(int64_t)i * MLP_HID + j] = h;
            }
        }
        __syncthreads();
    }
    if (wave != 0) return;  // the output layer's 16 columns are wave 0's
    fwd_layer<MLP_HID, 1>(act[1], a.wt + t_off[3], 0, acc);
    const float *bias = a.w + b_off[3];
    float se = 0.0f;
    LW_UNROLL for (int b = 0; b < 2; ++b) LW_UNROLL for (int r = 0; r < 4; ++r) {
        const int row = 16 * b + 4 * (lane >> 4) + r, j = lane & 15;
        const int32_t i = r0 + row;
        if (i >= a.n || j >= MLP_OUT) continue;
        const float y = acc[b][0][r] + bias[j];
        a.y[(int64_t)i * MLP_OUT + j] = y;
        if (a.target) {
            const float d = y - a.target[(int64_t)i * MLP_OUT + j];
            se += d * d;
            a.dy[(int64_t)i * MLP_OUT_PAD + j] = to_bf16(a.scale * d);
        }
    }
    if (a.target) {  // the workgroup's squared errors (wave 0 holds them all)
        LW_UNROLL for (int off = 32; off > 0; off >>= 1) se += __shfl_xor(se, off, 64);
        if (lane == 0) a.part[blockIdx.x] = se;
    }
}

// loss = mean over rows and the 6 outputs of (y - t)^2; dY (bf16, 16 wide) = d loss / d y, zero beyond column 6.
// One workgroup: the sum runs in a fixed order.
__global__ __launch_bounds__(1024) void k_mse(const float *y, const float *t, int32_t n, uint16_t *dY, float *loss) {
    __shared__ float part[1024];
    const float scale = 2.0f / (float)(6 * (int64_t)n);
    float s = 0.0f;
    for (int32_t i = threadIdx.x; i < n; i += 1024) {
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
    if (threadIdx.x < 64) {  // 16 partial sums a lane, then the wave's lanes in order
        float q = 0.0f;
        for (int k = 0; k < 16; ++k) q += part[threadIdx.x * 16 + k];
        part[threadIdx.x * 16] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float tot = 0.0f;
        for (int k = 0; k < 64; ++k) tot += part[k * 16];
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
    put_bf16(wb, p, to_bf16(wp));
}

// The backward's last kernel: each layer's weight-and-bias gradient block (W_l then b_l, contiguous in the
// parameter vector) as the sum of its splits' partial tiles in split order; with adam, the Adam step on it
// (k_adam's arithmetic); with loss, the training loss from the row tiles' squared errors, in tile order.
struct GradSum {
    const float *ws[4];
    int32_t splits[4];
    int32_t off[4];  // first parameter of the block
    int32_t len[4];  // its (rows + 1) x columns
    float *g;
    int32_t P, adam;
    float *w, *m, *v;
    uint16_t *wb;
    float lr, b1, b2, eps, c1, c2;
    const float *part;
    int32_t nparts;
    float inv_count;
    float *loss;
};

__global__ __launch_bounds__(256) void k_grad_sum(GradSum a) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && a.loss) {
        float s = 0.0f;
        for (int32_t k = 0; k < a.nparts; ++k) s += a.part[k];
        *a.loss = s * a.inv_count;
    }
    if (p >= a.P) return;
    int l = 3;
    while (l > 0 && p < a.off[l]) --l;
    const int64_t e = p - a.off[l], len = a.len[l];
    const float *ws = a.ws[l];
    const int32_t splits = a.splits[l];
    // every load (the splits' values, the moments, the weight) issued before any store
    float v[SPLIT_MAX];
    LW_UNROLL for (int z = 0; z < SPLIT_MAX; ++z) v[z] = z < splits ? ws[z * len + e] : 0.0f;
    float m0 = 0.0f, v0 = 0.0f, w0 = 0.0f;
    if (a.adam) {
        m0 = a.m[p];
        v0 = a.v[p];
        w0 = a.w[p];
    }
    float gp = 0.0f;  // the splits in split order
    LW_UNROLL for (int z = 0; z < SPLIT_MAX; ++z)
        if (z < splits) gp += v[z];
    a.g[p] = gp;
    if (a.adam) {
        const float mp = a.b1 * m0 + (1.0f - a.b1) * gp;
        const float vp = a.b2 * v0 + (1.0f - a.b2) * gp * gp;
        a.m[p] = mp;
        a.v[p] = vp;
        const float wp = w0 - a.lr * (mp / a.c1) / (sqrtf(vp / a.c2) + a.eps);
        a.w[p] = wp;
        put_bf16(a.wb, p, to_bf16(wp));
    }
}

__global__ void k_params_bf16(const float *w, uint16_t *wb, int32_t P) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P) put_bf16(wb, p, to_bf16(w[p]));
}

// The 128 histogram inputs of one frame (points in the sensor frame): 64 range bins over [0, r_max) and 64
// elevation bins over [-e_max, e_max), each as count * 64 / m (a flat histogram reads 1); out-of-range values
// go to the end bins. The 24 motion inputs come with the arguments.
__global__ __launch_bounds__(1024) void k_features(const float *xyz, int32_t m, const int32_t *m_dev, int32_t cap,
                                                   FeatArgs f, float *out) {
    __shared__ int32_t hr[64], he[64];
    if (m_dev) m = min(*m_dev, cap);
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

constexpr int64_t SPLIT_WS = (int64_t)SPLIT_MAX * (MLP_HID + 1) * MLP_HID;  // floats of the partial tiles

// An operand laid out as its fetch mode needs (16-byte aligned base; see fetch_frag).
template <int MODE>
bool layout_ok(const uint16_t *p, int64_t s_row, int64_t s_k, int32_t K) {
    if (((uintptr_t)p & 15) != 0) return false;
    if (MODE == ALONG) return s_k == 1 && (s_row & 7) == 0 && (K & 7) == 0;
    return s_row == 1 && (s_k & 7) == 0;
}

// The launch shape of a product (false if an operand's layout does not suit its fetch mode): for EPI_STORE
// with a workspace (room for SPLIT_WS floats), K split so that the grid holds >= 256 tiles of at most one
// stage each. A split product writes its partial tiles to ws (always, with force_ws), otherwise C.
template <int EPI, int AM, int BM>
bool plan(Gemm &g, float *ws, bool force_ws = false) {
    if (g.a_ones <= 0) g.a_ones = INT_MAX;
    if (!layout_ok<AM>(g.A, g.sa0, g.sa1, g.K) || !layout_ok<BM>(g.B, g.sb1, g.sb0, g.K)) return false;
    if (AM == ALONG && g.a_ones != INT_MAX) return false;
    g.gx = (g.N + GT - 1) / GT;
    g.gy = (g.M + GT - 1) / GT;
    const int32_t steps = (g.K + GK - 1) / GK;
    int32_t splits = 1;
    if (EPI == EPI_STORE && ws && (int64_t)g.M * g.N * SPLIT_MAX <= SPLIT_WS)
        splits = std::min(std::min(steps, SPLIT_MAX),
                          std::max((256 + g.gx * g.gy - 1) / (g.gx * g.gy), (g.K + GS - 1) / GS));
    const int32_t per = (steps + splits - 1) / splits;  // fragments of K a split
    g.kc = per * GK;
    g.gz = (steps + per - 1) / per;
    g.ws = (EPI == EPI_STORE && ws && (g.gz > 1 || force_ws)) ? ws : nullptr;
    return true;
}

// two planned products in one launch
template <int E1, int A1, int B1, int E2, int A2, int B2>
hipError_t gemm_pair(hipStream_t s, const Gemm &g1, const Gemm &g2) {
    const unsigned blocks = (unsigned)(g1.gx * g1.gy * g1.gz + g2.gx * g2.gy * g2.gz);
    hipLaunchKernelGGL((k_gemm_pair<E1, A1, B1, E2, A2, B2>), dim3(blocks), dim3(256), 0, s, g1, g2);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_features(hipStream_t s, const float *xyz, int32_t m, const int32_t *m_dev, int32_t cap,
                           const FeatArgs &f, float *out) {
    hipLaunchKernelGGL(k_features, dim3(1), dim3(1024), 0, s, xyz, m, m_dev, cap, f, out);
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
    float *Z[3] = {nullptr, nullptr, nullptr};
    float *ws = nullptr;  // split-K partial tiles of the weight gradients
    float *ystep = nullptr;  // the training step's outputs (n x 6)
    float *part = nullptr;   // the training step's squared errors per row tile
    // the captured training step (lw_mlp_train_step): replayed while its pointers and row count stay the same,
    // with the Adam node's scalars updated before each replay
    hipGraph_t step_graph = nullptr;
    hipGraphExec_t step_exec = nullptr;
    hipGraphNode_t adam_node = nullptr;
    hipKernelNodeParams adam_params{};
    lw::GradSum step_sum{};  // the captured step's gradient-sum arguments
    const void *step_key[4] = {nullptr, nullptr, nullptr, nullptr};
    int64_t step_ldx = 0;
    int32_t step_n = 0;
    int32_t rows = 0;  // rows of the last forward
};

namespace {



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
                    p->Z[0], p->Z[1], p->Z[2], p->ws, p->ystep, p->part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
}

template <class T>
hipError_t alloc(T **p, size_t n) {
    return hipMalloc(reinterpret_cast<void **>(p), n * sizeof(T));
}

// The fused forward over n rows (k_mlp_fwd), keeping X, Z and H for the backward; with a target, also the
// loss terms, d loss / d y and the workgroups' squared errors.
hipError_t forward_launch(lw_mlp *p, const float *x, int64_t ldx, int32_t n, float *y, const float *target) {
    FwdArgs a{};
    a.x = x;
    a.ldx = ldx;
    a.n = n;
    a.wt = p->wb + PARAMS;
    a.w = p->w;
    a.X = p->X;
    for (int l = 0; l < 3; ++l) {
        a.Z[l] = p->Z[l];
        a.H[l] = p->H[l];
    }
    a.y = y;
    a.target = target;
    a.dy = p->dY;
    a.part = p->part;
    a.scale = 2.0f / (float)(6 * (int64_t)n);
    hipLaunchKernelGGL(k_mlp_fwd, dim3((unsigned)((n + FR - 1) / FR)), dim3(256), 0, p->stream, a);
    p->rows = n;
    return hipGetLastError();
}

// The backward through the last forward's activations (n rows), every launch but the last: at each layer l
// from the top, one launch holds [dW_l; db_l] = [in, 1]^T dout_l -- (k_in + 1) x cols, K = n split into the
// layer's workspace; the bias gradient is the product's last row (a row of ones in A), and b_l follows W_l in
// the parameter vector -- beside dZ_{l-1} = (dout_l W_l^T) * gelu'(Z_{l-1}), which only needs the same dout_l;
// layer 0's weight gradient goes alone. gs receives the workspaces for k_grad_sum.
hipError_t backward_launch(lw_mlp *p, int32_t n, GradSum &gs) {
    static_assert(OFF_B1 == OFF_W1 + MLP_IN_PAD * MLP_HID && OFF_B2 == OFF_W2 + MLP_HID * MLP_HID &&
                      OFF_B3 == OFF_W3 + MLP_HID * MLP_HID && OFF_B4 == OFF_W4 + MLP_HID * MLP_OUT_PAD,
                  "each bias follows its weights");
    static const int32_t w_off[4] = {OFF_W1, OFF_W2, OFF_W3, OFF_W4};
    Gemm gw[4], gd[4];
    for (int l = 3; l >= 0; --l) {
        const uint16_t *in = l == 0 ? p->X : p->H[l - 1];
        const int32_t kin = l == 0 ? MLP_IN_PAD : MLP_HID;
        const int32_t wcols = l == 3 ? MLP_OUT_PAD : MLP_HID;
        const uint16_t *dout = l == 3 ? p->dY : p->dZ[(l + 1) & 1];  // dZ_l was written at layer l + 1
        const int32_t dstride = wcols, dcols = wcols;
        Gemm g{};
        g.A = in;
        g.sa0 = 1;
        g.sa1 = kin;
        g.a_ones = kin;
        g.B = dout;
        g.sb0 = dstride;
        g.sb1 = 1;
        g.M = kin + 1;
        g.N = dcols;
        g.K = n;
        g.cf = p->g + w_off[l];
        g.ldc = wcols;
        float *ws = p->ws + (int64_t)l * SPLIT_WS;
        if (!plan<EPI_STORE, ACROSS, ACROSS>(g, ws, true)) return hipErrorInvalidValue;
        gw[l] = g;
        gs.ws[l] = ws;
        gs.splits[l] = gw[l].gz;
        gs.off[l] = w_off[l];
        gs.len[l] = (kin + 1) * wcols;
        if (l == 0) break;
        Gemm d{};
        d.A = dout;
        d.sa0 = dstride;
        d.sa1 = 1;
        d.B = p->wb + w_off[l];
        d.sb0 = 1;
        d.sb1 = wcols;
        d.M = n;
        d.N = MLP_HID;
        d.K = dcols;
        d.zin = p->Z[l - 1];
        d.ldz = MLP_HID;
        d.cb = p->dZ[l & 1];
        d.ldh = MLP_HID;
        if (!plan<EPI_GELU_BWD, ALONG, ALONG>(d, nullptr)) return hipErrorInvalidValue;
        gd[l] = d;
    }
    for (int l = 3; l >= 1; --l) {
        hipError_t e = gemm_pair<EPI_STORE, ACROSS, ACROSS, EPI_GELU_BWD, ALONG, ALONG>(p->stream, gw[l], gd[l]);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_gemm<EPI_STORE, ACROSS, ACROSS>), dim3((unsigned)gw[0].gx, (unsigned)gw[0].gy, (unsigned)gw[0].gz), dim3(256),
                       0, p->stream, gw[0]);
    gs.g = p->g;
    gs.P = PARAMS;
    return hipGetLastError();
}

// A training step's launches but the last: the forward with the loss and d loss / d y in the output layer's
// epilogue, then the backward (backward_launch); gs also receives the loss reduction's arguments.
hipError_t train_launch(lw_mlp *p, const float *x, int64_t ldx, const float *target, int32_t n, float *loss,
                        GradSum &gs) {
    hipError_t er = forward_launch(p, x, ldx, n, p->ystep, target);
    if (er != hipSuccess) return er;
    er = backward_launch(p, n, gs);
    gs.part = p->part;
    gs.nparts = (n + FR - 1) / FR;
    gs.inv_count = 1.0f / (float)(6 * (int64_t)n);
    gs.loss = loss;
    return er;
}

int grad_sum_launch(lw_mlp *p, const GradSum &gs) {
    hipLaunchKernelGGL(k_grad_sum, dim3((PARAMS + 255) / 256), dim3(256), 0, p->stream, gs);
    return -(int)hipGetLastError();
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
    if (e == hipSuccess) e = alloc(&p->wb, PARAMS + T_END);  // the parameter vector's bf16 copy, then the transposed weights
    if (e == hipSuccess) e = alloc(&p->X, n * MLP_IN_PAD);
    for (int l = 0; l < 3 && e == hipSuccess; ++l) e = alloc(&p->H[l], n * MLP_HID);
    for (int l = 0; l < 3 && e == hipSuccess; ++l) e = alloc(&p->Z[l], n * MLP_HID);
    for (int l = 0; l < 2 && e == hipSuccess; ++l) e = alloc(&p->dZ[l], n * MLP_HID);
    if (e == hipSuccess) e = alloc(&p->dY, n * MLP_OUT_PAD);
    if (e == hipSuccess) e = alloc(&p->ws, (size_t)(4 * SPLIT_WS));  // one per layer
    if (e == hipSuccess) e = alloc(&p->part, n / FR + 1);
    // dY's columns 6..15 stay zero (the training step's loss epilogue writes columns 0..5 only)
    if (e == hipSuccess) e = hipMemsetAsync(p->dY, 0, n * MLP_OUT_PAD * sizeof(uint16_t), p->stream);
    if (e == hipSuccess) e = alloc(&p->ystep, n * MLP_OUT);
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
    MLP_CK(forward_launch(p, x, ldx, n, y, nullptr));
    return 0;
}

int lw_mlp_loss(lw_mlp *p, const float *y, const float *target, int32_t n, float *loss) {
    if (!p || !y || !target || !loss || n <= 0 || n > p->cap) return LW_EINVAL;
    hipLaunchKernelGGL(k_mse, dim3(1), dim3(1024), 0, p->stream, y, target, n, p->dY, loss);
    MLP_CK(hipGetLastError());
    return 0;
}

int lw_mlp_backward(lw_mlp *p, int32_t n) {
    if (!p || n <= 0 || n != p->rows) return LW_EINVAL;
    GradSum gs{};
    MLP_CK(backward_launch(p, n, gs));
    return grad_sum_launch(p, gs);
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
    // this step's Adam scalars (a captured step's last node gets them before each replay)
    const float c1 = 1.0f - (float)std::pow((double)beta1, (double)step);
    const float c2 = 1.0f - (float)std::pow((double)beta2, (double)step);
    auto set_adam = [&](GradSum &gs) {
        gs.adam = 1;
        gs.w = p->w;
        gs.m = p->m;
        gs.v = p->v;
        gs.wb = p->wb;
        gs.lr = lr;
        gs.b1 = beta1;
        gs.b2 = beta2;
        gs.eps = eps;
        gs.c1 = c1;
        gs.c2 = c2;
    };
    if (!use_graph || p->stream == nullptr) {  // a captured graph needs a stream other than the null stream
        GradSum gs{};
        MLP_CK(train_launch(p, x, ldx, target, n, loss, gs));
        set_adam(gs);
        return grad_sum_launch(p, gs);
    }
    const void *key[4] = {x, target, loss, nullptr};
    if (!p->step_exec || std::memcmp(key, p->step_key, sizeof(key)) != 0 || p->step_ldx != ldx || p->step_n != n) {
        drop_step_graph(p);
        MLP_CK(hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal));
        GradSum gs{};
        hipError_t e = train_launch(p, x, ldx, target, n, loss, gs);
        set_adam(gs);
        int rc = e == hipSuccess ? grad_sum_launch(p, gs) : -(int)e;
        hipGraph_t g = nullptr;
        const hipError_t e2 = hipStreamEndCapture(p->stream, &g);
        if (rc || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return rc ? rc : -(int)e2;
        }
        p->step_graph = g;
        p->step_sum = gs;
        // the gradient-sum node: the one kernel node whose arguments change from step to step
        size_t count = 0;
        MLP_CK(hipGraphGetNodes(g, nullptr, &count));
        hipGraphNode_t nodes[64];
        if (count > 64) {
            drop_step_graph(p);
            return LW_EINVAL;
        }
        MLP_CK(hipGraphGetNodes(g, nodes, &count));
        for (size_t k = 0; k < count; ++k) {
            hipGraphNodeType t;
            MLP_CK(hipGraphNodeGetType(nodes[k], &t));
            if (t != hipGraphNodeTypeKernel) continue;
            hipKernelNodeParams kp{};
            MLP_CK(hipGraphKernelNodeGetParams(nodes[k], &kp));
            if (kp.func == reinterpret_cast<const void *>(&k_grad_sum)) {
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
    GradSum gs = p->step_sum;
    set_adam(gs);
    void *args[] = {&gs};
    hipKernelNodeParams kp = p->adam_params;
    kp.kernelParams = args;
    kp.extra = nullptr;
    MLP_CK(hipGraphExecKernelNodeSetParams(p->step_exec, p->adam_node, &kp));
    MLP_CK(hipGraphLaunch(p->step_exec, p->stream));
    return 0;
}

}  // extern "C"
