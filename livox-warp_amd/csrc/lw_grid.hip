// The prior map's dense grid and the coarse scoring of global localisation (the reference's
// python/livox_warp/mapgrid.py:19-193 and localize.py:55-123).
//
// The grid holds, per 5 cm cell, the map point nearest the cell centre (seeded by the points in the cell,
// then jump-flooded over a band of about 0.8 m) and that distance in cm as a byte: the distance field the
// global search scores (position, heading) hypotheses with. Cells are x-fastest: c = (iz ny + iy) nx + ix.
//
// MI355X layout choices:
//  - the jump flood reads every seed of a step before it writes any (two seed arrays, swapped each step),
//    so a build is deterministic; the reference updates in place, and this order is one its launch admits.
//    A seed travels as a 16-byte record (index and coordinates), so a step's 27 neighbour reads are
//    coherent loads of neighbouring records, not gathers of random map points;
//  - k_score_grid gives the 64 lanes of a wave 64 consecutive positions and one heading at a time: the
//    caller lists positions x-fastest, so for every scan point the wave's 64 distance-field reads fall in a
//    few consecutive cache lines (0.25 m apart = 5 bytes). Each thread carries four positions and four
//    headings, so sixteen independent reads are in flight per point and a heading's products serve all
//    four positions; the scan's points are staged in LDS and read by broadcast; the loop has no branch;
//  - each hypothesis's sum runs over the points in order in one thread, float32 as the reference's, so the
//    scores are exact, not merely close.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <utility>

#include "../../include/livox_warp_amd.h"
#include "lw_common.h"

namespace lw {

namespace {

constexpr int GRID_IDX_BITS = 23;            // mapgrid.py:20
constexpr int32_t GRID_EMPTY = 0x7FFFFFFF;   // mapgrid.py:24
constexpr int SCORE_WAVES = 4;               // waves per workgroup, one heading group each
constexpr int SCORE_YPT = 4;                 // headings per thread
constexpr int SCORE_PPT = 4;                 // positions per thread
constexpr int SCORE_CHUNK = 1024;            // scan points staged in LDS per round

struct GridGeo {
    float ox, oy, oz, cell, inv;
    int32_t nx, ny, nz;
};

struct Mat34 {
    float m[12];
};

// One cell's seed during the flood: the map point's index (-1: none) with its coordinates beside it, so a
// step reads its 27 neighbours' seeds as coherent 16-byte records instead of gathering their points.
struct __align__(16) Seed {
    int32_t idx;
    float x, y, z;
};

__device__ inline int32_t grid_flat(const GridGeo &g, int32_t ix, int32_t iy, int32_t iz) {
    if (ix < 0 || iy < 0 || iz < 0 || ix >= g.nx || iy >= g.ny || iz >= g.nz) return -1;
    return (iz * g.ny + iy) * g.nx + ix;
}

__device__ inline int32_t grid_at(const GridGeo &g, float x, float y, float z) {
    return grid_flat(g, ifloor((x - g.ox) * g.inv), ifloor((y - g.oy) * g.inv), ifloor((z - g.oz) * g.inv));
}

__device__ inline void cell_ijk(const GridGeo &g, int32_t c, int32_t &ix, int32_t &iy, int32_t &iz) {
    ix = c % g.nx;
    const int32_t r = c / g.nx;
    iy = r % g.ny;
    iz = r / g.ny;
}

__device__ inline void cell_centre(const GridGeo &g, int32_t ix, int32_t iy, int32_t iz, float &x, float &y,
                                   float &z) {
    x = g.ox + ((float)ix + 0.5f) * g.cell;
    y = g.oy + ((float)iy + 0.5f) * g.cell;
    z = g.oz + ((float)iz + 0.5f) * g.cell;
}

__device__ inline float dist2_to(float x, float y, float z, const float *q) {
    const float d0 = x - q[0], d1 = y - q[1], d2 = z - q[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

__global__ __launch_bounds__(256) void k_grid_fill(int32_t *a, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        a[i] = v;
}

// mapgrid.py:103-112: each point claims its cell; the smallest (mm to the centre, index) key wins.
__global__ __launch_bounds__(256) void k_grid_seed(const float *xyz, int32_t n, GridGeo g, int32_t *nn) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = xyz + 3 * (int64_t)i;
    const int32_t c = grid_at(g, p[0], p[1], p[2]);
    if (c < 0) return;
    int32_t ix, iy, iz;
    float cx, cy, cz;
    cell_ijk(g, c, ix, iy, iz);
    cell_centre(g, ix, iy, iz, cx, cy, cz);
    const float cc[3] = {cx, cy, cz};
    const int32_t dmm = min((int32_t)(sqrtf(dist2_to(p[0], p[1], p[2], cc)) * 1000.0f), 255);
    atomicMin(nn + c, (dmm << GRID_IDX_BITS) | i);
}

// mapgrid.py:115-123: the winning key's point index (-1 for an unclaimed cell), as a seed record.
__global__ __launch_bounds__(256) void k_grid_unpack(const int32_t *__restrict__ keys, const float *__restrict__ xyz,
                                                     int64_t n, Seed *__restrict__ rec) {
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = keys[c];
        Seed r{-1, 0.0f, 0.0f, 0.0f};
        if (v != GRID_EMPTY) {
            r.idx = v & ((1 << GRID_IDX_BITS) - 1);
            const float *q = xyz + 3 * (int64_t)r.idx;
            r.x = q[0];
            r.y = q[1];
            r.z = q[2];
        }
        rec[c] = r;
    }
}

__device__ inline float dist2_xyz(float x, float y, float z, const Seed &q) {
    const float d0 = x - q.x, d1 = y - q.y, d2 = z - q.z;
    return d0 * d0 + d1 * d1 + d2 * d2;
}

// mapgrid.py:126-155: the nearest of the cell's own seed and the 26 seeds `step` cells away (src -> dst).
// Launched over (x tiles of 256, rows, slabs): the cell's coordinates come from the launch, not divisions.
__global__ __launch_bounds__(256) void k_grid_jfa(const Seed *__restrict__ src, Seed *__restrict__ dst, GridGeo g,
                                                  int32_t step) {
    const int32_t ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= g.nx) return;
    const int32_t c = (iz * g.ny + iy) * g.nx + ix;
    float px, py, pz;
    cell_centre(g, ix, iy, iz, px, py, pz);
    Seed best = src[c];
    float bd = 1.0e30f;
    if (best.idx >= 0) bd = dist2_xyz(px, py, pz, best);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int32_t j = grid_flat(g, ix + dx * step, iy + dy * step, iz + dz * step);
                if (j < 0 || j == c) continue;
                const Seed s = src[j];
                if (s.idx >= 0 && s.idx != best.idx) {
                    const float d2 = dist2_xyz(px, py, pz, s);
                    if (d2 < bd) {
                        bd = d2;
                        best = s;
                    }
                }
            }
    dst[c] = best;
}

// The same step through LDS, for steps up to JFA_MAX_STEP cells. A step of s only ever combines cells whose
// coordinates agree mod s, so a workgroup takes JFA_TX consecutive cells of JFA_TY rows y0, y0 + s, ...,
// y0 + (JFA_TY - 1) s, and walks the planes z = rz, rz + s, ... of its residue rz. It keeps three planes of
// that lattice in LDS -- z - s, z and z + s, each with a halo of s cells in x and one lattice row in y
// (outside the grid: idx -1, which the step skips as it skips an empty seed) -- and loads each plane once, the
// next one while it computes the current: a seed record is read (JFA_TX + 2 s)(JFA_TY + 2) / (JFA_TX JFA_TY)
// times, 1.3-1.9x for steps 1-16, instead of 27 times. Same neighbour order and ties as k_grid_jfa, so the
// same result.
constexpr int JFA_TX = 64, JFA_TY = 8, JFA_MAX_STEP = 16, JFA_LX = JFA_TX + 2 * JFA_MAX_STEP, JFA_LY = JFA_TY + 2;
constexpr int JFA_PREFETCH = (JFA_LX * JFA_LY + 255) / 256;

// Rows of the lattice of residue ry (mod s) and the workgroups per residue that cover them.
inline int32_t jfa_y_tiles(int32_t ny, int32_t s) { return ((ny + s - 1) / s + JFA_TY - 1) / JFA_TY; }

// Plane zz of a slab tile's lattice (with its halo) into registers, then into an LDS slot.
__device__ inline void jfa_fetch(const Seed *__restrict__ src, const GridGeo &g, int32_t x0, int32_t y0, int32_t s,
                                 int32_t zz, Seed (&pre)[JFA_PREFETCH]) {
    const int32_t lxs = JFA_TX + 2 * s, n_load = lxs * JFA_LY;
#pragma unroll
    for (int k = 0; k < JFA_PREFETCH; ++k) {
        const int t = threadIdx.x + 256 * k;
        int32_t j = -1;
        if (t < n_load) j = grid_flat(g, x0 - s + t % lxs, y0 + (t / lxs - 1) * s, zz);
        pre[k] = j >= 0 ? src[j] : Seed{-1, 0.0f, 0.0f, 0.0f};
    }
}

__device__ inline void jfa_store(Seed (*plane)[JFA_LX], int32_t s, const Seed (&pre)[JFA_PREFETCH]) {
    const int32_t lxs = JFA_TX + 2 * s, n_load = lxs * JFA_LY;
#pragma unroll
    for (int k = 0; k < JFA_PREFETCH; ++k) {
        const int t = threadIdx.x + 256 * k;
        if (t < n_load) plane[t / lxs][t % lxs] = pre[k];
    }
}

__global__ __launch_bounds__(256) void k_grid_jfa_slab(const Seed *__restrict__ src, Seed *__restrict__ dst,
                                                       GridGeo g, int32_t s, int32_t y_tiles) {
    __shared__ Seed win[3][JFA_LY][JFA_LX];
    const int32_t ry = blockIdx.y / y_tiles, ty = blockIdx.y % y_tiles, rz = blockIdx.z;
    const int32_t x0 = blockIdx.x * JFA_TX, y0 = ry + ty * JFA_TY * s;
    if (y0 >= g.ny || rz >= g.nz) return;  // the whole workgroup: no row of this tile in the grid
    Seed pre[JFA_PREFETCH];
    int lo = 0, mid = 1, hi = 2;  // slots of planes z - s, z, z + s
    jfa_fetch(src, g, x0, y0, s, rz - s, pre);
    jfa_store(win[lo], s, pre);
    jfa_fetch(src, g, x0, y0, s, rz, pre);
    jfa_store(win[mid], s, pre);
    jfa_fetch(src, g, x0, y0, s, rz + s, pre);
    jfa_store(win[hi], s, pre);
    __syncthreads();
    for (int32_t z = rz; z < g.nz; z += s) {
        jfa_fetch(src, g, x0, y0, s, z + 2 * s, pre);  // in flight while this plane is computed
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            const int lx = threadIdx.x % JFA_TX, ly = threadIdx.x / JFA_TX + 4 * half;
            const int32_t ix = x0 + lx, iy = y0 + ly * s;
            if (ix < g.nx && iy < g.ny) {
                float px, py, pz;
                cell_centre(g, ix, iy, z, px, py, pz);
                Seed best = win[mid][ly + 1][lx + s];
                float bd = 1.0e30f;
                if (best.idx >= 0) bd = dist2_xyz(px, py, pz, best);
                for (int dz = -1; dz <= 1; ++dz) {
                    const int slot = dz < 0 ? lo : (dz == 0 ? mid : hi);
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            if (dx == 0 && dy == 0 && dz == 0) continue;
                            const Seed sd = win[slot][ly + 1 + dy][lx + s + dx * s];
                            if (sd.idx >= 0 && sd.idx != best.idx) {
                                const float d2 = dist2_xyz(px, py, pz, sd);
                                if (d2 < bd) {
                                    bd = d2;
                                    best = sd;
                                }
                            }
                        }
                }
                dst[(z * g.ny + iy) * g.nx + ix] = best;
            }
        }
        __syncthreads();  // every lane is done with plane z - s
        jfa_store(win[lo], s, pre);  // it becomes plane z + 2 s
        const int t = lo;
        lo = mid;
        mid = hi;
        hi = t;
        __syncthreads();
    }
}

// mapgrid.py:158-167 with the seed indices written out: nn per cell, and the cell centre's distance to its
// seed in cm (rounded), 255 beyond the band.
__global__ __launch_bounds__(256) void k_grid_df(const Seed *__restrict__ rec, GridGeo g, int32_t *__restrict__ nn,
                                                 uint8_t *__restrict__ df) {
    const int32_t ix = blockIdx.x * blockDim.x + threadIdx.x, iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= g.nx) return;
    const int32_t c = (iz * g.ny + iy) * g.nx + ix;
    const Seed s = rec[c];
    nn[c] = s.idx;
    if (s.idx < 0) {
        df[c] = 255;
        return;
    }
    float cx, cy, cz;
    cell_centre(g, ix, iy, iz, cx, cy, cz);
    const float d = sqrtf(dist2_xyz(cx, cy, cz, s)) * 100.0f;
    df[c] = (uint8_t)min((int32_t)(d + 0.5f), 255);
}

// mapgrid.py:65-100 nearest_point: exact within one cell, else the nearest of the 27 cells' seeds, else -1.
__device__ int32_t nearest_point(const GridGeo &g, const int32_t *start, const int32_t *nn, const float *xyz, float wx,
                                 float wy, float wz) {
    const int32_t ix = ifloor((wx - g.ox) * g.inv), iy = ifloor((wy - g.oy) * g.inv), iz = ifloor((wz - g.oz) * g.inv);
    int32_t best = -1, seed = -1;
    float bd = g.cell * g.cell, sd = 1.0e30f;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int32_t c = grid_flat(g, ix + dx, iy + dy, iz + dz);
                if (c < 0) continue;
                const int32_t k1 = start[c + 1];
                for (int32_t k = start[c]; k < k1; ++k) {
                    const float d2 = dist2_to(wx, wy, wz, xyz + 3 * (int64_t)k);
                    if (d2 < bd) {
                        bd = d2;
                        best = k;
                    }
                }
                if (best < 0) {
                    const int32_t s = nn[c];
                    if (s >= 0) {
                        const float d2 = dist2_to(wx, wy, wz, xyz + 3 * (int64_t)s);
                        if (d2 < sd) {
                            sd = d2;
                            seed = s;
                        }
                    }
                }
            }
    return best >= 0 ? best : seed;
}

// mapgrid.py:170-193 k_changes_T: each point placed by T, its distance to the map, 1e3 beyond the band.
__global__ __launch_bounds__(256) void k_grid_distances(const float *pts, int32_t n, Mat34 T, GridGeo g,
                                                        const int32_t *start, const int32_t *nn, const float *xyz,
                                                        float *out) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = pts + 3 * (int64_t)i;
    float w[3];
    for (int r = 0; r < 3; ++r) w[r] = T.m[4 * r] * p[0] + T.m[4 * r + 1] * p[1] + T.m[4 * r + 2] * p[2] + T.m[4 * r + 3];
    const int32_t s = nearest_point(g, start, nn, xyz, w[0], w[1], w[2]);
    out[i] = s >= 0 ? sqrtf(dist2_to(w[0], w[1], w[2], xyz + 3 * (int64_t)s)) : 1.0e3f;
}

// Stage points [base, base + cnt) into LDS (x, y, z interleaved).
__device__ inline void stage_points(float *sp, const float *pts, int32_t base, int32_t cnt) {
    for (int k = threadIdx.x; k < 3 * cnt; k += blockDim.x) sp[k] = pts[3 * (int64_t)base + k];
}

// The soft count's lookups in three phases, so a thread's reads are all in flight before any is used: the cell
// of (wx, wy) at row iz (a cell outside the grid reads cell 0 and is masked off), the distance-field bytes,
// then the term 1 - d^2 / tau^2 where d < tau, else +0 (adding +0 to a sum that starts at +0 leaves it as it
// was, so each sum is the reference's, without a branch).
// floor(x) as int32 in one instruction (v_cvt_flr_i32_f32: the same value as (int32_t)floorf(x) for finite x in
// the int32 range, which every grid coordinate here is).
__device__ inline int32_t ifloor1(float x) {
    int32_t r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// The cell index in 24-bit multiplies: exact while nx < 2^24 and ny nz < 2^24 (checked by the launcher), and
// ncell < 2^31; a coordinate outside the grid gives a garbage index that the mask turns into 0.
__device__ inline uint32_t soft_cell(const GridGeo &g, float wx, float wy, int32_t iz, bool &ok) {
    const int32_t ix = ifloor1((wx - g.ox) * g.inv), iy = ifloor1((wy - g.oy) * g.inv);
    ok = ((uint32_t)ix < (uint32_t)g.nx) & ((uint32_t)iy < (uint32_t)g.ny) & ((uint32_t)iz < (uint32_t)g.nz);
    const uint32_t row = __umul24((uint32_t)iz, (uint32_t)g.ny) + (uint32_t)iy;
    const uint32_t c = __umul24(row, (uint32_t)g.nx) + (uint32_t)ix;
    return c & (0u - (uint32_t)ok);  // a mask, not a branch
}

__device__ inline float soft_value(uint32_t byte, bool ok, float tau_cm, float inv_t2) {
    const float d = (float)byte;
    const float t = 1.0f - d * d * inv_t2;
    return (ok & (d < tau_cm)) ? t : 0.0f;
}

// localize.py:55-92 k_score_grid. Workgroup (x, y): positions SCORE_PPT 64 x .. (lane l holds positions
// base + l, base + 64 + l, ...: a wave reads 64 consecutive positions at once), headings of groups
// SCORE_WAVES y .. (one group of SCORE_YPT per wave). The heading products of a point are shared by the
// thread's positions. Positions outside the grid or within free_cm of a mapped surface score -1.
__global__ __launch_bounds__(64 * SCORE_WAVES) void k_score_grid(const float *__restrict__ pts, int32_t n_pts,
                                                                 const float *__restrict__ pos, int32_t n_pos,
                                                                 int32_t n_yaw, float yaw0, float dyaw, GridGeo g,
                                                                 const uint8_t *__restrict__ df, float tau_cm,
                                                                 float free_cm, const int32_t *__restrict__ out_row,
                                                                 float *__restrict__ out) {
    __shared__ float sp[3 * SCORE_CHUNK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int32_t iy0 = (blockIdx.y * SCORE_WAVES + wave) * SCORE_YPT;
    int64_t row[SCORE_PPT];
    float cx[SCORE_PPT], cy[SCORE_PPT], cz[SCORE_PPT];
    bool act[SCORE_PPT];
    bool any = false;
#pragma unroll
    for (int q = 0; q < SCORE_PPT; ++q) {
        const int32_t ip = (blockIdx.x * SCORE_PPT + q) * 64 + lane;
        cx[q] = cy[q] = cz[q] = 0.0f;
        act[q] = false;
        row[q] = ip < n_pos ? (out_row ? (int64_t)out_row[ip] : (int64_t)ip) : 0;
        if (ip < n_pos && iy0 < n_yaw) {
            cx[q] = pos[3 * (int64_t)ip];
            cy[q] = pos[3 * (int64_t)ip + 1];
            cz[q] = pos[3 * (int64_t)ip + 2];
            const int32_t cc = grid_at(g, cx[q], cy[q], cz[q]);
            act[q] = cc >= 0 && !((float)df[cc] < free_cm);
            if (!act[q])
                for (int y = 0; y < SCORE_YPT && iy0 + y < n_yaw; ++y) out[row[q] * n_yaw + iy0 + y] = -1.0f;
        }
        any = any || act[q];
    }
    float co[SCORE_YPT], si[SCORE_YPT], s[SCORE_PPT][SCORE_YPT];
#pragma unroll
    for (int y = 0; y < SCORE_YPT; ++y) {
        double sd, cd;
        sincos((double)(yaw0 + (float)(iy0 + y) * dyaw), &sd, &cd);
        co[y] = (float)cd;
        si[y] = (float)sd;
#pragma unroll
        for (int q = 0; q < SCORE_PPT; ++q) s[q][y] = 0.0f;
    }
    const float inv_t2 = 1.0f / (tau_cm * tau_cm);
    for (int32_t base = 0; base < n_pts; base += SCORE_CHUNK) {
        const int32_t cnt = min(SCORE_CHUNK, n_pts - base);
        __syncthreads();
        stage_points(sp, pts, base, cnt);
        __syncthreads();
        if (!any) continue;
        for (int32_t k = 0; k < cnt; ++k) {
            const float px = sp[3 * k], py = sp[3 * k + 1], pz = sp[3 * k + 2];
            int32_t iz[SCORE_PPT];
#pragma unroll
            for (int q = 0; q < SCORE_PPT; ++q) iz[q] = ifloor1((cz[q] + pz - g.oz) * g.inv);
            uint32_t c[SCORE_PPT][SCORE_YPT];
            bool ok[SCORE_PPT][SCORE_YPT];
#pragma unroll
            for (int y = 0; y < SCORE_YPT; ++y) {
                const float a = co[y] * px, b = si[y] * py, e = si[y] * px, f = co[y] * py;
#pragma unroll
                for (int q = 0; q < SCORE_PPT; ++q) c[q][y] = soft_cell(g, cx[q] + a - b, cy[q] + e + f, iz[q], ok[q][y]);
            }
            uint32_t v[SCORE_PPT][SCORE_YPT];
#pragma unroll
            for (int q = 0; q < SCORE_PPT; ++q)
#pragma unroll
                for (int y = 0; y < SCORE_YPT; ++y) v[q][y] = df[c[q][y]];
#pragma unroll
            for (int q = 0; q < SCORE_PPT; ++q)
#pragma unroll
                for (int y = 0; y < SCORE_YPT; ++y) s[q][y] += soft_value(v[q][y], ok[q][y], tau_cm, inv_t2);
        }
    }
#pragma unroll
    for (int q = 0; q < SCORE_PPT; ++q)
        if (act[q])
            for (int y = 0; y < SCORE_YPT && iy0 + y < n_yaw; ++y) out[row[q] * n_yaw + iy0 + y] = s[q][y];
}

// localize.py:95-123 k_score_list: one thread per (x, y, z, yaw) hypothesis.
__global__ __launch_bounds__(256) void k_score_list(const float *__restrict__ pts, int32_t n_pts,
                                                    const float *__restrict__ hyp, int32_t n_hyp, GridGeo g,
                                                    const uint8_t *__restrict__ df, float tau_cm,
                                                    float *__restrict__ out) {
    __shared__ float sp[3 * SCORE_CHUNK];
    const int32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = h < n_hyp;
    float cx = 0.0f, cy = 0.0f, cz = 0.0f, co = 1.0f, si = 0.0f, s = 0.0f;
    if (active) {
        const float4 q = reinterpret_cast<const float4 *>(hyp)[h];
        cx = q.x;
        cy = q.y;
        cz = q.z;
        double sd, cd;
        sincos((double)q.w, &sd, &cd);
        co = (float)cd;
        si = (float)sd;
    }
    const float inv_t2 = 1.0f / (tau_cm * tau_cm);
    for (int32_t base = 0; base < n_pts; base += SCORE_CHUNK) {
        const int32_t cnt = min(SCORE_CHUNK, n_pts - base);
        __syncthreads();
        stage_points(sp, pts, base, cnt);
        __syncthreads();
        if (!active) continue;
        for (int32_t k = 0; k < cnt; ++k) {
            const float px = sp[3 * k], py = sp[3 * k + 1], pz = sp[3 * k + 2];
            bool ok;
            const uint32_t c = soft_cell(g, cx + co * px - si * py, cy + si * px + co * py,
                                         ifloor1((cz + pz - g.oz) * g.inv), ok);
            s += soft_value(df[c], ok, tau_cm, inv_t2);
        }
    }
    if (active) out[h] = s;
}

// ---- point-to-plane ICP over the grid (localize.py:126-210) ----------------------------------------------

constexpr int PLANE_CHUNKS = 64;  // point chunks per hypothesis, one workgroup each
constexpr int PLANE_SUMS = 30;    // H (21), b (6), count, weighted error, weight sum
constexpr float PLANAR = 0.5f;    // mapgrid.py:26

__device__ inline double wave_sum_f64(double v) {
    _Pragma("unroll") for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// localize.py:126-168 k_icp_batch for the points of chunk blockIdx.x of hypothesis blockIdx.y: each term in
// float32 as the reference forms it, widened to float64 and summed per thread, then per wave and per
// workgroup in a fixed order, into the chunk's partial row (PLANE_SUMS doubles, 32 apart).
__global__ __launch_bounds__(256) void k_plane_icp(const float *__restrict__ pts, int32_t n_pts,
                                                   const float *__restrict__ Tk, GridGeo g,
                                                   const int32_t *__restrict__ start, const int32_t *__restrict__ nn,
                                                   const float *__restrict__ xyz, const float *__restrict__ nrm,
                                                   const float *__restrict__ planar, float max_d2, float c2,
                                                   float alpha, double *__restrict__ partials) {
    __shared__ double red[4][PLANE_SUMS];
    const int h = blockIdx.y, chunk = blockIdx.x;
    const float *T = Tk + 16 * h;
    const int32_t lo = (int32_t)((int64_t)n_pts * chunk / PLANE_CHUNKS),
                  hi = (int32_t)((int64_t)n_pts * (chunk + 1) / PLANE_CHUNKS);
    double acc[PLANE_SUMS];
    _Pragma("unroll") for (int k = 0; k < PLANE_SUMS; ++k) acc[k] = 0.0;
    for (int32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const float *p = pts + 3 * (int64_t)i;
        float w[3];
        for (int r = 0; r < 3; ++r) w[r] = T[4 * r] * p[0] + T[4 * r + 1] * p[1] + T[4 * r + 2] * p[2] + T[4 * r + 3];
        const int32_t s = nearest_point(g, start, nn, xyz, w[0], w[1], w[2]);
        if (s < 0) continue;
        const float *q = xyz + 3 * (int64_t)s;
        const float r0 = w[0] - q[0], r1 = w[1] - q[1], r2 = w[2] - q[2];
        if (r0 * r0 + r1 * r1 + r2 * r2 > max_d2) continue;
        const float *n = nrm + 3 * (int64_t)s;
        float M[3][3] = {{alpha, 0.0f, 0.0f}, {0.0f, alpha, 0.0f}, {0.0f, 0.0f, alpha}};
        if (planar[s] > PLANAR && sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) > 0.5f) {
            const float f = 1.0f - alpha;  // the normal at full weight, along the surface only a little
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) M[a][b] = n[a] * n[b] * f + M[a][b];
        } else {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) M[a][b] = M[a][b] * 4.0f;  // a weak point-to-point pull
        }
        const float Mr[3] = {M[0][0] * r0 + M[0][1] * r1 + M[0][2] * r2, M[1][0] * r0 + M[1][1] * r1 + M[1][2] * r2,
                             M[2][0] * r0 + M[2][1] * r1 + M[2][2] * r2};
        const float d2 = r0 * Mr[0] + r1 * Mr[1] + r2 * Mr[2];
        float wgt = 1.0f / (1.0f + d2 / c2);
        wgt = wgt * wgt;  // Geman-McClure
        const float J[3][6] = {{1.0f, 0.0f, 0.0f, 0.0f, w[2], -w[1]},
                               {0.0f, 1.0f, 0.0f, -w[2], 0.0f, w[0]},
                               {0.0f, 0.0f, 1.0f, w[1], -w[0], 0.0f}};
        int idx = 0;
        _Pragma("unroll") for (int a = 0; a < 6; ++a) {
            _Pragma("unroll") for (int b = a; b < 6; ++b) {
                float v = M[0][0] * J[0][a] * J[0][b] + M[0][1] * J[0][a] * J[1][b] + M[0][2] * J[0][a] * J[2][b];
                v += M[1][0] * J[1][a] * J[0][b] + M[1][1] * J[1][a] * J[1][b] + M[1][2] * J[1][a] * J[2][b];
                v += M[2][0] * J[2][a] * J[0][b] + M[2][1] * J[2][a] * J[1][b] + M[2][2] * J[2][a] * J[2][b];
                acc[idx++] += (double)(wgt * v);
            }
        }
        _Pragma("unroll") for (int a = 0; a < 6; ++a)
            acc[21 + a] += (double)(wgt * (Mr[0] * J[0][a] + Mr[1] * J[1][a] + Mr[2] * J[2][a]));
        acc[27] += 1.0;
        acc[28] += (double)(wgt * (r0 * Mr[0] + r1 * Mr[1] + r2 * Mr[2]));
        acc[29] += (double)wgt;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    _Pragma("unroll") for (int k = 0; k < PLANE_SUMS; ++k) {
        const double v = wave_sum_f64(acc[k]);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < PLANE_SUMS) {
        const int k = threadIdx.x;
        partials[((int64_t)h * PLANE_CHUNKS + chunk) * 32 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    }
}

// localize.py:171-180 k_reduce_batch: hypothesis blockIdx.x's chunks summed in chunk order into its
// accumulator row (64 doubles: H 0-20, b 21-26, count 27, weighted error 28, weight sum 30, the rest 0).
__global__ __launch_bounds__(64) void k_plane_reduce(const double *__restrict__ partials, double *__restrict__ out) {
    const int h = blockIdx.x, j = threadIdx.x;
    const int src = j < 29 ? j : (j == 30 ? 29 : -1);  // the weight sum sits at ACC_WEIGHT = 30, ACC_RMS stays 0
    double s = 0.0;
    if (src >= 0)
        for (int c = 0; c < PLANE_CHUNKS; ++c) s += partials[((int64_t)h * PLANE_CHUNKS + c) * 32 + src];
    out[(int64_t)h * 64 + j] = s;
}

// localize.py:183-206 k_eval_batch: per hypothesis, the scan points whose nearest map point is within
// radii[0], radii[1], radii[2] (counts zeroed by the launcher; integer sums, so the counts are exact).
__global__ __launch_bounds__(256) void k_plane_eval(const float *__restrict__ pts, int32_t n_pts,
                                                    const float *__restrict__ Tk, GridGeo g,
                                                    const int32_t *__restrict__ start, const int32_t *__restrict__ nn,
                                                    const float *__restrict__ xyz, float rad0, float rad1, float rad2,
                                                    int32_t *__restrict__ counts) {
    const int h = blockIdx.y, chunk = blockIdx.x;
    const float *T = Tk + 16 * h;
    const int32_t lo = (int32_t)((int64_t)n_pts * chunk / PLANE_CHUNKS),
                  hi = (int32_t)((int64_t)n_pts * (chunk + 1) / PLANE_CHUNKS);
    int32_t c0 = 0, c1 = 0, c2 = 0;
    for (int32_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const float *p = pts + 3 * (int64_t)i;
        float w[3];
        for (int r = 0; r < 3; ++r) w[r] = T[4 * r] * p[0] + T[4 * r + 1] * p[1] + T[4 * r + 2] * p[2] + T[4 * r + 3];
        const int32_t s = nearest_point(g, start, nn, xyz, w[0], w[1], w[2]);
        if (s < 0) continue;
        const float d = sqrtf(dist2_to(w[0], w[1], w[2], xyz + 3 * (int64_t)s));
        c0 += d < rad0;
        c1 += d < rad1;
        c2 += d < rad2;
    }
    _Pragma("unroll") for (int off = 32; off > 0; off >>= 1) {
        c0 += __shfl_xor(c0, off, 64);
        c1 += __shfl_xor(c1, off, 64);
        c2 += __shfl_xor(c2, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (c0) atomicAdd(counts + 3 * h, c0);
        if (c1) atomicAdd(counts + 3 * h + 1, c1);
        if (c2) atomicAdd(counts + 3 * h + 2, c2);
    }
}

bool geo_of(const float *geo, const int32_t *dims, GridGeo &g) {
    if (!geo || !dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return false;
    if ((int64_t)dims[0] * dims[1] * dims[2] > (int64_t)INT32_MAX - 1) return false;
    if (dims[0] >= (1 << 24) || (int64_t)dims[1] * dims[2] >= (1 << 24)) return false;  // the 24-bit index
    if (dims[1] > 65535 || dims[2] > 65535) return false;  // rows and slabs index the launch grid
    if (!(geo[3] > 0.0f) || !(geo[4] > 0.0f)) return false;
    g = GridGeo{geo[0], geo[1], geo[2], geo[3], geo[4], dims[0], dims[1], dims[2]};
    return true;
}

int launched() {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : -(int)e;
}

unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

}  // namespace

}  // namespace lw

using lw::GridGeo;

extern "C" int64_t lw_grid_work_bytes(const int32_t *dims) {
    if (!dims || dims[0] <= 0 || dims[1] <= 0 || dims[2] <= 0) return -1;
    return 2 * (int64_t)sizeof(lw::Seed) * dims[0] * dims[1] * dims[2];
}

extern "C" int lw_grid_build(const float *xyz, int32_t n, const float *geo, const int32_t *dims, const int32_t *steps,
                             int32_t n_steps, int32_t *nn, void *work, int64_t work_bytes, uint8_t *df, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n < 0 || n >= (1 << lw::GRID_IDX_BITS) || n_steps < 0 || n_steps > 32 || !nn ||
        !work || !df || (n > 0 && !xyz) || (n_steps > 0 && !steps) || ((uintptr_t)work & 15))
        return LW_EINVAL;
    const int64_t ncell = (int64_t)g.nx * g.ny * g.nz;
    if (work_bytes < 2 * (int64_t)sizeof(lw::Seed) * ncell) return LW_EINVAL;
    for (int k = 0; k < n_steps; ++k)
        if (steps[k] <= 0) return LW_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    lw::Seed *a = static_cast<lw::Seed *>(work), *b = a + ncell;
    const unsigned fill_blocks = (unsigned)std::min<int64_t>(lw::blocks_for(ncell), 65536);
    // the seeds' keys go through nn, which the last kernel overwrites with the indices
    hipLaunchKernelGGL(lw::k_grid_fill, dim3(fill_blocks), dim3(256), 0, s, nn, ncell, lw::GRID_EMPTY);
    if (n > 0) hipLaunchKernelGGL(lw::k_grid_seed, dim3(lw::blocks_for(n)), dim3(256), 0, s, xyz, n, g, nn);
    hipLaunchKernelGGL(lw::k_grid_unpack, dim3(fill_blocks), dim3(256), 0, s, nn, xyz, ncell, a);
    const dim3 cells(lw::blocks_for(g.nx), (unsigned)g.ny, (unsigned)g.nz);
    for (int k = 0; k < n_steps; ++k) {
        const int32_t st = steps[k];
        if (st <= lw::JFA_MAX_STEP) {
            const int32_t yt = lw::jfa_y_tiles(g.ny, st);
            const dim3 slabs((unsigned)((g.nx + lw::JFA_TX - 1) / lw::JFA_TX), (unsigned)(st * yt), (unsigned)st);
            hipLaunchKernelGGL(lw::k_grid_jfa_slab, slabs, dim3(256), 0, s, a, b, g, st, yt);
        } else {
            hipLaunchKernelGGL(lw::k_grid_jfa, cells, dim3(256), 0, s, a, b, g, st);
        }
        std::swap(a, b);
    }
    hipLaunchKernelGGL(lw::k_grid_df, cells, dim3(256), 0, s, a, g, nn, df);
    return lw::launched();
}

extern "C" int64_t lw_plane_work_bytes(int32_t k) {
    return k < 0 ? -1 : (int64_t)k * lw::PLANE_CHUNKS * 32 * (int64_t)sizeof(double);
}

extern "C" int lw_plane_icp(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo,
                            const int32_t *dims, const int32_t *start, const int32_t *nn, const float *xyz,
                            const float *nrm, const float *planar, float max_d2, float c2, float alpha, void *work,
                            int64_t work_bytes, double *out, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n_pts < 0 || k < 0 || k > 65535 || !(c2 > 0.0f)) return LW_EINVAL;
    if (k == 0) return 0;
    if ((n_pts > 0 && !pts) || !T || !start || !nn || !xyz || !nrm || !planar || !work || !out ||
        work_bytes < lw_plane_work_bytes(k))
        return LW_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    double *partials = static_cast<double *>(work);
    hipLaunchKernelGGL(lw::k_plane_icp, dim3(lw::PLANE_CHUNKS, (unsigned)k), dim3(256), 0, s, pts, n_pts, T, g, start,
                       nn, xyz, nrm, planar, max_d2, c2, alpha, partials);
    hipLaunchKernelGGL(lw::k_plane_reduce, dim3((unsigned)k), dim3(64), 0, s, partials, out);
    return lw::launched();
}

// ---- lw_plane_icp without a workspace argument ---------------------------------------------------------------
// The partial rows live in a file-scope device array, one copy per GPU for the life of the process. Host code
// never passes the array's name to a kernel (on the host it is a stand-in, not a device address): the entry point
// looks up each device's address once, with that device current, and passes that.
static __device__ __align__(256) double scratch[(size_t)LW_PLANE_NOALLOC_MAX * lw::PLANE_CHUNKS * 32];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch[SCRATCH_DEVICES] = {};  // per device ordinal, filled on the first call there

int scratch_of_current_device(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return -(int)e;
    if (dev < 0 || dev >= SCRATCH_DEVICES) return LW_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return e != hipSuccess ? -(int)e : LW_EINVAL;
        g_scratch[dev] = p;
    }
    *out = static_cast<double *>(g_scratch[dev]);
    return 0;
}
}  // namespace

extern "C" int lw_plane_icp_noalloc(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo,
                                    const int32_t *dims, const int32_t *start, const int32_t *nn, const float *xyz,
                                    const float *nrm, const float *planar, float max_d2, float c2, float alpha,
                                    double *out, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n_pts < 0 || k < 0 || !(c2 > 0.0f)) return LW_EINVAL;
    if (k > LW_PLANE_NOALLOC_MAX) return LW_ERANGE;
    if (k == 0) return 0;
    if ((n_pts > 0 && !pts) || !T || !start || !nn || !xyz || !nrm || !planar || !out) return LW_EINVAL;
    double *partials = nullptr;
    if (int rc = scratch_of_current_device(&partials)) return rc;
    // every partial row's 30 sums are written by its chunk's workgroup before the reduction reads them, so the
    // scratch needs no clearing between calls
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lw::k_plane_icp, dim3(lw::PLANE_CHUNKS, (unsigned)k), dim3(256), 0, s, pts, n_pts, T, g, start,
                       nn, xyz, nrm, planar, max_d2, c2, alpha, partials);
    hipLaunchKernelGGL(lw::k_plane_reduce, dim3((unsigned)k), dim3(64), 0, s, partials, out);
    return lw::launched();
}

extern "C" int lw_plane_eval(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo,
                             const int32_t *dims, const int32_t *start, const int32_t *nn, const float *xyz,
                             const float *radii, int32_t *counts, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n_pts < 0 || k < 0 || k > 65535 || !radii) return LW_EINVAL;
    if (k == 0) return 0;
    if ((n_pts > 0 && !pts) || !T || !start || !nn || !xyz || !counts) return LW_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(lw::k_grid_fill, dim3(1), dim3(256), 0, s, counts, (int64_t)3 * k, 0);
    hipLaunchKernelGGL(lw::k_plane_eval, dim3(lw::PLANE_CHUNKS, (unsigned)k), dim3(256), 0, s, pts, n_pts, T, g, start,
                       nn, xyz, radii[0], radii[1], radii[2], counts);
    return lw::launched();
}

extern "C" int lw_grid_distances(const float *pts, int32_t n, const float *T, const float *geo, const int32_t *dims,
                                 const int32_t *start, const int32_t *nn, const float *xyz, float *out, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n < 0 || !T) return LW_EINVAL;
    if (n == 0) return 0;
    if (!pts || !start || !nn || !xyz || !out) return LW_EINVAL;
    lw::Mat34 m;
    for (int k = 0; k < 12; ++k) m.m[k] = T[k];
    hipLaunchKernelGGL(lw::k_grid_distances, dim3(lw::blocks_for(n)), dim3(256), 0, (hipStream_t)stream, pts, n, m, g,
                       start, nn, xyz, out);
    return lw::launched();
}

extern "C" int lw_score_grid(const float *pts, int32_t n_pts, const float *pos, int32_t n_pos, int32_t n_yaw,
                             float yaw0, float dyaw, const float *geo, const int32_t *dims, const uint8_t *df,
                             float tau_cm, float free_cm, const int32_t *out_row, float *out, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n_pts < 0 || n_pos < 0 || n_yaw <= 0 || !(tau_cm > 0.0f)) return LW_EINVAL;
    if (n_pos == 0) return 0;
    if ((n_pts > 0 && !pts) || !pos || !df || !out) return LW_EINVAL;
    const int32_t groups = (n_yaw + lw::SCORE_YPT - 1) / lw::SCORE_YPT;
    const dim3 grid((unsigned)((n_pos + 64 * lw::SCORE_PPT - 1) / (64 * lw::SCORE_PPT)),
                    (unsigned)((groups + lw::SCORE_WAVES - 1) / lw::SCORE_WAVES));
    hipLaunchKernelGGL(lw::k_score_grid, grid, dim3(64 * lw::SCORE_WAVES), 0, (hipStream_t)stream, pts, n_pts, pos,
                       n_pos, n_yaw, yaw0, dyaw, g, df, tau_cm, free_cm, out_row, out);
    return lw::launched();
}

extern "C" int lw_score_list(const float *pts, int32_t n_pts, const float *hyp, int32_t n_hyp, const float *geo,
                             const int32_t *dims, const uint8_t *df, float tau_cm, float *out, void *stream) {
    GridGeo g;
    if (!lw::geo_of(geo, dims, g) || n_pts < 0 || n_hyp < 0 || !(tau_cm > 0.0f)) return LW_EINVAL;
    if (n_hyp == 0) return 0;
    if ((n_pts > 0 && !pts) || !hyp || !df || !out || ((uintptr_t)hyp & 15)) return LW_EINVAL;
    hipLaunchKernelGGL(lw::k_score_list, dim3(lw::blocks_for(n_hyp)), dim3(256), 0, (hipStream_t)stream, pts, n_pts,
                       hyp, n_hyp, g, df, tau_cm, out);
    return lw::launched();
}
