/*
 * CPU restatement of the reference's prior-map grid and its coarse scoring -- TEST INFRASTRUCTURE ONLY.
 *
 * Tests, the smoke check and the CPU baseline leg of the benchmark call it as the checker; the product
 * never links or calls it. Float32 throughout, each multiply and add rounded on its own (built with
 * -ffp-contract=off), in the order the reference's expressions are written, left to right.
 *
 * Reference files (paths relative to the reference repository root):
 *   python/livox_warp/mapgrid.py   :19-26 constants, :38-62 grid_flat, grid_at, cell_centre,
 *                                   :65-100 nearest_point, :103-112 k_seed, :115-123 k_unpack,
 *                                   :126-155 k_jfa, :158-167 k_df, :170-193 k_changes_T, :275-290 build
 *   python/livox_warp/localize.py  :55-92 k_score_grid, :95-123 k_score_list, :126-180 k_icp_batch and
 *                                   k_reduce_batch, :183-206 k_eval_batch
 *   python/livox_warp/odom.py      :234-256 accumulate (the ACC layout, :72-82)
 *
 * Choices where the reference leaves the bits open ("parity unpinned" at the bit level, pinned by the
 * properties the tests check instead):
 *   - k_jfa updates the seeds in place, so a cell may read a neighbour already updated in the same step
 *     and the result depends on the order the device runs threads in. This restatement (and the device)
 *     reads every seed of a step before writing any: one admissible order of the reference's launch.
 *   - wp.cos / wp.sin of the float32 heading: here the float64 cos / sin of that float32 value, rounded
 *     to float32 (within an ulp of any float32 cos), so device and oracle agree bit for bit.
 *   - wp.transform_point: row r is ((m[r][0]*p0 + m[r][1]*p1) + m[r][2]*p2) + m[r][3].
 *
 * Geometry: geo = {origin x, y, z, cell, inv_cell}, dims = {nx, ny, nz}; cell c = (iz*ny + iy)*nx + ix.
 */
#include <math.h>
#include <stdint.h>

#define GRID_IDX_BITS 23
#define GRID_EMPTY 0x7FFFFFFF

static int64_t grid_flat(const int32_t *dims, int64_t ix, int64_t iy, int64_t iz)
{
    if (ix < 0 || iy < 0 || iz < 0 || ix >= dims[0] || iy >= dims[1] || iz >= dims[2]) return -1;
    return (iz * dims[1] + iy) * dims[0] + ix;
}

static int64_t grid_at(const float *geo, const int32_t *dims, const float *w)
{
    return grid_flat(dims, (int64_t)(int32_t)floorf((w[0] - geo[0]) * geo[4]),
                     (int64_t)(int32_t)floorf((w[1] - geo[1]) * geo[4]),
                     (int64_t)(int32_t)floorf((w[2] - geo[2]) * geo[4]));
}

static void cell_ijk(const int32_t *dims, int64_t c, int64_t *ix, int64_t *iy, int64_t *iz)
{
    *ix = c % dims[0];
    int64_t r = c / dims[0];
    *iy = r % dims[1];
    *iz = r / dims[1];
}

static void cell_centre(const float *geo, const int32_t *dims, int64_t c, float *p)
{
    int64_t ix, iy, iz;
    cell_ijk(dims, c, &ix, &iy, &iz);
    p[0] = geo[0] + ((float)ix + 0.5f) * geo[3];
    p[1] = geo[1] + ((float)iy + 0.5f) * geo[3];
    p[2] = geo[2] + ((float)iz + 0.5f) * geo[3];
}

static float dist2(const float *a, const float *b)
{
    float d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + d1 * d1 + d2 * d2;
}

/* mapgrid.py:103-112 k_seed: each point claims its cell, the smallest (mm to the centre, index) key wins. */
void lwo_grid_seed(const float *xyz, int32_t n, const float *geo, const int32_t *dims, int32_t *nn)
{
    for (int32_t i = 0; i < n; ++i) {
        const float *p = xyz + 3 * (int64_t)i;
        int64_t c = grid_at(geo, dims, p);
        if (c < 0) continue;
        float cc[3];
        cell_centre(geo, dims, c, cc);
        int32_t dmm = (int32_t)(sqrtf(dist2(p, cc)) * 1000.0f);
        if (dmm > 255) dmm = 255;
        int32_t key = (dmm << GRID_IDX_BITS) | i;
        if (key < nn[c]) nn[c] = key;
    }
}

/* mapgrid.py:126-155 k_jfa, every seed of the step read before any is written (src -> dst). */
void lwo_grid_jfa(const int32_t *src, int32_t *dst, const float *xyz, const float *geo, const int32_t *dims,
                  int32_t step)
{
    int64_t ncell = (int64_t)dims[0] * dims[1] * dims[2];
    for (int64_t c = 0; c < ncell; ++c) {
        float p[3];
        int64_t ix, iy, iz;
        cell_centre(geo, dims, c, p);
        cell_ijk(dims, c, &ix, &iy, &iz);
        int32_t best = src[c];
        float bd = 1.0e30f;
        if (best >= 0) bd = dist2(p, xyz + 3 * (int64_t)best);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int64_t j = grid_flat(dims, ix + dx * step, iy + dy * step, iz + dz * step);
                    if (j < 0 || j == c) continue;
                    int32_t s = src[j];
                    if (s >= 0 && s != best) {
                        float d2 = dist2(p, xyz + 3 * (int64_t)s);
                        if (d2 < bd) {
                            bd = d2;
                            best = s;
                        }
                    }
                }
        dst[c] = best;
    }
}

/* mapgrid.py:275-290 build: fill, seed, unpack (:115-123), the jump-flood steps, the distance field
 * (:158-167). steps: the step list (8, 4, 2, 1, 1 in the reference); tmp: a second seed array. */
void lwo_grid_build(const float *xyz, int32_t n, const float *geo, const int32_t *dims, const int32_t *steps,
                    int32_t n_steps, int32_t *nn, int32_t *tmp, uint8_t *df)
{
    int64_t ncell = (int64_t)dims[0] * dims[1] * dims[2];
    int32_t *a = (n_steps % 2) ? tmp : nn, *b = (n_steps % 2) ? nn : tmp;
    for (int64_t c = 0; c < ncell; ++c) a[c] = GRID_EMPTY;
    lwo_grid_seed(xyz, n, geo, dims, a);
    for (int64_t c = 0; c < ncell; ++c) a[c] = a[c] == GRID_EMPTY ? -1 : (a[c] & ((1 << GRID_IDX_BITS) - 1));
    for (int32_t k = 0; k < n_steps; ++k) {
        lwo_grid_jfa(a, b, xyz, geo, dims, steps[k]);
        int32_t *t = a;
        a = b;
        b = t;
    }
    for (int64_t c = 0; c < ncell; ++c) {
        int32_t s = nn[c];
        if (s < 0) {
            df[c] = 255;
            continue;
        }
        float cc[3];
        cell_centre(geo, dims, c, cc);
        const float *q = xyz + 3 * (int64_t)s;
        float d0 = cc[0] - q[0], d1 = cc[1] - q[1], d2 = cc[2] - q[2];
        float d = sqrtf(d0 * d0 + d1 * d1 + d2 * d2) * 100.0f;
        int32_t v = (int32_t)(d + 0.5f);
        df[c] = (uint8_t)(v < 255 ? v : 255);
    }
}

/* mapgrid.py:65-100 nearest_point: the exact nearest point within one cell of w, else the nearest of the
 * 27 cells' seeds, else -1. */
int32_t lwo_grid_nearest(const float *w, const float *geo, const int32_t *dims, const int32_t *start,
                         const int32_t *nn, const float *xyz)
{
    int64_t ix = (int32_t)floorf((w[0] - geo[0]) * geo[4]);
    int64_t iy = (int32_t)floorf((w[1] - geo[1]) * geo[4]);
    int64_t iz = (int32_t)floorf((w[2] - geo[2]) * geo[4]);
    int32_t best = -1, seed = -1;
    float bd = geo[3] * geo[3], sd = 1.0e30f;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int64_t c = grid_flat(dims, ix + dx, iy + dy, iz + dz);
                if (c < 0) continue;
                for (int32_t k = start[c]; k < start[c + 1]; ++k) {
                    float d2 = dist2(w, xyz + 3 * (int64_t)k);
                    if (d2 < bd) {
                        bd = d2;
                        best = k;
                    }
                }
                if (best < 0) {
                    int32_t s = nn[c];
                    if (s >= 0) {
                        float d2 = dist2(w, xyz + 3 * (int64_t)s);
                        if (d2 < sd) {
                            sd = d2;
                            seed = s;
                        }
                    }
                }
            }
    return best >= 0 ? best : seed;
}

/* mapgrid.py:170-193 k_changes_T: each point placed by T (row-major 4x4), its distance to the nearest map
 * point, 1e3 when there is none within the band. */
void lwo_grid_distances(const float *pts, int32_t n, const float *T, const float *geo, const int32_t *dims,
                        const int32_t *start, const int32_t *nn, const float *xyz, float *out)
{
    for (int32_t i = 0; i < n; ++i) {
        const float *p = pts + 3 * (int64_t)i;
        float w[3];
        for (int r = 0; r < 3; ++r) w[r] = T[4 * r] * p[0] + T[4 * r + 1] * p[1] + T[4 * r + 2] * p[2] + T[4 * r + 3];
        int32_t s = lwo_grid_nearest(w, geo, dims, start, nn, xyz);
        out[i] = s >= 0 ? sqrtf(dist2(w, xyz + 3 * (int64_t)s)) : 1.0e3f;
    }
}

/* The soft count of one hypothesis (localize.py:79-92, 112-123): points rotated by the heading (cy, sy)
 * about z and moved to c, each within tau cm of the map (by the distance field) adding 1 - d^2 / tau^2. */
static float soft_count(const float *pts, int32_t n_pts, const float *c, float cy, float sy, const float *geo,
                        const int32_t *dims, const uint8_t *df, float tau_cm)
{
    float inv_t2 = 1.0f / (tau_cm * tau_cm), s = 0.0f;
    for (int32_t k = 0; k < n_pts; ++k) {
        const float *p = pts + 3 * (int64_t)k;
        float w[3] = {c[0] + cy * p[0] - sy * p[1], c[1] + sy * p[0] + cy * p[1], c[2] + p[2]};
        int64_t j = grid_at(geo, dims, w);
        if (j >= 0) {
            float d = (float)df[j];
            if (d < tau_cm) s += 1.0f - d * d * inv_t2;
        }
    }
    return s;
}

/* localize.py:55-92 k_score_grid: hypothesis h = position h / n_yaw, heading yaw0 + (h % n_yaw) dyaw; -1
 * where the position is outside the grid or within free_cm of a mapped surface. */
void lwo_score_grid(const float *pts, int32_t n_pts, const float *pos, int32_t n_pos, int32_t n_yaw, float yaw0,
                    float dyaw, const float *geo, const int32_t *dims, const uint8_t *df, float tau_cm,
                    float free_cm, float *out)
{
    for (int64_t h = 0; h < (int64_t)n_pos * n_yaw; ++h) {
        int64_t ip = h / n_yaw, iyaw = h - ip * n_yaw;
        const float *c = pos + 3 * ip;
        int64_t cc = grid_at(geo, dims, c);
        if (cc < 0 || (float)df[cc] < free_cm) {
            out[h] = -1.0f;
            continue;
        }
        float yaw = yaw0 + (float)iyaw * dyaw;
        out[h] = soft_count(pts, n_pts, c, (float)cos((double)yaw), (float)sin((double)yaw), geo, dims, df, tau_cm);
    }
}

/* localize.py:95-123 k_score_list: the soft count of each (x, y, z, yaw) hypothesis. */
void lwo_score_list(const float *pts, int32_t n_pts, const float *hyp, int32_t n_hyp, const float *geo,
                    const int32_t *dims, const uint8_t *df, float tau_cm, float *out)
{
    for (int32_t h = 0; h < n_hyp; ++h) {
        const float *q = hyp + 4 * (int64_t)h;
        out[h] = soft_count(pts, n_pts, q, (float)cos((double)q[3]), (float)sin((double)q[3]), geo, dims, df, tau_cm);
    }
}

static void xform_row(const float *T, const float *p, float *w)
{
    for (int r = 0; r < 3; ++r) w[r] = T[4 * r] * p[0] + T[4 * r + 1] * p[1] + T[4 * r + 2] * p[2] + T[4 * r + 3];
}

/* localize.py:126-180 k_icp_batch then k_reduce_batch, in point order: hypothesis h's accumulator row out[64 h ..]
 * (H upper triangle 0-20, b 21-26, count 27, weighted error 28, weight sum 30; odom.py:73-82), each term in
 * float32 as the reference forms it, summed in float64. T: k row-major float32 4x4 poses. */
void lwo_plane_icp(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo, const int32_t *dims,
                   const int32_t *start, const int32_t *nn, const float *xyz, const float *nrm, const float *planar,
                   float max_d2, float c2, float alpha, double *out)
{
    for (int32_t h = 0; h < k; ++h) {
        double *acc = out + 64 * (int64_t)h;
        for (int j = 0; j < 64; ++j) acc[j] = 0.0;
        for (int32_t i = 0; i < n_pts; ++i) {
            float w[3];
            xform_row(T + 16 * (int64_t)h, pts + 3 * (int64_t)i, w);
            int32_t s = lwo_grid_nearest(w, geo, dims, start, nn, xyz);
            if (s < 0) continue;
            const float *q = xyz + 3 * (int64_t)s;
            float r[3] = {w[0] - q[0], w[1] - q[1], w[2] - q[2]};
            if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] > max_d2) continue;
            const float *n = nrm + 3 * (int64_t)s;
            float M[3][3] = {{alpha, 0.0f, 0.0f}, {0.0f, alpha, 0.0f}, {0.0f, 0.0f, alpha}};
            if (planar[s] > 0.5f && sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) > 0.5f) {
                float f = 1.0f - alpha;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) M[a][b] = n[a] * n[b] * f + M[a][b];
            } else {
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) M[a][b] = M[a][b] * 4.0f;
            }
            float Mr[3];
            for (int a = 0; a < 3; ++a) Mr[a] = M[a][0] * r[0] + M[a][1] * r[1] + M[a][2] * r[2];
            float d2 = r[0] * Mr[0] + r[1] * Mr[1] + r[2] * Mr[2];
            float wgt = 1.0f / (1.0f + d2 / c2);
            wgt = wgt * wgt;
            float J[3][6] = {{1.0f, 0.0f, 0.0f, 0.0f, w[2], -w[1]},
                             {0.0f, 1.0f, 0.0f, -w[2], 0.0f, w[0]},
                             {0.0f, 0.0f, 1.0f, w[1], -w[0], 0.0f}};
            int idx = 0;
            for (int a = 0; a < 6; ++a)
                for (int b = a; b < 6; ++b) {
                    float v = M[0][0] * J[0][a] * J[0][b] + M[0][1] * J[0][a] * J[1][b] + M[0][2] * J[0][a] * J[2][b];
                    v += M[1][0] * J[1][a] * J[0][b] + M[1][1] * J[1][a] * J[1][b] + M[1][2] * J[1][a] * J[2][b];
                    v += M[2][0] * J[2][a] * J[0][b] + M[2][1] * J[2][a] * J[1][b] + M[2][2] * J[2][a] * J[2][b];
                    acc[idx++] += (double)(wgt * v);
                }
            for (int a = 0; a < 6; ++a) acc[21 + a] += (double)(wgt * (Mr[0] * J[0][a] + Mr[1] * J[1][a] + Mr[2] * J[2][a]));
            acc[27] += 1.0;
            acc[28] += (double)(wgt * (r[0] * Mr[0] + r[1] * Mr[1] + r[2] * Mr[2]));
            acc[30] += (double)wgt;
        }
    }
}

/* localize.py:183-206 k_eval_batch: per hypothesis, the points whose nearest map point is within each radius. */
void lwo_plane_eval(const float *pts, int32_t n_pts, const float *T, int32_t k, const float *geo, const int32_t *dims,
                    const int32_t *start, const int32_t *nn, const float *xyz, const float *radii, int32_t *counts)
{
    for (int32_t h = 0; h < k; ++h) {
        counts[3 * h] = counts[3 * h + 1] = counts[3 * h + 2] = 0;
        for (int32_t i = 0; i < n_pts; ++i) {
            float w[3];
            xform_row(T + 16 * (int64_t)h, pts + 3 * (int64_t)i, w);
            int32_t s = lwo_grid_nearest(w, geo, dims, start, nn, xyz);
            if (s < 0) continue;
            float d = sqrtf(dist2(w, xyz + 3 * (int64_t)s));
            for (int r = 0; r < 3; ++r)
                if (d < radii[r]) counts[3 * h + r] += 1;
        }
    }
}
