// The integration map the odometry's frames feed (the reference's Pipeline map: python/livox_warp/gpu.py:373-418,
// 420-595, 1213-1219, 1304-1312, 1346-1372, 1421-1452): voxel records keyed and probed like the registration
// map, the device pose table its inserts and rays read, and free-space carving.
//
// Layout in HBM: one 64-byte record per slot (MapRec: key, hit count, carving misses, the float32 sums of x, y,
// z and reflectivity, first and last time), so an insert's claim and its five atomics, and a ray's probe and
// the centroid it reads, each touch one cache line; the reference keeps six parallel arrays. Poses: 65536
// row-major float32 4x4 matrices indexed by frame id & 65535, scattered from the host's edits. Bricks: one bit
// per hashed 8x8x8-voxel brick (cap / 8 bits, L2-resident), set when a voxel in the brick is claimed, so a ray
// steps through empty bricks without probing the table (a clear bit means no voxel of the brick is in the map;
// a set bit, shared by colliding bricks, only means the voxels are probed).
#include <climits>
#include <cmath>

#include "lw_common.h"
#include "lw_gn.h"
#include "lw_imap.h"

namespace lw {

namespace {

constexpr int IMAP_BLOCK = 256;

// The brick bit of the voxel with key k: its 8x8x8 brick (each 21-bit key field divided by 8, so voxels whose
// keys coincide share a brick), hashed onto nbits (a power of two).
__device__ inline uint32_t brick_bit(int64_t key, int32_t nbits) {
    const uint64_t k = (uint64_t)key;
    const uint64_t b = ((k & KEY_MASK) >> 3) | ((((k >> KEY_BITS) & KEY_MASK) >> 3) << KEY_BITS) |
                       ((((k >> (2 * KEY_BITS)) & KEY_MASK) >> 3) << (2 * KEY_BITS));
    return (uint32_t)(slot_hash((int64_t)b) & (nbits - 1));
}

__device__ inline void brick_set(uint32_t *bricks, uint32_t b) { atomicOr(&bricks[b >> 5], 1u << (b & 31)); }

__global__ void k_fill_u32(uint32_t *p, int64_t n, uint32_t v) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// The brick bits of every occupied slot (after records were loaded from the host).
__global__ void k_brick_rebuild(const MapRec *recs, int64_t cap, uint32_t *bricks, int32_t nbits) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < cap; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t key = recs[s].key;
        if (key != -1) brick_set(bricks, brick_bit(key, nbits));
    }
}

__global__ void k_imap_clear(MapRec *recs, int64_t cap, int32_t *stats) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < cap; s += (int64_t)gridDim.x * blockDim.x) {
        MapRec r;
        r.key = -1;
        r.cnt = 0;
        r.miss = 0;
        r.acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        r.first = 1.0e30f;
        r.last = 0.0f;
        for (int k = 0; k < 6; ++k) r.pad[k] = 0;
        recs[s] = r;
    }
    if (blockIdx.x == 0 && threadIdx.x < 2) stats[threadIdx.x] = 0;
}

// gpu.py:1346-1365 (k_pose_scatter, pose part): pose idx[k] = mats[k].
__global__ void k_pose_scatter(const int32_t *idx, const float *mats, int32_t n, float *poses) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < (int64_t)n * 16;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = k / 16, c = k % 16;
        poses[(int64_t)(idx[e] & (IMAP_POSE_SLOTS - 1)) * 16 + c] = mats[k];
    }
}

// gpu.py:156-167 (keep_raw): the raw-return filters.
__device__ inline bool keep_raw(float px, float py, float pz, uint32_t a, const MapView &v) {
    const float r = length3(px, py, pz);
    if (r < v.min_range || r > v.max_range) return false;
    if ((int)(a & 255u) < v.min_refl) return false;
    if (((v.noise_mask >> (tag_of(a) & 3)) & 1) == 0) return false;
    if (((v.ret_mask >> (((int)(a >> 16) & 255) & 3)) & 1) == 0) return false;
    return true;
}

// gpu.py:373-413 (k_map_insert): the point through its frame's pose, claimed at its voxel's key; hits keep
// counting, the float32 sums stop after IMAP_COUNT_CAP of them, first and last time by atomic min and max.
//
// Points arrive in scan order, so neighbouring lanes often fall in one voxel: a wave first combines each run of
// consecutive lanes with the same key (a segmented shuffle reduction of the sums, the count and the time
// range), and the run's first lane claims the slot and makes one set of atomics for the whole run. The count
// still grows by one per point; atomicAdd on it hands out the run's range of old counts, and exactly the
// points below IMAP_COUNT_CAP add to the sums: all of the run at once, or, for the run that crosses the cap,
// its first (cap - old) points one by one. Which points make up the capped sums depends on the arrival order,
// as it does in the reference's per-point atomics.
__global__ __launch_bounds__(IMAP_BLOCK) void k_imap_insert(const float *sx, const uint32_t *sa, const float *st,
                                                            const int32_t *sf, int32_t n, const float *poses,
                                                            MapView v, float inv_voxel, int32_t mask, MapRec *recs,
                                                            int32_t *stats, uint32_t *bricks, int32_t nbits) {
    const int32_t i = blockIdx.x * IMAP_BLOCK + threadIdx.x;
    const int lane = lane_id();
    bool live = i < n;
    float wx = 0.0f, wy = 0.0f, wz = 0.0f, refl = 0.0f, t = 0.0f;
    int64_t key = -2 - lane;  // a dead lane's key matches no other lane's (voxel keys are non-negative)
    if (live) {
        const float px = sx[3 * (int64_t)i], py = sx[3 * (int64_t)i + 1], pz = sx[3 * (int64_t)i + 2];
        const uint32_t a = sa[i];
        live = keep_raw(px, py, pz, a, v);
        if (live) {
            const float *T = poses + (int64_t)(sf[i] & (IMAP_POSE_SLOTS - 1)) * 16;
            transform_point(T, px, py, pz, wx, wy, wz);
            key = voxel_key_ijk(ifloor(wx * inv_voxel), ifloor(wy * inv_voxel), ifloor(wz * inv_voxel));
            refl = (float)(a & 255u);
            t = st[i];
        }
    }
    // runs of equal keys: a run starts where the key differs from the lane below's
    const int64_t below = __shfl_up(key, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || below != key);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const int head = 63 - __clzll(heads & upto);
    const uint64_t after = heads & ~upto;
    const int last = after ? __ffsll((long long)after) - 2 : 63;
    // segmented sums over lanes [lane, last]: the run's totals end up at its head
    float rx = wx, ry = wy, rz = wz, rr = refl, t_min = live ? t : 1.0e30f, t_max = live ? t : -1.0e30f;
    int32_t cnt = live ? 1 : 0;
    LW_UNROLL for (int off = 1; off < 64; off <<= 1) {
        const float ox = __shfl_down(rx, off, 64), oy = __shfl_down(ry, off, 64), oz = __shfl_down(rz, off, 64);
        const float orr = __shfl_down(rr, off, 64), omin = __shfl_down(t_min, off, 64);
        const float omax = __shfl_down(t_max, off, 64);
        const int32_t oc = __shfl_down(cnt, off, 64);
        if (lane + off <= last) {
            rx = rx + ox;
            ry = ry + oy;
            rz = rz + oz;
            rr = rr + orr;
            t_min = fminf(t_min, omin);
            t_max = fmaxf(t_max, omax);
            cnt += oc;
        }
    }
    // the head claims the slot (probing as find_slot does) and counts the whole run
    int32_t slot = -1, old = 0;
    if (lane == head && live) {
        const int32_t h = slot_hash(key) & mask;
        for (int probe = 0; probe < PROBES; ++probe) {
            MapRec *r = &recs[(h + probe) & mask];
            const unsigned long long prev = atomicCAS(reinterpret_cast<unsigned long long *>(&r->key),
                                                      0xFFFFFFFFFFFFFFFFull, (unsigned long long)key);
            if ((int64_t)prev == -1 || (int64_t)prev == key) {
                if ((int64_t)prev == -1) {
                    atomicAdd(&stats[0], 1);
                    brick_set(bricks, brick_bit(key, nbits));
                }
                slot = (h + probe) & mask;
                old = atomicAdd(&r->cnt, cnt);
                atomicMax(&r->last, t_max);
                atomicMin(&r->first, t_min);
                break;
            }
        }
        if (slot < 0) atomicAdd(&stats[1], cnt);  // no slot along this probe chain
    }
    slot = __shfl(slot, head, 64);
    old = __shfl(old, head, 64);
    if (slot < 0 || !live) return;
    MapRec *r = &recs[slot];
    const int32_t room = IMAP_COUNT_CAP - old;  // points of this run that still add to the sums
    const int32_t run = __shfl(cnt, head, 64);
    if (room >= run) {
        if (lane == head) {
            unsafeAtomicAdd(&r->acc.x, rx);
            unsafeAtomicAdd(&r->acc.y, ry);
            unsafeAtomicAdd(&r->acc.z, rz);
            unsafeAtomicAdd(&r->acc.w, rr);
        }
    } else if (lane - head < room) {  // the run that crosses the cap: its first points, one by one
        unsafeAtomicAdd(&r->acc.x, wx);
        unsafeAtomicAdd(&r->acc.y, wy);
        unsafeAtomicAdd(&r->acc.z, wz);
        unsafeAtomicAdd(&r->acc.w, refl);
    }
}

// gpu.py:247-265 (find_slot) on the records: the slot holding key, or -1.
__device__ inline int32_t imap_find(const MapRec *recs, int32_t mask, int64_t key) {
    const int32_t h = slot_hash(key) & mask;
    for (int probe = 0; probe < PROBES; ++probe) {
        const int32_t s = (h + probe) & mask;
        const int64_t k = recs[s].key;
        if (k == key) return s;
        if (k == -1) return -1;
    }
    return -1;
}

// gpu.py:415-418 (map_centroid).
__device__ inline void imap_centroid(const MapRec &r, float &cx, float &cy, float &cz) {
    const float m = 1.0f / (float)(r.cnt < IMAP_COUNT_CAP ? r.cnt : IMAP_COUNT_CAP);
    cx = r.acc.x * m;
    cy = r.acc.y * m;
    cz = r.acc.z * m;
}

// A slot's first 32 bytes: {key lo, key hi, hits, misses} and the sums.
__device__ inline void load_head(const MapRec *recs, int32_t s, int4 &lo, float4 &acc) {
    const int4 *p = reinterpret_cast<const int4 *>(recs + s);
    lo = p[0];
    acc = reinterpret_cast<const float4 *>(p)[1];
}

// gpu.py:420-467 (local_plane): the normal of the surface through voxel (ix, iy, iz) from the centroids of
// the occupied voxels around it (relative to origin o), w = 2 for a surface, 1 for at least 4 occupied voxels
// that are not one, 0 for fewer. The 27 neighbours go in three slabs of nine (gx = -1, 0, 1): a slab's home
// records are loaded together, its keys not at home probe on together, and the centroids then enter the
// sums in the reference's (gx, gy, gz) order. The float32 covariance is formed as the reference forms it;
// its eigen-decomposition is float64 Jacobi (the reference's wp.eig3 is float32), as for the voxel precision.
__device__ inline float4 local_plane(int32_t ix, int32_t iy, int32_t iz, float ox, float oy, float oz,
                                     const MapRec *recs, int32_t mask) {
    float cs[3] = {0.0f, 0.0f, 0.0f};
    float cc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    float cn = 0.0f;
    LW_UNROLL for (int gx = -1; gx < 2; ++gx) {
        int64_t key[9];
        int32_t h[9];
        int4 lo[9];
        float4 acc[9];
        LW_UNROLL for (int c = 0; c < 9; ++c) {
            key[c] = voxel_key_ijk(ix + gx, iy + c / 3 - 1, iz + c % 3 - 1);
            h[c] = slot_hash(key[c]) & mask;
        }
        LW_UNROLL for (int c = 0; c < 9; ++c) load_head(recs, h[c], lo[c], acc[c]);
        // keys whose home holds another key probe on, one slot further per round, all of the slab together
        uint32_t need = 0u;
        LW_UNROLL for (int c = 0; c < 9; ++c) {
            const int64_t k = (int64_t)(((uint64_t)(uint32_t)lo[c].y << 32) | (uint64_t)(uint32_t)lo[c].x);
            need |= (k != key[c] && k != -1 ? 1u : 0u) << c;
        }
        for (int probe = 1; need != 0u && probe < PROBES; ++probe) {
            LW_UNROLL for (int c = 0; c < 9; ++c) {
                if (!((need >> c) & 1u)) continue;
                int4 l;
                float4 q;
                load_head(recs, (h[c] + probe) & mask, l, q);
                const int64_t k = (int64_t)(((uint64_t)(uint32_t)l.y << 32) | (uint64_t)(uint32_t)l.x);
                if (k == key[c] || k == -1) {
                    lo[c] = l;
                    acc[c] = q;
                    need &= ~(1u << c);
                }
            }
        }
        LW_UNROLL for (int c = 0; c < 9; ++c) {
            const int64_t k = (int64_t)(((uint64_t)(uint32_t)lo[c].y << 32) | (uint64_t)(uint32_t)lo[c].x);
            const bool found = ((need >> c) & 1u) == 0u && k == key[c];  // past PROBES: not found
            if (found && lo[c].z > 0) {
                const float m = 1.0f / (float)(lo[c].z < IMAP_COUNT_CAP ? lo[c].z : IMAP_COUNT_CAP);
                const float q[3] = {acc[c].x * m - ox, acc[c].y * m - oy, acc[c].z * m - oz};
                for (int a = 0; a < 3; ++a) cs[a] = cs[a] + q[a];
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) cc[3 * a + b] = cc[3 * a + b] + q[a] * q[b];
                cn = cn + 1.0f;
            }
        }
    }
    if (cn < 4.0f) return make_float4(0.0f, 0.0f, 1.0f, 0.0f);
    const float mu[3] = {cs[0] / cn, cs[1] / cn, cs[2] / cn};
    double cov[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) cov[3 * a + b] = (double)(cc[3 * a + b] / cn - mu[a] * mu[b]);
    double ev[3], Q[9];
    gn::sym_eig<3>(cov, ev, Q);
    int k = 0;
    if (ev[1] < ev[k]) k = 1;
    if (ev[2] < ev[k]) k = 2;
    const int j = (k + 1) % 3, j2 = (k + 2) % 3;
    const double mid = ev[j] < ev[j2] ? ev[j] : ev[j2];
    const float ok = ev[k] < 0.25 * mid ? 2.0f : 1.0f;
    return make_float4((float)Q[k], (float)Q[3 + k], (float)Q[6 + k], ok);
}

// gpu.py:470-595 (k_carve): walk every `every`-th return's ray through the map (a 3D DDA from the sensor
// origin) up to a margin short of the return, and add `every` misses to each occupied voxel whose surface
// patch the ray pierces. A ray takes CARVE_LANES lanes: each runs the same float32 DDA (so every lane sees the
// reference's voxel sequence) in rounds of CARVE_LANES steps, and lane j looks up the round's j-th voxel, so
// the lanes of a ray wait on memory together, once a round, instead of once a voxel.
constexpr int CARVE_LANES = 16;
__global__ __launch_bounds__(IMAP_BLOCK) void k_imap_carve(const float *sx, const uint32_t *sa, const int32_t *sf,
                                                           int32_t n, int32_t every, const float *poses,
                                                           float inv_voxel, int32_t mask, float max_range,
                                                           int32_t max_steps, MapRec *recs, const uint32_t *bricks,
                                                           int32_t nbits) {
    const int64_t gt = (int64_t)blockIdx.x * IMAP_BLOCK + threadIdx.x;
    const int32_t j = (int32_t)(gt % CARVE_LANES);
    const int64_t i = gt / CARVE_LANES * every;
    if (i >= n) return;
    if (tag_of(sa[i]) != 0) return;
    const float *T = poses + (int64_t)(sf[i] & (IMAP_POSE_SLOTS - 1)) * 16;
    const float ox = T[3], oy = T[7], oz = T[11];
    float wx, wy, wz;
    transform_point(T, sx[3 * i], sx[3 * i + 1], sx[3 * i + 2], wx, wy, wz);
    float dx = wx - ox, dy = wy - oy, dz = wz - oz;
    const float L = length3(dx, dy, dz);
    if (L < IMAP_CARVE_MIN_RANGE || L > max_range) return;
    dx = dx / L;
    dy = dy / L;
    dz = dz / L;
    const float voxel = 1.0f / inv_voxel;
    float margin = fmaxf(IMAP_CARVE_MARGIN_VOXELS * voxel, IMAP_CARVE_MARGIN_FRACTION * L);
    const float4 pe = local_plane(ifloor(wx * inv_voxel), ifloor(wy * inv_voxel), ifloor(wz * inv_voxel), wx, wy, wz,
                                  recs, mask);
    if (pe.w > 0.5f) {
        float c = pe.x * dx;
        c = c + pe.y * dy;
        c = c + pe.z * dz;
        const float cosine = fabsf(c);
        margin = fmaxf(margin, IMAP_CARVE_GRAZING_VOXELS * voxel / fmaxf(cosine, IMAP_CARVE_MIN_COSINE));
    }
    const float stop = (L - margin) * inv_voxel;
    if (stop <= 1.0f) return;
    const float px = ox * inv_voxel, py = oy * inv_voxel, pz = oz * inv_voxel;
    int32_t ix = ifloor(px), iy = ifloor(py), iz = ifloor(pz);
    int32_t sxs = 1, sys = 1, szs = 1;
    float tmx = 1.0e30f, tmy = 1.0e30f, tmz = 1.0e30f, tdx = 1.0e30f, tdy = 1.0e30f, tdz = 1.0e30f;
    if (dx > 1.0e-9f) {
        tmx = ((float)(ix + 1) - px) / dx;
        tdx = 1.0f / dx;
    } else if (dx < -1.0e-9f) {
        sxs = -1;
        tmx = (px - (float)ix) / (-dx);
        tdx = -1.0f / dx;
    }
    if (dy > 1.0e-9f) {
        tmy = ((float)(iy + 1) - py) / dy;
        tdy = 1.0f / dy;
    } else if (dy < -1.0e-9f) {
        sys = -1;
        tmy = (py - (float)iy) / (-dy);
        tdy = -1.0f / dy;
    }
    if (dz > 1.0e-9f) {
        tmz = ((float)(iz + 1) - pz) / dz;
        tdz = 1.0f / dz;
    } else if (dz < -1.0e-9f) {
        szs = -1;
        tmz = (pz - (float)iz) / (-dz);
        tdz = -1.0f / dz;
    }
    int32_t bx = INT_MIN, by = 0, bz = 0;  // the brick of this lane's last voxel, and whether it is empty
    bool empty = true;
    int32_t step = 0;
    bool done = false;
    while (!done) {
        // a round: CARVE_LANES steps of the DDA (the same on every lane of the ray); this lane keeps the j-th
        int32_t vx = 0, vy = 0, vz = 0;
        bool have = false;
        LW_UNROLL for (int u = 0; u < CARVE_LANES; ++u) {
            if (!done) {
                if (step >= max_steps) {
                    done = true;
                } else {
                    ++step;
                    float t;
                    if (tmx < tmy && tmx < tmz) {
                        t = tmx;
                        ix += sxs;
                        tmx = tmx + tdx;
                    } else if (tmy < tmz) {
                        t = tmy;
                        iy += sys;
                        tmy = tmy + tdy;
                    } else {
                        t = tmz;
                        iz += szs;
                        tmz = tmz + tdz;
                    }
                    if (t >= stop) {
                        done = true;
                    } else if (u == j) {
                        vx = ix;
                        vy = iy;
                        vz = iz;
                        have = true;
                    }
                }
            }
        }
        if (!have) continue;
        if ((vx >> 3) != bx || (vy >> 3) != by || (vz >> 3) != bz) {
            bx = vx >> 3;
            by = vy >> 3;
            bz = vz >> 3;
            const uint32_t b = brick_bit(voxel_key_ijk(vx, vy, vz), nbits);
            empty = ((bricks[b >> 5] >> (b & 31)) & 1u) == 0u;
        }
        if (empty) continue;  // no voxel of this brick is in the map
        const int32_t s = imap_find(recs, mask, voxel_key_ijk(vx, vy, vz));
        if (s < 0) continue;
        const MapRec r = recs[s];
        if (r.cnt <= 0) continue;
        float cx, cy, cz;
        imap_centroid(r, cx, cy, cz);
        const float rx = cx - ox, ry = cy - oy, rz = cz - oz;
        bool pierced = false;
        const float4 pl = local_plane(vx, vy, vz, cx, cy, cz, recs, mask);
        if (pl.w > 1.5f) {
            float dn = dx * pl.x;
            dn = dn + dy * pl.y;
            dn = dn + dz * pl.z;
            if (fabsf(dn) > 1.0e-4f) {
                float rn = rx * pl.x;
                rn = rn + ry * pl.y;
                rn = rn + rz * pl.z;
                const float f = rn / dn;
                const float x0 = dx * f - rx, x1 = dy * f - ry, x2 = dz * f - rz;
                if (length3(x0, x1, x2) < IMAP_CARVE_PIERCE_PLANE * voxel) pierced = true;
            }
        } else {
            float rd = rx * dx;
            rd = rd + ry * dy;
            rd = rd + rz * dz;
            const float q0 = rx - rd * dx, q1 = ry - rd * dy, q2 = rz - rd * dz;
            if (length3(q0, q1, q2) < IMAP_CARVE_PIERCE_POINT * voxel) pierced = true;
        }
        if (pierced) atomicAdd(&recs[s].miss, every);
    }
}

// gpu.py:1367-1372 (k_rebase_clock): first and last time of every occupied slot onto the new clock.
__global__ void k_imap_rebase(MapRec *recs, int64_t cap, float shift) {
    for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < cap; s += (int64_t)gridDim.x * blockDim.x) {
        if (recs[s].key == -1) continue;
        recs[s].first = recs[s].first - shift;
        recs[s].last = recs[s].last - shift;
    }
}

int grid_of(int64_t n, int64_t max_grid) {
    int64_t g = (n + IMAP_BLOCK - 1) / IMAP_BLOCK;
    if (g < 1) g = 1;
    return (int)(g < max_grid ? g : max_grid);
}

}  // namespace

int32_t imap_brick_bits(int64_t cap) { return cap / 8 > 1024 ? (int32_t)(cap / 8) : 1024; }

hipError_t imap_clear(hipStream_t s, MapRec *recs, int64_t cap, int32_t *stats, uint32_t *bricks) {
    hipLaunchKernelGGL(k_imap_clear, dim3(grid_of(cap, 8192)), dim3(IMAP_BLOCK), 0, s, recs, cap, stats);
    const int64_t words = imap_brick_bits(cap) / 32;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(words, 4096)), dim3(IMAP_BLOCK), 0, s, bricks, words, 0u);
    return hipGetLastError();
}

hipError_t imap_brick_rebuild(hipStream_t s, const MapRec *recs, int64_t cap, uint32_t *bricks) {
    const int64_t words = imap_brick_bits(cap) / 32;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(words, 4096)), dim3(IMAP_BLOCK), 0, s, bricks, words, 0u);
    hipLaunchKernelGGL(k_brick_rebuild, dim3(grid_of(cap, 8192)), dim3(IMAP_BLOCK), 0, s, recs, cap, bricks,
                       imap_brick_bits(cap));
    return hipGetLastError();
}

hipError_t imap_pose_scatter(hipStream_t s, const int32_t *idx, const float *mats, int32_t n, float *poses) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_pose_scatter, dim3(grid_of((int64_t)n * 16, 65535)), dim3(IMAP_BLOCK), 0, s, idx, mats, n,
                       poses);
    return hipGetLastError();
}

hipError_t imap_insert(hipStream_t s, const float *sx, const uint32_t *sa, const float *st, const int32_t *sf,
                       int32_t n, const float *poses, const MapView &v, float inv_voxel, int32_t mask, MapRec *recs,
                       int32_t *stats, uint32_t *bricks) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_imap_insert, dim3(grid_of(n, INT_MAX)), dim3(IMAP_BLOCK), 0, s, sx, sa, st, sf, n, poses, v,
                       inv_voxel, mask, recs, stats, bricks, imap_brick_bits((int64_t)mask + 1));
    return hipGetLastError();
}

hipError_t imap_carve(hipStream_t s, const float *sx, const uint32_t *sa, const int32_t *sf, int32_t n, int32_t every,
                      const float *poses, float inv_voxel, int32_t mask, float max_range, int32_t max_steps,
                      MapRec *recs, const uint32_t *bricks) {
    if (n <= 0) return hipSuccess;
    const int64_t rays = ((int64_t)n + every - 1) / every;
    hipLaunchKernelGGL(k_imap_carve, dim3(grid_of(rays * CARVE_LANES, INT_MAX)), dim3(IMAP_BLOCK), 0, s, sx, sa, sf, n, every, poses,
                       inv_voxel, mask, max_range, max_steps, recs, bricks, imap_brick_bits((int64_t)mask + 1));
    return hipGetLastError();
}

hipError_t imap_rebase(hipStream_t s, MapRec *recs, int64_t cap, float shift) {
    hipLaunchKernelGGL(k_imap_rebase, dim3(grid_of(cap, 8192)), dim3(IMAP_BLOCK), 0, s, recs, cap, shift);
    return hipGetLastError();
}

}  // namespace lw

// ---- the C ABI (include/livox_warp_amd.h) -----------------------------------------------------------------

#include <cstring>
#include <new>

#include "../../include/livox_warp_amd.h"

struct lw_imap {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t cap = 0;
    lw::MapRec *recs = nullptr;
    int32_t *stats = nullptr;
    float *poses = nullptr;    // IMAP_POSE_SLOTS x 16
    uint32_t *bricks = nullptr;  // imap_brick_bits(cap) bits
    int32_t *up_idx = nullptr;  // pose upload staging on the device
    float *up_mat = nullptr;
    void *pin = nullptr;       // pinned host staging: the pose uploads, then the stats readback
    hipEvent_t up_done = nullptr;  // the last pose upload has left the pinned staging
};

namespace {

constexpr size_t IMAP_PIN = (size_t)lw::IMAP_POSE_SLOTS * (16 * sizeof(float) + sizeof(int32_t));

#define IMAP_CK(x)                               \
    do {                                         \
        hipError_t e_ = (x);                     \
        if (e_ != hipSuccess) return -(int)e_;   \
    } while (0)

void imap_free(lw_imap *m) {
    if (m->recs) (void)hipFree(m->recs);
    if (m->stats) (void)hipFree(m->stats);
    if (m->bricks) (void)hipFree(m->bricks);
    if (m->poses) (void)hipFree(m->poses);
    if (m->up_idx) (void)hipFree(m->up_idx);
    if (m->up_mat) (void)hipFree(m->up_mat);
    if (m->pin) (void)hipHostFree(m->pin);
    if (m->up_done) (void)hipEventDestroy(m->up_done);
}

__global__ void k_pose_identity(float *poses) {
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < (int64_t)lw::IMAP_POSE_SLOTS * 16) poses[k] = (k % 16) % 5 == 0 ? 1.0f : 0.0f;
}

}  // namespace

extern "C" {

int lw_imap_create(lw_imap **out, int64_t slots, void *stream) {
    if (!out || slots < 64 || slots > (1ll << 27) || (slots & (slots - 1)) != 0) return LW_EINVAL;
    lw_imap *m = new (std::nothrow) lw_imap();
    if (!m) return LW_EINVAL;
    auto fail = [&](hipError_t e) {
        imap_free(m);
        delete m;
        return -(int)e;
    };
    hipError_t e;
    if ((e = hipGetDevice(&m->device)) != hipSuccess) return fail(e);
    m->stream = (hipStream_t)stream;
    m->cap = slots;
    if ((e = hipMalloc((void **)&m->recs, (size_t)slots * sizeof(lw::MapRec))) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->stats, 64)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->bricks, (size_t)lw::imap_brick_bits(slots) / 8)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->poses, (size_t)lw::IMAP_POSE_SLOTS * 16 * sizeof(float))) != hipSuccess)
        return fail(e);
    if ((e = hipMalloc((void **)&m->up_idx, (size_t)lw::IMAP_POSE_SLOTS * sizeof(int32_t))) != hipSuccess)
        return fail(e);
    if ((e = hipMalloc((void **)&m->up_mat, (size_t)lw::IMAP_POSE_SLOTS * 16 * sizeof(float))) != hipSuccess)
        return fail(e);
    if ((e = hipHostMalloc(&m->pin, IMAP_PIN, hipHostMallocDefault)) != hipSuccess) return fail(e);
    if ((e = hipEventCreateWithFlags(&m->up_done, hipEventDisableTiming)) != hipSuccess) return fail(e);
    hipLaunchKernelGGL(k_pose_identity, dim3((lw::IMAP_POSE_SLOTS * 16 + 255) / 256), dim3(256), 0, m->stream,
                       m->poses);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = lw::imap_clear(m->stream, m->recs, m->cap, m->stats, m->bricks)) != hipSuccess) return fail(e);
    if ((e = hipEventRecord(m->up_done, m->stream)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(m->stream)) != hipSuccess) return fail(e);
    *out = m;
    return 0;
}

int lw_imap_destroy(lw_imap *m) {
    if (!m) return LW_EINVAL;
    (void)hipStreamSynchronize(m->stream);
    imap_free(m);
    delete m;
    return 0;
}

int lw_imap_clear(lw_imap *m) {
    if (!m) return LW_EINVAL;
    IMAP_CK(lw::imap_clear(m->stream, m->recs, m->cap, m->stats, m->bricks));
    return 0;
}

int lw_imap_set_poses(lw_imap *m, const int32_t *idx, const float *mats, int32_t n) {
    if (!m || n < 0 || n > lw::IMAP_POSE_SLOTS || (n > 0 && (!idx || !mats))) return LW_EINVAL;
    if (n == 0) return 0;
    IMAP_CK(hipEventSynchronize(m->up_done));  // the previous upload has left the staging
    char *p = (char *)m->pin;
    std::memcpy(p, idx, (size_t)n * sizeof(int32_t));
    std::memcpy(p + (size_t)lw::IMAP_POSE_SLOTS * sizeof(int32_t), mats, (size_t)n * 16 * sizeof(float));
    IMAP_CK(hipMemcpyAsync(m->up_idx, p, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, m->stream));
    IMAP_CK(hipMemcpyAsync(m->up_mat, p + (size_t)lw::IMAP_POSE_SLOTS * sizeof(int32_t), (size_t)n * 16 * sizeof(float),
                           hipMemcpyHostToDevice, m->stream));
    IMAP_CK(hipEventRecord(m->up_done, m->stream));
    IMAP_CK(lw::imap_pose_scatter(m->stream, m->up_idx, m->up_mat, n, m->poses));
    return 0;
}

int lw_imap_insert(lw_imap *m, const float *sx, const uint32_t *sa, const float *st, const int32_t *sf, int32_t n,
                   const lw_map_view *v, float voxel) {
    if (!m || !v || n < 0 || (n > 0 && (!sx || !sa || !st || !sf))) return LW_EINVAL;
    const lw::MapView mv{v->min_range, v->max_range, v->min_refl, v->noise_mask, v->ret_mask};
    const float inv = 1.0f / (voxel > 1e-4f ? voxel : 1e-4f);  // 1.0 / max(voxel, 1e-4) (gpu.py:1430)
    IMAP_CK(lw::imap_insert(m->stream, sx, sa, st, sf, n, m->poses, mv, inv, (int32_t)(m->cap - 1), m->recs, m->stats,
                            m->bricks));
    return 0;
}

int lw_imap_carve(lw_imap *m, const float *sx, const uint32_t *sa, const int32_t *sf, int32_t n, float voxel,
                  int32_t every, float max_range, int32_t max_steps) {
    if (!m || n < 0 || max_steps < 0 || (n > 0 && (!sx || !sa || !sf))) return LW_EINVAL;
    every = every > 1 ? every : 1;  // max(1, int(every)) (gpu.py:1443)
    const float inv = 1.0f / (voxel > 1e-4f ? voxel : 1e-4f);
    IMAP_CK(lw::imap_carve(m->stream, sx, sa, sf, n, every, m->poses, inv, (int32_t)(m->cap - 1), max_range, max_steps,
                           m->recs, m->bricks));
    return 0;
}

int lw_imap_rebase_clock(lw_imap *m, float shift) {
    if (!m) return LW_EINVAL;
    IMAP_CK(lw::imap_rebase(m->stream, m->recs, m->cap, shift));
    return 0;
}

int lw_imap_stats(lw_imap *m, int32_t *out) {
    if (!m || !out) return LW_EINVAL;
    IMAP_CK(hipEventSynchronize(m->up_done));
    IMAP_CK(hipMemcpyAsync(m->pin, m->stats, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, m->stream));
    IMAP_CK(hipStreamSynchronize(m->stream));
    std::memcpy(out, m->pin, 2 * sizeof(int32_t));
    return 0;
}

int lw_imap_read(lw_imap *m, void *out, int64_t bytes) {
    if (!m || !out || bytes != m->cap * (int64_t)sizeof(lw::MapRec)) return LW_EINVAL;
    IMAP_CK(hipStreamSynchronize(m->stream));
    IMAP_CK(hipMemcpy(out, m->recs, (size_t)bytes, hipMemcpyDeviceToHost));
    return 0;
}

int lw_imap_load(lw_imap *m, const void *in, int64_t bytes) {
    if (!m || !in || bytes != m->cap * (int64_t)sizeof(lw::MapRec)) return LW_EINVAL;
    IMAP_CK(hipStreamSynchronize(m->stream));
    IMAP_CK(hipMemcpy(m->recs, in, (size_t)bytes, hipMemcpyHostToDevice));
    IMAP_CK(lw::imap_brick_rebuild(m->stream, m->recs, m->cap, m->bricks));
    return 0;
}

}  // extern "C"
