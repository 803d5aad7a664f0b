// The closest-mean neighbour search shared by the odometry accumulation (lw_icp.hip) and the batched
// localisation (lw_loc.hip): odom.py:308-334 with find_slot (gpu.py:254-265).
//
// A lane resolves a chunk of C neighbour keys together. Round 0 loads every key's home slot record (slot_hash
// & mask; 32 B: key, voxel id, count, mean, displacement bound) at once, with no branch around the loads, so
// they are all in flight together; a record holding the key is a hit. A key whose home slot holds another key
// can only lie within the home's displacement bound (Slot::disp, the largest probe distance of the keys homed
// there, kept by the inserts): a bound of 0 (the usual case, and always for an empty slot) ends the lookup as
// a miss at once, otherwise the slots up to the bound are read, the records of every key of the chunk still
// unresolved at a depth loaded together. The slot found is the one find_slot's linear probe reaches
// (gpu.py:254-265): keys are never removed, so a key homed at h sits within h .. h + disp(h).
//
// With BITMAP set, the occupancy bitmap is read first and a key whose home bit is clear loads slot 0's record
// instead of its own (cached by every lane, so it costs no memory traffic) and is a miss: one more dependent
// load, for far less traffic on a large, sparse table (the batched localisation's throughput-bound launch);
// the odometry's latency-bound loop reads the home records directly.
//
// The closest mean is chosen by (squared distance, position in the reference's neighbour order), so the
// order in which candidates are evaluated does not matter: the result is the serial loop's strict-< choice.
#pragma once

#include "lw_common.h"
#include "lw_gn.h"

namespace lw {

struct Cand {
    float d2;
    int32_t s;  // slot, -1 for none
    int32_t j;  // position in the reference's neighbour order
    float mx, my, mz;
    int32_t vid;  // the slot's voxel id (read with the record, so the terms need no second lookup)
};

struct Table {
    const Slot *slots;
    const uint32_t *occ;  // one bit per slot
    int32_t mask;
};

// A slot as two 16-byte loads: lo = {key lo, key hi, voxel id, count bits}, hi = {mean x, y, z, disp bits}.
__device__ inline void load_slot(const Slot *slots, int32_t s, int4 &lo, float4 &hi) {
    const int4 *p = reinterpret_cast<const int4 *>(slots + s);
    lo = p[0];
    hi = reinterpret_cast<const float4 *>(p)[1];
}

__device__ inline int64_t key_of(const int4 &lo) {
    return (int64_t)(((uint64_t)(uint32_t)lo.y << 32) | (uint64_t)(uint32_t)lo.x);
}

__device__ inline bool bit_of(uint32_t word, int32_t s) { return ((word >> (s & 31)) & 1u) != 0u; }

// One candidate: skip count 0 (odom.py:324), then keep it if it is closer, or as close and earlier.
__device__ inline void consider(float wx, float wy, float wz, int32_t s, int32_t j, const int4 &lo, const float4 &hi,
                                Cand &best) {
    if (!(__int_as_float(lo.w) > 0.0f)) return;
    float d0 = wx - hi.x, d1 = wy - hi.y, d2 = wz - hi.z;
    float dd = d0 * d0;
    dd = dd + d1 * d1;
    dd = dd + d2 * d2;
    // the serial loop starts from best_d2 = max_d2 and takes strictly smaller distances (odom.py:326-330)
    const bool take = best.s < 0 ? dd < best.d2 : (dd < best.d2 || (dd == best.d2 && j < best.j));
    if (take) {
        best.d2 = dd;
        best.s = s;
        best.j = j;
        best.mx = hi.x;
        best.my = hi.y;
        best.mz = hi.z;
        best.vid = lo.z;
    }
}

// The closest mean over neighbours [j_lo, j_hi) of the (2R+1)^3 block around the point's voxel (dx outer, dz
// inner), in chunks of C. hits counts the occupied neighbours found.
template <int R, int C, bool BITMAP>
__device__ inline void search_part(const Table &t, float inv_v, float wx, float wy, float wz, int j_lo, int j_hi,
                                   Cand &best, int32_t &hits) {
    constexpr int D = 2 * R + 1;
    static_assert(C <= 32, "one bit per key of a chunk");
    const int32_t bx = ifloor(wx * inv_v), by = ifloor(wy * inv_v), bz = ifloor(wz * inv_v);
    for (int j0 = j_lo; j0 < j_hi; j0 += C) {
        int64_t key[C];
        int32_t h[C];
        LW_UNROLL for (int c = 0; c < C; ++c) {
            const int j = j0 + c;  // past j_hi: a real voxel's key, looked up and then ignored
            const int dx = j / (D * D) - R, dy = (j / D) % D - R, dz = j % D - R;
            key[c] = voxel_key_ijk(bx + dx, by + dy, bz + dz);
            h[c] = slot_hash(key[c]) & t.mask;
        }
        uint32_t live = 0u;  // keys of this chunk inside [j_lo, j_hi) (and, with BITMAP, with their home bit set)
        LW_UNROLL for (int c = 0; c < C; ++c) live |= (j0 + c < j_hi ? 1u : 0u) << c;
        if constexpr (BITMAP) {
            uint32_t w[C];
            LW_UNROLL for (int c = 0; c < C; ++c) w[c] = t.occ[h[c] >> 5];
            LW_UNROLL for (int c = 0; c < C; ++c) live &= ~((bit_of(w[c], h[c]) ? 0u : 1u) << c);
        }
        // round 0: the home slots, every load issued before any is used
        uint32_t need = 0u;
        int32_t dmax[C];
        {
            int4 lo[C];
            float4 hi[C];
            LW_UNROLL for (int c = 0; c < C; ++c)
                load_slot(t.slots, BITMAP && !((live >> c) & 1u) ? 0 : h[c], lo[c], hi[c]);
            LW_UNROLL for (int c = 0; c < C; ++c) {
                dmax[c] = 0;
                if (!((live >> c) & 1u)) continue;
                if (key_of(lo[c]) != key[c]) {
                    dmax[c] = __float_as_int(hi[c].w);  // the home's displacement bound
                    if (dmax[c] > 0) need |= 1u << c;
                    continue;
                }
                ++hits;
                consider(wx, wy, wz, h[c], j0 + c, lo[c], hi[c], best);
            }
        }
        // the displaced keys: every unresolved key of the chunk one slot further per round
        for (int probe = 1; need != 0u && probe < PROBES; ++probe) {
            int4 lo[C];
            float4 hi[C];
            LW_UNROLL for (int c = 0; c < C; ++c) load_slot(t.slots, (h[c] + probe) & t.mask, lo[c], hi[c]);
            LW_UNROLL for (int c = 0; c < C; ++c) {
                if (!((need >> c) & 1u)) continue;
                if (key_of(lo[c]) == key[c]) {
                    need &= ~(1u << c);
                    ++hits;
                    consider(wx, wy, wz, (h[c] + probe) & t.mask, j0 + c, lo[c], hi[c], best);
                } else if (probe >= dmax[c]) {
                    need &= ~(1u << c);  // past the bound: not in the map
                }
            }
        }
    }
}

}  // namespace lw
