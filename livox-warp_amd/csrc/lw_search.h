// The closest-mean neighbour search shared by the odometry accumulation (lw_icp.hip) and the batched
// localisation (lw_loc.hip): odom.py:308-334 with find_slot (gpu.py:254-265).
//
// A lane resolves a chunk of C neighbour keys together. Keys and hashes come from the point's own voxel key
// by addition: neighbour j's key is key0 + off(j), off(j) = dx + dy 2^21 + dz 2^42, and its hash product is
// key0 * HASH_MUL + off(j) * HASH_MUL (mod 2^64), both offsets read from a small LDS table (NbrTable); that
// holds whenever the point's voxel indices lie at least R inside the 21-bit key range, and a wave with a
// point outside it computes every key and hash directly (voxel_key_ijk, slot_hash), so the keys are the
// reference's in every case. Round 0 loads every key's home slot record (slot_hash & mask; 32 B: key, voxel
// id, count, mean, displacement bound) at once, with no branch around the loads, and the tests that follow
// are selects, not branches. A record holding the key is a hit. A key whose home slot holds another key can
// only lie within the home's displacement bound (Slot::disp, the largest probe distance of the keys homed
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

// A slot as two 16-byte loads: lo = {key lo, key hi, voxel id, count bits}, hi = {mean x, y, z, disp bits}. s is
// a slot index (0 <= s < 2^27): its byte offset fits 32 bits, so the load takes the table base as a scalar.
__device__ inline void load_slot(const Slot *slots, int32_t s, int4 &lo, float4 &hi) {
    const int4 *p = reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(slots) + ((uint32_t)s << 5));
    lo = p[0];
    hi = reinterpret_cast<const float4 *>(p)[1];
}

__device__ inline int64_t key_of(const int4 &lo) {
    return (int64_t)(((uint64_t)(uint32_t)lo.y << 32) | (uint64_t)(uint32_t)lo.x);
}

__device__ inline bool bit_of(uint32_t word, int32_t s) { return ((word >> (s & 31)) & 1u) != 0u; }

// The squared distance from the point to a mean, in the reference's order of float32 operations.
__device__ inline float dist2(float wx, float wy, float wz, float mx, float my, float mz) {
    const float d0 = wx - mx, d1 = wy - my, d2 = wz - mz;
    float dd = d0 * d0;
    dd = dd + d1 * d1;
    dd = dd + d2 * d2;
    return dd;
}

// A candidate of positive count (odom.py:324) at squared distance dd is taken when it is closer, or as close
// and earlier in the order, than the best so far (the serial loop starts from best_d2 = max_d2 and takes
// strictly smaller distances, odom.py:326-330). Without branches.
__device__ inline void take_if_closer(bool ok, float dd, int32_t s, int32_t j, float mx, float my, float mz,
                                      int32_t vid, Cand &best) {
    const bool closer = dd < best.d2 || (best.s >= 0 && dd == best.d2 && j < best.j);
    const bool take = ok && closer;
    best.d2 = take ? dd : best.d2;
    best.s = take ? s : best.s;
    best.j = take ? j : best.j;
    best.mx = take ? mx : best.mx;
    best.my = take ? my : best.my;
    best.mz = take ? mz : best.mz;
    best.vid = take ? vid : best.vid;
}

// One candidate record: taken when ok (the key was found), its count is positive and it is the closer one.
__device__ inline void consider(bool ok, float wx, float wy, float wz, int32_t s, int32_t j, const int4 &lo,
                                const float4 &hi, Cand &best) {
    take_if_closer(ok && __int_as_float(lo.w) > 0.0f, dist2(wx, wy, wz, hi.x, hi.y, hi.z), s, j, hi.x, hi.y, hi.z,
                   lo.z, best);
}

// What a search can hand to a recorder: every key of round 0, found (ok) or not, and then each key a probe finds,
// as slot s at position j with its record.
struct NoRec {
    __device__ void add(bool, int32_t, int32_t, const int4 &, const float4 &) {}
};

// Key and hash-product offsets of the neighbours of either reach (dx outer, dz inner), in LDS: rows
// [0, NBR_ROWS_1) for reach 1, then [NBR_ROWS_1, NBR_ROWS_1 + NBR_ROWS_2) for reach 2; past (2R+1)^3 the rows
// hold the offsets the same formula gives (the lookups of a padded chunk are made and then ignored).
struct NbrOff {
    uint64_t mul;  // off * HASH_MUL (mod 2^64)
    uint64_t key;  // off = dx + dy 2^21 + dz 2^42 (two's complement)
};
constexpr int NBR_ROWS_1 = 48, NBR_ROWS_2 = 144, NBR_ROWS = NBR_ROWS_1 + NBR_ROWS_2;

// Every lane of the workgroup takes part; the caller synchronises before the first search.
__device__ inline void nbr_fill(NbrOff *tab) {
    for (int r = threadIdx.x; r < NBR_ROWS; r += blockDim.x) {
        const int R = r < NBR_ROWS_1 ? 1 : 2, D = 2 * R + 1, j = r < NBR_ROWS_1 ? r : r - NBR_ROWS_1;
        const int dx = j / (D * D) - R, dy = (j / D) % D - R, dz = j % D - R;
        const uint64_t off = (uint64_t)(int64_t)dx + ((uint64_t)(int64_t)dy << KEY_BITS) +
                             ((uint64_t)(int64_t)dz << (2 * KEY_BITS));
        tab[r] = NbrOff{off * HASH_MUL, off};
    }
}

// The closest mean over neighbours [j_lo, j_hi) of the (2R+1)^3 block around the point's voxel (dx outer, dz
// inner), in chunks of C. hits counts the occupied neighbours found. tab: the LDS offset table (nbr_fill).
template <int R, int C, bool BITMAP, class Rec = NoRec>
__device__ inline void search_part(const Table &t, const NbrOff *tab, float inv_v, float wx, float wy, float wz,
                                   int j_lo, int j_hi, Cand &best, int32_t &hits, Rec rec = Rec{}) {
    constexpr int D = 2 * R + 1;
    constexpr int ROW0 = R == 1 ? 0 : NBR_ROWS_1;
    static_assert(C <= 32, "one bit per key of a chunk");
    static_assert(D * D * D + C - 1 <= (R == 1 ? NBR_ROWS_1 : NBR_ROWS_2), "a chunk starting below (2R+1)^3 stays in the table");
    const int32_t bx = ifloor(wx * inv_v), by = ifloor(wy * inv_v), bz = ifloor(wz * inv_v);
    // inside the key range by at least R on every axis: keys and hashes by addition
    constexpr int32_t LO = R - (int32_t)KEY_BIAS, HI = (int32_t)KEY_BIAS - 1 - R;
    const bool inside = bx >= LO && bx <= HI && by >= LO && by <= HI && bz >= LO && bz <= HI;
    const bool fast = __ballot(!inside) == 0ull;
    const uint64_t key0 = (uint64_t)voxel_key_ijk(bx, by, bz);
    const uint64_t mul0 = key0 * HASH_MUL;
    for (int j0 = j_lo; j0 < j_hi; j0 += C) {
        int64_t key[C];
        int32_t h[C];
        if (fast) {
            LW_UNROLL for (int c = 0; c < C; ++c) {
                const NbrOff o = tab[ROW0 + j0 + c];
                key[c] = (int64_t)(key0 + o.key);
                const uint64_t p = mul0 + o.mul;
                h[c] = (int32_t)(((uint32_t)p ^ (uint32_t)(p >> 29)) & 0x7FFFFFFFu) & t.mask;
            }
        } else {
            LW_UNROLL for (int c = 0; c < C; ++c) {
                const int j = j0 + c;
                const int dx = j / (D * D) - R, dy = (j / D) % D - R, dz = j % D - R;
                key[c] = voxel_key_ijk(bx + dx, by + dy, bz + dz);
                h[c] = slot_hash(key[c]) & t.mask;
            }
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
                const bool in = ((live >> c) & 1u) != 0u;
                const bool same = key_of(lo[c]) == key[c];
                dmax[c] = __float_as_int(hi[c].w);  // the home's displacement bound
                need |= (in && !same && dmax[c] > 0 ? 1u : 0u) << c;
                hits += in && same ? 1 : 0;
                consider(in && same, wx, wy, wz, h[c], j0 + c, lo[c], hi[c], best);
                rec.add(in && same, h[c], j0 + c, lo[c], hi[c]);
            }
        }
        // the displaced keys: every unresolved key of the chunk one slot further per round
        for (int probe = 1; need != 0u && probe < PROBES; ++probe) {
            int4 lo[C];
            float4 hi[C];
            LW_UNROLL for (int c = 0; c < C; ++c) load_slot(t.slots, (h[c] + probe) & t.mask, lo[c], hi[c]);
            LW_UNROLL for (int c = 0; c < C; ++c) {
                const bool in = ((need >> c) & 1u) != 0u;
                const bool same = key_of(lo[c]) == key[c];
                hits += in && same ? 1 : 0;
                consider(in && same, wx, wy, wz, (h[c] + probe) & t.mask, j0 + c, lo[c], hi[c], best);
                if (in && same) rec.add(true, (h[c] + probe) & t.mask, j0 + c, lo[c], hi[c]);
                // found, or past the bound: not in the map
                need &= ~((in && (same || probe >= dmax[c]) ? 1u : 0u) << c);
            }
        }
    }
}

}  // namespace lw
