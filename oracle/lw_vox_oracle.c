/*
 * CPU restatement of the reference's bulk voxelisation of a prior map -- TEST INFRASTRUCTURE ONLY.
 *
 * Tests and the smoke check call it as the checker; the product never links or calls it. Warp's CPU
 * backend runs a launch as a serial loop over thread ids, so every atomic below is a plain
 * read-modify-write in index order. Float32 throughout, each multiply and add rounded on its own (built
 * with -ffp-contract=off), in the order the reference writes them.
 *
 * Reference files (paths relative to the reference repository root):
 *   python/livox_warp/prior_map.py :31-36 key layout and MAX_PROBES (128), :39-47 voxel_key_unpack,
 *                                   :50-52 voxel_centre, :55-91 k_vox_insert, :94-97 k_vox_flag,
 *                                   :100-123 k_vox_emit, :169-203 voxel_downsample (flag, exclusive scan,
 *                                   emit: the voxels come out in the table's slot order)
 *   python/livox_warp/gpu.py       :231-265 voxel_key_ijk, slot_hash
 */
#include <math.h>
#include <stdint.h>

#define LWO_VOX_PROBES 128 /* prior_map.py:36 */
#define LWO_KEY_BITS 21
#define LWO_KEY_BIAS (1LL << 20)
#define LWO_KEY_MASK ((1LL << 21) - 1LL)

int64_t lwo_voxel_key_ijk(int32_t ix, int32_t iy, int32_t iz);
int32_t lwo_slot_hash(int64_t k);

/* prior_map.py:50-52 */
static void centre(int32_t ix, int32_t iy, int32_t iz, float voxel, float *c) {
    c[0] = ((float)ix + 0.5f) * voxel;
    c[1] = ((float)iy + 0.5f) * voxel;
    c[2] = ((float)iz + 0.5f) * voxel;
}

/* prior_map.py:55-91: each point's offset from its voxel's centre, colour and intensity added to the slot
 * claimed for its key (128 probes); stats[0] voxels claimed, stats[1] points dropped. rgb / inten may be
 * null (zeros). keys: cap int64 (-1 empty); rsum: cap x 3; csum: cap x 4; cnt: cap. */
void lwo_vox_insert(const float *xyz, const float *inten, const float *rgb, int64_t n, float inv_v, float voxel,
                    int64_t mask, int64_t *keys, float *rsum, float *csum, int32_t *cnt, int32_t *stats) {
    for (int64_t i = 0; i < n; ++i) {
        const float *p = xyz + 3 * i;
        const int32_t ix = (int32_t)floorf(p[0] * inv_v), iy = (int32_t)floorf(p[1] * inv_v),
                      iz = (int32_t)floorf(p[2] * inv_v);
        const int64_t key = lwo_voxel_key_ijk(ix, iy, iz);
        float c[3];
        centre(ix, iy, iz, voxel, c);
        const int64_t h = (int64_t)lwo_slot_hash(key) & mask;
        int done = 0;
        for (int probe = 0; probe < LWO_VOX_PROBES; ++probe) {
            const int64_t s = (h + probe) & mask;
            const int64_t prev = keys[s];
            if (prev == -1 || prev == key) {
                if (prev == -1) {
                    keys[s] = key;
                    stats[0] += 1;
                }
                for (int k = 0; k < 3; ++k) rsum[3 * s + k] += p[k] - c[k];
                for (int k = 0; k < 3; ++k) csum[4 * s + k] += rgb ? rgb[3 * i + k] : 0.0f;
                csum[4 * s + 3] += inten ? inten[i] : 0.0f;
                cnt[s] += 1;
                done = 1;
                break;
            }
        }
        if (!done) stats[1] += 1;
    }
}

/* prior_map.py:94-123 with voxel_downsample's exclusive scan (:192-201): the claimed slots in slot order as
 * mean position (centre + offset sum / count), mean colour and intensity, count. Returns the voxels written. */
int64_t lwo_vox_emit(const int64_t *keys, const float *rsum, const float *csum, const int32_t *cnt, int64_t cap,
                     float voxel, float *out_xyz, float *out_col, int32_t *out_cnt) {
    int64_t j = 0;
    for (int64_t s = 0; s < cap; ++s) {
        if (keys[s] == -1) continue;
        const int64_t k = keys[s];
        const int32_t ix = (int32_t)((k & LWO_KEY_MASK) - LWO_KEY_BIAS);
        const int32_t iy = (int32_t)(((k >> LWO_KEY_BITS) & LWO_KEY_MASK) - LWO_KEY_BIAS);
        const int32_t iz = (int32_t)(((k >> (2 * LWO_KEY_BITS)) & LWO_KEY_MASK) - LWO_KEY_BIAS);
        float c[3];
        centre(ix, iy, iz, voxel, c);
        const float fc = (float)cnt[s];
        for (int a = 0; a < 3; ++a) out_xyz[3 * j + a] = c[a] + rsum[3 * s + a] / fc;
        for (int a = 0; a < 4; ++a) out_col[4 * j + a] = csum[4 * s + a] / fc;
        out_cnt[j] = cnt[s];
        ++j;
    }
    return j;
}
