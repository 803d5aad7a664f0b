// The integration map's records, view and launchers (lw_imap.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lw {

constexpr int IMAP_POSE_SLOTS = 1 << 16;  // gpu.py:56 (POSE_SLOTS)
constexpr int IMAP_COUNT_CAP = 4096;      // gpu.py:55 (MAP_COUNT_CAP)
// gpu.py:67-73: where a ray stops short of its return, and what counts as piercing a voxel
constexpr float IMAP_CARVE_MIN_RANGE = 0.05f;
constexpr float IMAP_CARVE_MARGIN_VOXELS = 3.0f;
constexpr float IMAP_CARVE_MARGIN_FRACTION = 0.03f;
constexpr float IMAP_CARVE_GRAZING_VOXELS = 1.5f;
constexpr float IMAP_CARVE_MIN_COSINE = 0.05f;
constexpr float IMAP_CARVE_PIERCE_PLANE = 0.6f;
constexpr float IMAP_CARVE_PIERCE_POINT = 0.5f;

// One slot of the integration map (64 B): the voxel key (-1: empty), hits, carving misses, the float32 sums of
// x, y, z and reflectivity over the first IMAP_COUNT_CAP hits, and the first and last time it was hit.
struct __align__(64) MapRec {
    int64_t key;
    int32_t cnt;
    int32_t miss;
    float4 acc;
    float first;
    float last;
    int32_t pad[6];
};
static_assert(sizeof(MapRec) == 64, "one record per 64 bytes");

// The raw-return filters an insert applies (the keep_raw fields of the reference's View, gpu.py:77-99).
struct MapView {
    float min_range, max_range;
    int32_t min_refl, noise_mask, ret_mask;
};

// The brick bits a map of cap slots keeps (a power of two, at least 1024).
int32_t imap_brick_bits(int64_t cap);
hipError_t imap_clear(hipStream_t s, MapRec *recs, int64_t cap, int32_t *stats, uint32_t *bricks);
hipError_t imap_brick_rebuild(hipStream_t s, const MapRec *recs, int64_t cap, uint32_t *bricks);
hipError_t imap_pose_scatter(hipStream_t s, const int32_t *idx, const float *mats, int32_t n, float *poses);
hipError_t imap_insert(hipStream_t s, const float *sx, const uint32_t *sa, const float *st, const int32_t *sf,
                       int32_t n, const float *poses, const MapView &v, float inv_voxel, int32_t mask, MapRec *recs,
                       int32_t *stats, uint32_t *bricks);
hipError_t imap_carve(hipStream_t s, const float *sx, const uint32_t *sa, const int32_t *sf, int32_t n, int32_t every,
                      const float *poses, float inv_voxel, int32_t mask, float max_range, int32_t max_steps,
                      MapRec *recs, const uint32_t *bricks);
hipError_t imap_rebase(hipStream_t s, MapRec *recs, int64_t cap, float shift);

}  // namespace lw
