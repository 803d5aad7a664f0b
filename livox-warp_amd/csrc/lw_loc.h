// Launch description of the batched localisation kernel (lw_loc.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct lw_ctx;

namespace lw {

struct LocLaunch {
    const float *scans;       // all query scans, float32 x 3, concatenated (device)
    const int32_t *scan_off;  // Q + 1 offsets into scans (device)
    const int32_t *order;     // the Q queries in processing order, longest first (device)
    const double *hyp_T;      // Q * H hypotheses, 4x4 float64 row-major (device)
    double *out_T;            // Q * H refined poses (device)
    double *out_score;        // Q * H: sum of robust weights of the last pass (device)
    int32_t *out_count;       // Q * H: matched points of the last pass (device)
    int32_t *out_iters;       // Q * H: steps taken (device)
    unsigned long long *prof; // measurement counters, or null (part of a captured launch's identity)
    int32_t Q, H, iters, min_count, block, grid, chunk;
    float inv_v, max_d2;
    double damping;
    float kappa[16];
};

hipError_t launch_loc(lw_ctx *c, const LocLaunch &L);
// The compiled kernel of a tile (block 64 / 128 / 256 / 512, chunk 7 / 14): registers, scratch, LDS.
hipError_t loc_tile_attributes(int32_t block, int32_t chunk, hipFuncAttributes *fa);

}  // namespace lw
