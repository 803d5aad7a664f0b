// Shapes of the learned pose-increment prior (lw_mlp.hip) and its per-frame input features.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lw {

constexpr int MLP_MOTION = 24;                  // the last 4 frame increments, 6 each
constexpr int MLP_IN = MLP_MOTION + 64 + 64;    // + range and elevation histograms = 152
constexpr int MLP_IN_PAD = 160;                 // K of the first layer, a multiple of 32
constexpr int MLP_HID = 256;
constexpr int MLP_OUT = 6;
constexpr int MLP_OUT_PAD = 16;

struct FeatArgs {
    float r_max;               // range histogram over [0, r_max)
    float e_max;               // elevation histogram over [-e_max, e_max) (radians)
    float motion[MLP_MOTION];  // the motion inputs, already scaled
};

// The 152 inputs of one frame into out (device): motion from the arguments, histograms of xyz[0..m). m_dev,
// when given, holds the frame's count on the device (taken up to cap) in place of m.
hipError_t launch_features(hipStream_t s, const float *xyz, int32_t m, const int32_t *m_dev, int32_t cap,
                           const FeatArgs &f, float *out);

}  // namespace lw
