// Layout of the learned warm start's parameters and device buffers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "relax_device.h"

namespace tgh {

constexpr uint32_t kFeatureNeighbors = 16;
constexpr uint32_t kFeatureWidth = 64;  // 16 * 3 neighbour features, tangent, motion history, zeros
constexpr uint32_t kFeatTangent = 48;   // 3: unit tangent
constexpr uint32_t kFeatMotion = 51;    // 3: unit direction of the recent midpoint motion
constexpr uint32_t kFeatMotionLog = 54; // 1: log10(recent motion in diameters) / 8 + 1
constexpr float kMinMotion = 1.0e-12f;  // recent motion below this (in diameters) counts as none

// f32 parameter vector: W1 [64 out][64 in], b1 [64], W2 [64][64], b2 [64],
// W3 [16 out][64 in] (3 outputs used), b3 [16]; gradients use the same
// layout, followed by the loss in a partial slot.
constexpr uint32_t kOffW1 = 0;
constexpr uint32_t kOffB1 = kOffW1 + 64 * 64;
constexpr uint32_t kOffW2 = kOffB1 + 64;
constexpr uint32_t kOffB2 = kOffW2 + 64 * 64;
constexpr uint32_t kOffW3 = kOffB2 + 64;
constexpr uint32_t kOffB3 = kOffW3 + 16 * 64;
constexpr uint32_t kMlpParams = kOffB3 + 16;
constexpr uint32_t kOffLoss = kMlpParams;
constexpr uint32_t kMlpPartialStride = kMlpParams + 4;

struct MlpDevice {
    float *master;          // kMlpParams f32
    float *adam_m, *adam_v; // kMlpParams f32 each
    const float *b1, *b2, *b3;  // views into master
    uint16_t *w1, *w2, *w3;     // bf16 [out][in]
    uint16_t *w2n, *w3n;        // bf16 [in][out]
};

hipError_t launch_mfma_probe(const uint16_t *A, const uint16_t *B, float *D, hipStream_t s);
// Features of every segment; `history` (midpoints at the last mark, or
// null) gives the motion terms, and `scale` (may be null) receives each
// segment's motion length in diameters times its diameter (0: no motion).
hipError_t launch_features(const WorldDev &w, const float4 *history, uint16_t *X, float *scale, hipStream_t s);
// Midpoints; .w = diameter times `scale[s]` when given, else the diameter.
hipError_t launch_midpoints(const WorldDev &w, const float *scale, float4 *mid, hipStream_t s);
// Targets: (midpoint now - then) / then.w (zero rows when then.w is 0).
hipError_t launch_targets(const WorldDev &w, const float4 *mid_then, float4 *T, hipStream_t s);
hipError_t launch_mlp_forward(const MlpDevice &m, const uint16_t *X, float4 *Y, uint32_t rows, hipStream_t s);
hipError_t launch_mlp_gradients(const MlpDevice &m, const uint16_t *X, const float4 *T, const uint32_t *idx,
                                uint32_t rows, uint32_t slots, float *partial, float *grad, hipStream_t s);
hipError_t launch_mlp_adam(const MlpDevice &m, const float *grad, float lr, uint64_t step, hipStream_t s);
hipError_t launch_mlp_publish(const MlpDevice &m, hipStream_t s);
// Moves each vertex by `scale` times the mean of its segments' predicted
// displacements, each prediction in units of that segment's `seg_scale`.
hipError_t launch_apply_proposal(const WorldDev &w, const float4 *Y, const float *seg_scale, float4 *backup,
                                 float scale, hipStream_t s);
hipError_t launch_restore(const WorldDev &w, const float4 *backup, hipStream_t s);
// restores the vertices of the specimens with revert[specimen] != 0
hipError_t launch_restore_specimens(const WorldDev &w, const float4 *backup, const uint32_t *revert,
                                    hipStream_t s);

}  // namespace tgh
