// mlp.hpp - engine-side entry points of the learned cut proposal (mlp.hip).
#pragma once
#include "ctx.hpp"

namespace cuacd {

constexpr int MLP_SAMPLES = 256;  // surface samples per part

// A part mesh on the device and its box (the sampling frame).
struct MlpSamplePart {
    const float* v;
    const int* t;
    int nv, nt;
    float lo[3], hi[3];
};

// d_out [P * 256 * 3]: area-stratified samples of each part in its box frame
// (center, half the largest extent). Stream-ordered.
int mlp_sample_dev(gpu_ctx_t ctx, const MlpSamplePart* parts, int P, float* d_out);

// Predicted scores d_out [C] of candidate planes d_planes [C x 4] (unit
// normal and offset in their part's frame; part_of ascending). Stream-ordered.
int mlp_forward_dev(gpu_ctx_t ctx, gpu_mlp* mp, const float* d_pts, int P, const float* d_planes, const int* d_part,
                    int C, float* d_out);

// Tuning knob 6 (gpu_set_tuning): inference through the fused point-layer
// kernel (1) or the layer-by-layer GEMMs (0); the scores are the same.
int mlp_fused_on();

}  // namespace cuacd
