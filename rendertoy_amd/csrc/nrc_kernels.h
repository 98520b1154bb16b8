// nrc_kernels.h -- radiance-cache device structures and launchers.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rendertoy_amd.h"

namespace rtamd {

struct DevScene;

constexpr int kNrcLevels = 16;
constexpr int kNrcParams = 5 * 4096 + 16 * 64;   // W0..W4 [64x64], W5 [16x64] (rows 3..15 zero)
constexpr int kNrcMaxV = 16;                     // vertices of a training path
// Query / sample record, 20 floats: position 3, outgoing direction 3, facing
// normal 3, roughness, albedo 3, bounces left after the vertex (path vertices
// the estimator may still add), 0, path slot (render queries: uint bits),
// target radiance 3 (training samples), 0.
constexpr int kNrcFloats = 20;
constexpr int kRecBounces = 13;
constexpr int kRecSlot = 15;
constexpr int kRecTarget = 16;
constexpr float kNrcBounceScale = 1.0f / 16.0f;  // input feature = bounces left / 16

struct NrcGrid {
    _Float16 *t16;          // [levels][T][2] fp16 features (forward)
    float *grad;            // [levels][T][2] fp32 gradient (scaled)
    uint32_t tmask;         // T - 1
    uint32_t side[kNrcLevels];
    uint32_t dense[kNrcLevels];
    float res[kNrcLevels];
    float bmin[3], binv[3];
    size_t grid_params;     // levels * T * 2
    // gradient shards of the dense levels (see nrc_grid_scatter_kernel):
    // kGradShards copies of [dense level l at doff[l]: side^3 x 2 floats]
    float *shard;
    uint32_t doff[kNrcLevels];
    uint32_t dense_floats;  // floats per shard
};

// Copies of the dense levels' gradient that the scatter spreads its atomics
// over (lane % kGradShards): coarse levels see dozens of atomics per entry per
// step, which serialise on one address.
constexpr uint32_t kGradShards = 16;
static_assert((kGradShards & (kGradShards - 1)) == 0, "kGradShards is a power of two");

struct NrcTrain {
    float *w, *m, *v;           // MLP fp32 master + Adam moments
    _Float16 *w16, *wt16;       // fp16 copies (W_j and W_j^T)
    float *grid, *grid_m, *grid_v;
    _Float16 *act, *dout;       // per-wave activations / loss gradient
    float *loss_part;           // per-wave loss sums
    float *wpart;               // per-wave weight-gradient partials
    float *red;                 // weight gradient reduced in 32 groups of waves
    float *dgrid;               // per-sample input gradients of the 32 grid features
    float *pred_out;            // optional: predictions of the last forward
};

struct NrcGen {
    rt_camera cam;
    float divW, divH;
    uint32_t npix, npaths, frame, max_depth, first_vertex, capacity;
    uint32_t key0, key1;
    float target_clamp;
};

hipError_t nrc_launch_infer(const NrcGrid &g, const _Float16 *w16, const float *queries, const uint32_t *count_ptr,
                            uint32_t count, float4 *out, float4 *rad, const float4 *thr, uint32_t grid, int mode,
                            hipStream_t s);
hipError_t nrc_launch_train_step(const NrcGrid &g, NrcTrain &t, const float *samples, uint32_t n, hipStream_t s);
hipError_t nrc_launch_adam(const NrcGrid &g, NrcTrain &t, uint32_t n, int step, float lr_mlp, float lr_grid,
                           bool update_grid, hipStream_t s);
hipError_t nrc_launch_sync_weights(const NrcGrid &g, NrcTrain &t, hipStream_t s);
hipError_t nrc_launch_zero(float *p, size_t n, hipStream_t s);
hipError_t nrc_launch_gen(const DevScene &sc, const NrcGen &gp, float *out, uint32_t *count, uint32_t grid,
                          uint32_t *ovf_ref, float *ovf_t, int64_t ovf_stride, hipStream_t s);
int nrc_gen_blocks_per_cu();

}  // namespace rtamd
