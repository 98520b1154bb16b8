// kernels.h -- host-callable launchers of the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rendertoy_amd.h"

namespace rtamd {

struct DevScene;

struct AovDev {
    uint32_t *id;
    float *t;
    float *normal;
    float *depth;
    float *albedo;
    float *mask;
};

hipError_t launch_reset(uint32_t *p, uint32_t n, hipStream_t s, uint32_t *q = nullptr, uint32_t m = 0);
int aov_blocks_per_cu(bool count);
// counters / stat: this launch's band counters (8 x 32 words, zero) and
// statistics (8 u64, zero); next_counters / next_stat are cleared by this
// launch for the next one.
hipError_t launch_aov(const DevScene &sc, const rt_camera &cam, const AovDev &out, uint32_t passes,
                      uint32_t *counters, uint32_t *next_counters, uint32_t grid, uint32_t *ovf_ref, float *ovf_t,
                      int64_t ovf_stride, unsigned long long *stat, unsigned long long *next_stat, bool count,
                      hipStream_t s);

hipError_t launch_rays(const DevScene &sc, const float *src, const float *dir, const uint32_t *exclude,
                       float4 *out, uint32_t n, uint32_t *fetch, uint32_t grid, uint32_t *ovf_ref, float *ovf_t,
                       int64_t ovf_stride, hipStream_t s);

}  // namespace rtamd
