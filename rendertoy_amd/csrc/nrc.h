// nrc.h -- the opaque rt_cache behind the C ABI.
#pragma once
#include <vector>

#include "nrc_kernels.h"
#include "scene.h"

struct rt_cache {
    rt_scene *scene = nullptr;     // null once the scene is destroyed
    int device = 0;
    rt_cache_params prm{};
    rtamd::NrcGrid grid{};
    rtamd::NrcTrain tr{};
    std::vector<rtamd::DevBuf> bufs;
    float *samples = nullptr;      // training samples (capacity x 16 floats)
    uint32_t capacity = 0;         // samples the buffers hold
    uint32_t *count = nullptr;     // training-sample counter
    int step = 0;                  // Adam steps taken
    uint32_t frame = 0;            // training-path stream position
    ~rt_cache();
};
