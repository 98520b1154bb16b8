// scene.h -- the opaque rt_scene behind the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/rendertoy_amd.h"
#include "dev_trace.h"
#include "kernels.h"

namespace rtamd {

struct PtState;
void set_error(const std::string &msg);

// Device allocation owned by a scene; freed in ~rt_scene.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

}  // namespace rtamd

struct rt_scene {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<rtamd::DevBuf> owned;   // scene geometry
    rtamd::DevScene ds{};
    uint32_t num_tris = 0;
    uint32_t num_materials = 0;
    uint32_t bvh_nodes = 0, bvh_depth = 0, max_stack = 0;
    double build_ms = 0.0;
    float bbox[6] = {0, 0, 0, 0, 0, 0};   // scene bounds (min xyz, max xyz)
    bool counting = false;
    // work queue counters (8 x 128-byte lines) and statistics
    uint32_t *d_counters = nullptr;
    unsigned long long *d_stat = nullptr;
    // AOV launches: two sets of band counters and statistics; launch k uses set
    // k % 2 and clears the other for launch k + 1, so no reset launch is needed
    uint32_t *d_aov_ctr = nullptr;            // 2 x 256
    unsigned long long *d_aov_stat = nullptr; // 2 x 8
    uint32_t aov_parity = 0;
    // traversal stack overflow region
    uint32_t *d_ovf_ref = nullptr;
    float *d_ovf_t = nullptr;
    size_t ovf_bytes = 0;
    // staging for rt_intersect
    void *d_rays = nullptr;
    size_t ray_bytes = 0;
    // AOV outputs
    rtamd::AovDev aov{};
    uint32_t aov_w = 0, aov_h = 0;
    std::vector<rtamd::DevBuf> aov_bufs;
    // path tracer state
    rtamd::PtState *pt = nullptr;
    rt_stats stats{};
    // radiance caches built on this scene and batches that hold it (both are
    // detached when the scene goes first)
    std::vector<rt_cache *> caches;
    std::vector<rt_batch *> batches;

    ~rt_scene();
};

namespace rtamd {
// Called by rt_scene_destroy: the scene's radiance caches and batches stop referring to it.
void detach_caches(rt_scene *scene);
void detach_batches(rt_scene *scene);
// A render context without geometry (stream, events, path state) on a device.
rt_scene *make_render_context(int device);
}  // namespace rtamd
