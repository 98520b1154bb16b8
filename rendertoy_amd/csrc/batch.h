// batch.h -- the opaque rt_batch behind the C ABI: many scenes path-traced
// through one wavefront (one set of bounce-loop launches, one hipGraph).
#pragma once
#include <vector>

#include "pt_kernels.h"
#include "scene.h"

struct rt_batch {
    rt_scene *ctx = nullptr;                 // render context: stream, events, overflow region, path state, stats
    std::vector<rt_scene *> scenes;          // members, not owned
    std::vector<rtamd::BatchScene> h_table;  // geometry handles and cameras, uploaded before every frame
    std::vector<uint32_t> h_off;             // pixel offsets (n + 1 entries), then primary-tile offsets (n + 1)
    rtamd::BatchScene *d_table = nullptr;
    uint32_t *d_off = nullptr;
    uint32_t total_pix = 0, total_tiles = 0;
    bool any_emitters = false;
    bool detached = false;                   // a member scene was destroyed: only rt_batch_destroy is allowed
    ~rt_batch();
};
