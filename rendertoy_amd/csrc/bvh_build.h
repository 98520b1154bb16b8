// bvh_build.h -- host BVH4 construction (see bvh_build.cpp).
#pragma once
#include <mutex>
#include <vector>

#include "rt_layout.h"

namespace rtamd {

struct Bvh4 {
    std::vector<Node4> nodes;      // node 0 = root
    std::vector<uint32_t> order;   // leaf order -> triangle index
    uint32_t depth = 0;            // levels of Node4
    uint32_t max_stack = 0;        // deepest traversal stack any ray can need
    double ms = 0.0;
};

// v0e1e2: [n][9] = v0, e1, e2 per triangle (the cached Triangle fields).
Bvh4 build_bvh4(uint32_t n, const float *v0e1e2);

}  // namespace rtamd
