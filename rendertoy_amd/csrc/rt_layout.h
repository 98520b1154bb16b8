// rt_layout.h -- device data layout of a flattened RenderToy World.
//
// HBM layout (all arrays 16-byte aligned, indexed as noted):
//   nodes   Node4[num_nodes]       128 B each = one L2 line; node 0 is the root.
//   tris    float4[3*num_tris]     48 B per triangle in BVH leaf order:
//                                   {v0.xyz, e1.x} {e1.yz, e2.xy} {e2.z, id, -, -}
//                                   v0/e1/e2 exactly as Triangle::UpdateCache
//                                   computes them (src/object.cpp:224-243).
//   geo     float4[num_tris]       by triangle index: cached normal.xyz, area
//   info    uint32[num_tris]       by triangle index: material | smooth << 31
//   vnorm   float[9*num_tris]      by triangle index: Triangle::norm (smooth meshes only)
//   mats    rt_material[num_mats]  96 B each
//   emit    EmitterRec[num_emit]   64 B each, World::emissive_triangles order
#pragma once
#include <stdint.h>

namespace rtamd {

constexpr uint32_t kNone = 0xFFFFFFFFu;     // "no triangle"
constexpr uint32_t kEmptyRef = 0x7FFFFFFFu; // unused child slot
constexpr uint32_t kLeafBit = 0x80000000u;  // child ref: leaf = kLeafBit | start << 3 | (count-1)
constexpr uint32_t kMaxLeaf = 8;
constexpr float kFloatInfinity = 1e7f;      // include/RenderToy/rtmath.h:23

// Four child boxes in SoA form plus four child references.
struct alignas(128) Node4 {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

struct alignas(16) EmitterRec {
    float v0[3], e1[3], e2[3];
    float normal[3];
    float area;
    uint32_t tri;
    uint32_t pad[2];
};
static_assert(sizeof(EmitterRec) == 64, "EmitterRec is 64 bytes");

inline uint32_t make_leaf(uint32_t start, uint32_t count) { return kLeafBit | (start << 3) | (count - 1); }

}  // namespace rtamd
