// bvh_build.cpp -- host construction of the BVH4 the traversal kernels walk.
//
// Replaces the reference's octree (src/bvh.cpp:106-264), whose node boxes are
// all seeded at the world origin and whose traversal is a heap-ordered
// best-first search.  Here: a binned-SAH binary tree (32 bins per axis,
// leaves of at most kMaxLeaf triangles) collapsed into 4-wide nodes of one
// 128-byte line each, nodes laid out depth-first so a subtree is contiguous.
//
// Closest-hit results do not depend on the tree: traversal keeps the minimum
// (t, triangle index) over every triangle whose box the ray reaches, which is
// the reference's brute-force order (bvh.cpp:268-286).  Boxes are built from
// v0, v0+e1, v0+e2 -- the triangle Moller-Trumbore actually tests -- rounded
// outward and padded so the culling never drops a hit.
#include "bvh_build.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <stdexcept>

namespace rtamd {
namespace {

struct Aabb {
    float lo[3], hi[3];
    void reset() {
        for (int k = 0; k < 3; ++k) { lo[k] = std::numeric_limits<float>::infinity(); hi[k] = -lo[k]; }
    }
    void grow(const Aabb &b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow_pt(const float *p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {
    Aabb box;
    int left = -1, right = -1;
    uint32_t start = 0, count = 0;
    bool leaf() const { return left < 0; }
};

constexpr int kBins = 32;

struct Builder {
    const std::vector<Aabb> &boxes;
    std::vector<float> cen;          // 3 per triangle
    std::vector<uint32_t> &idx;
    std::vector<BNode> nodes;
    std::mutex mu;

    Builder(const std::vector<Aabb> &b, std::vector<uint32_t> &order) : boxes(b), idx(order) {}

    int alloc() {
        std::lock_guard<std::mutex> g(mu);
        nodes.emplace_back();
        return (int)nodes.size() - 1;
    }

    // Builds the subtree over idx[begin, end); returns its node index.
    int build(uint32_t begin, uint32_t end, int depth) {
        int me = alloc();
        Aabb box, cbox;
        box.reset();
        cbox.reset();
        for (uint32_t i = begin; i < end; ++i) {
            box.grow(boxes[idx[i]]);
            cbox.grow_pt(&cen[3 * idx[i]]);
        }
        uint32_t n = end - begin;
        auto make_leaf = [&]() {
            std::lock_guard<std::mutex> g(mu);
            nodes[me].box = box;
            nodes[me].start = begin;
            nodes[me].count = n;
            return me;
        };
        if (n <= 2) {
            if (n == 1) return make_leaf();
        }
        // binned SAH over the centroid box
        int best_axis = -1, best_bin = -1;
        float best_cost = std::numeric_limits<float>::infinity();
        for (int ax = 0; ax < 3; ++ax) {
            float ext = cbox.hi[ax] - cbox.lo[ax];
            if (!(ext > 0.0f)) continue;
            float scale = kBins / ext;
            Aabb bb[kBins];
            uint32_t bc[kBins] = {};
            for (int b = 0; b < kBins; ++b) bb[b].reset();
            for (uint32_t i = begin; i < end; ++i) {
                int b = std::min(kBins - 1, (int)((cen[3 * idx[i] + ax] - cbox.lo[ax]) * scale));
                bb[b].grow(boxes[idx[i]]);
                bc[b]++;
            }
            float ra[kBins];
            uint32_t rc[kBins];
            Aabb acc;
            acc.reset();
            uint32_t cnt = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]);
                cnt += bc[b];
                ra[b] = acc.area();
                rc[b] = cnt;
            }
            acc.reset();
            cnt = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.grow(bb[b]);
                cnt += bc[b];
                if (cnt == 0 || rc[b + 1] == 0) continue;
                float cost = acc.area() * cnt + ra[b + 1] * rc[b + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = ax; best_bin = b; }
            }
        }
        float pa = box.area();
        float leaf_cost = (float)n;
        float split_cost = pa > 0.0f ? 1.0f + best_cost / pa : std::numeric_limits<float>::infinity();
        if (best_axis < 0 || (n <= 4 && leaf_cost <= split_cost)) {
            if (n <= kMaxLeaf) return make_leaf();
        }
        uint32_t mid;
        if (best_axis >= 0) {
            float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
            float scale = kBins / ext;
            float lo = cbox.lo[best_axis];
            auto it = std::partition(idx.begin() + begin, idx.begin() + end, [&](uint32_t t) {
                int b = std::min(kBins - 1, (int)((cen[3 * t + best_axis] - lo) * scale));
                return b <= best_bin;
            });
            mid = (uint32_t)(it - idx.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        } else {
            // all centroids coincide: split by count along the longest axis of the box
            mid = begin + n / 2;
        }
        int l, r;
        if (n > 65536 && depth < 6) {
            auto fut = std::async(std::launch::async, [&]() { return build(begin, mid, depth + 1); });
            r = build(mid, end, depth + 1);
            l = fut.get();
        } else {
            l = build(begin, mid, depth + 1);
            r = build(mid, end, depth + 1);
        }
        std::lock_guard<std::mutex> g(mu);
        nodes[me].box = box;
        nodes[me].left = l;
        nodes[me].right = r;
        return me;
    }
};

// Outward rounding of a double to float.
inline float down(double x) {
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -std::numeric_limits<float>::infinity());
    return f;
}
inline float up(double x) {
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

struct Collapser {
    const std::vector<BNode> &b2;
    std::vector<Node4> &out;
    uint32_t max_depth = 0;

    // Returns (node index, stack bound of its subtree).
    std::pair<uint32_t, uint32_t> collapse(int root, uint32_t depth) {
        int kids[4];
        int nk = 0;
        if (b2[root].leaf()) {
            kids[nk++] = root;
        } else {
            kids[nk++] = b2[root].left;
            kids[nk++] = b2[root].right;
            while (nk < 4) {
                int best = -1;
                float ba = -1.0f;
                for (int i = 0; i < nk; ++i)
                    if (!b2[kids[i]].leaf() && b2[kids[i]].box.area() > ba) { ba = b2[kids[i]].box.area(); best = i; }
                if (best < 0) break;
                int c = kids[best];
                kids[best] = b2[c].left;
                kids[nk++] = b2[c].right;
            }
        }
        uint32_t me = (uint32_t)out.size();
        out.emplace_back();
        max_depth = std::max(max_depth, depth + 1);
        Node4 nd;
        std::memset(&nd, 0, sizeof(nd));
        uint32_t sub = 0;
        for (int j = 0; j < 4; ++j) {
            if (j >= nk) {
                nd.child[j] = kEmptyRef;
                nd.lox[j] = nd.loy[j] = nd.loz[j] = 1.0f;
                nd.hix[j] = nd.hiy[j] = nd.hiz[j] = -1.0f;
                continue;
            }
            const BNode &c = b2[kids[j]];
            float pad[3];
            for (int k = 0; k < 3; ++k) {
                float m = std::max(std::fabs(c.box.lo[k]), std::fabs(c.box.hi[k]));
                pad[k] = m * 0x1p-19f + 1e-30f;
            }
            nd.lox[j] = c.box.lo[0] - pad[0];
            nd.loy[j] = c.box.lo[1] - pad[1];
            nd.loz[j] = c.box.lo[2] - pad[2];
            nd.hix[j] = c.box.hi[0] + pad[0];
            nd.hiy[j] = c.box.hi[1] + pad[1];
            nd.hiz[j] = c.box.hi[2] + pad[2];
            if (c.leaf()) {
                nd.child[j] = make_leaf(c.start, c.count);
            } else {
                auto r = collapse(kids[j], depth + 1);
                nd.child[j] = r.first;
                sub = std::max(sub, r.second);
            }
        }
        out[me] = nd;
        return {me, (uint32_t)(nk - 1) + sub};
    }
};

}  // namespace

Bvh4 build_bvh4(uint32_t n, const float *v0e1e2) {
    auto t0 = std::chrono::steady_clock::now();
    Bvh4 res;
    if (n == 0) {
        Node4 nd;
        std::memset(&nd, 0, sizeof(nd));
        for (int j = 0; j < 4; ++j) {
            nd.child[j] = kEmptyRef;
            nd.lox[j] = nd.loy[j] = nd.loz[j] = 1.0f;
            nd.hix[j] = nd.hiy[j] = nd.hiz[j] = -1.0f;
        }
        res.nodes.push_back(nd);
        res.depth = 1;
        res.max_stack = 0;
        return res;
    }
    if (n >= (1u << 28)) throw std::runtime_error("scene exceeds 2^28 triangles");
    std::vector<Aabb> boxes(n);
    res.order.resize(n);
    std::vector<uint32_t> &order = res.order;
    for (uint32_t i = 0; i < n; ++i) {
        const float *p = v0e1e2 + 9 * (size_t)i;
        double a[3][3];
        for (int k = 0; k < 3; ++k) {
            a[0][k] = p[k];
            a[1][k] = (double)p[k] + (double)p[3 + k];
            a[2][k] = (double)p[k] + (double)p[6 + k];
        }
        for (int k = 0; k < 3; ++k) {
            double lo = std::min(a[0][k], std::min(a[1][k], a[2][k]));
            double hi = std::max(a[0][k], std::max(a[1][k], a[2][k]));
            boxes[i].lo[k] = down(lo);
            boxes[i].hi[k] = up(hi);
        }
        order[i] = i;
    }
    Builder b(boxes, order);
    b.cen.resize(3 * (size_t)n);
    for (uint32_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) b.cen[3 * i + k] = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
    b.nodes.reserve(2 * (size_t)n);
    int root = b.build(0, n, 0);
    if (root != 0) throw std::runtime_error("bvh root is not node 0");
    Collapser c{b.nodes, res.nodes};
    res.nodes.reserve(b.nodes.size() / 2 + 1);
    auto r = c.collapse(0, 0);
    res.depth = c.max_depth;
    res.max_stack = r.second;
    res.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return res;
}

}  // namespace rtamd
