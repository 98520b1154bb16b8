// aov.hip -- primary-ray AOV passes in one persistent launch.
//
// Replaces the per-pixel loops of NormalRenderer::Render
// (src/renderer.cpp:186-229), DepthBufferRenderer::Render (:138-174),
// AlbedoRenderer::Render (:391-427) and IntersectTestRenderer::Render (:20-54).
// Each wave owns an 8x8 pixel tile at a time (coherent primary rays share BVH
// nodes).  Tiles are cut into 8 contiguous bands, band b drained first by the
// workgroups on XCD b (round-robin dispatch puts workgroup g on XCD g % 8)
// through one counter per band, then helped by the others: one atomic per
// tile and no static dealing, which measured faster than BandSched for these
// 32k coherent tiles per 1080p frame.
#include "dev_trace.h"
#include "kernels.h"

namespace rtamd {

template <bool COUNT>
RT_TRACE_KERNEL void aov_kernel(DevScene sc, rt_camera cam, AovDev out, uint32_t passes, uint32_t *counters,
                                uint32_t *next_counters, uint32_t tiles_x, uint32_t ntiles, uint32_t *ovf_ref,
                                float *ovf_t, int64_t ovf_stride, unsigned long long *stat,
                                unsigned long long *next_stat) {
    __shared__ uint64_t s_stk[kLdsStack * 256];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + tid;
    if (blockIdx.x == 0) {   // the next launch's counters and statistics (this launch uses the other set)
        next_counters[tid] = 0u;
        if (tid < 8u) next_stat[tid] = 0ull;
    }
    Stack stk(s_stk + tid, (int)blockDim.x, ovf_ref + gtid, ovf_t + gtid, ovf_stride);
    const float divW = 1.0f / float(cam.width);
    const float divH = 1.0f / float(cam.height);
    const float div_fmn = 1.0f / (cam.far_clip - cam.near_clip);
    uint32_t n_nodes = 0, n_tris = 0;
    const uint32_t xcd = blockIdx.x & 7u;
    for (uint32_t k = 0; k < 8; ++k) {
        const uint32_t band = (xcd + k) & 7u;
        const uint32_t b0 = band_begin(ntiles, band), b1 = band_begin(ntiles, band + 1);
        while (true) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(counters + band * 32u, 1u);
            t = __shfl(t, 0);
            if (t >= b1 - b0) break;
            const uint32_t tile = b0 + t;
            const uint32_t x = (tile % tiles_x) * 8u + (lane & 7u);
            const uint32_t y = (tile / tiles_x) * 8u + (lane >> 3);
            if (x >= cam.width || y >= cam.height) continue;
            V3 o, d;
            camera_ray(cam, divW, divH, x, y, o, d);
            Hit h;
            trace<false, COUNT>(sc, o, d, kNone, kNone, 0.0f, stk, h, n_nodes, n_tris);
            const size_t p = (size_t)y * cam.width + x;
            const bool hit = h.id != kNone;
            if (passes & RT_AOV_ID) out.id[p] = h.id;
            if (passes & RT_AOV_T) out.t[p] = hit ? h.t : kFloatInfinity;
            if (passes & RT_AOV_MASK) out.mask[p] = hit ? 1.0f : 0.0f;
            if (passes & RT_AOV_DEPTH) {
                float dv = (h.t - cam.near_clip) * div_fmn;
                dv = (dv < 0.0f) ? 0.0f : ((1.0f < dv) ? 1.0f : dv);
                out.depth[p] = hit ? dv : 1.0f;
            }
            if (passes & RT_AOV_NORMAL) {
                V3 r = v3(0.0f);
                if (hit) {
                    V3 n = shading_normal(sc, h.id, h.u, h.v);
                    float4 g = sc.geo[h.id];
                    if (dot(V3{g.x, g.y, g.z}, neg(d)) < 0.0f) n = neg(n);
                    r = divs(add(n, v3(1.0f)), 2.0f);
                }
                out.normal[3 * p + 0] = r.x;
                out.normal[3 * p + 1] = r.y;
                out.normal[3 * p + 2] = r.z;
            }
            if (passes & RT_AOV_ALBEDO) {
                V3 r = v3(0.0f);
                if (hit) {
                    const rt_material &m = sc.mats[sc.info[h.id] & 0x7FFFFFFFu];
                    r = V3{m.base_color[0], m.base_color[1], m.base_color[2]};
                }
                out.albedo[3 * p + 0] = r.x;
                out.albedo[3 * p + 1] = r.y;
                out.albedo[3 * p + 2] = r.z;
            }
        }
    }
    if (COUNT) {
        atomicAdd(stat + 0, (unsigned long long)n_nodes);
        atomicAdd(stat + 1, (unsigned long long)n_tris);
    }
}

// BVH::Intersect over an explicit ray batch; a wave's work item is 64 rays.
RT_TRACE_KERNEL void ray_kernel(DevScene sc, const float *src, const float *dir,
                                                  const uint32_t *exclude, float4 *out, uint32_t n,
                                                  uint32_t *fetch, uint32_t *ovf_ref, float *ovf_t,
                                                  int64_t ovf_stride) {
    __shared__ uint64_t s_stk[kLdsStack * 256];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + tid;
    Stack stk(s_stk + tid, (int)blockDim.x, ovf_ref + gtid, ovf_t + gtid, ovf_stride);
    uint32_t nn = 0, nt = 0;
    BandSched bs = band_sched((n + 63u) / 64u, fetch, sc.share_rays);
    while (true) {
        const uint32_t chunk = band_next(bs);
        if (chunk == kNone) break;
        const uint32_t i = chunk * 64u + lane;
        if (i >= n) continue;
        const V3 o = {src[3 * (size_t)i], src[3 * (size_t)i + 1], src[3 * (size_t)i + 2]};
        const V3 d = {dir[3 * (size_t)i], dir[3 * (size_t)i + 1], dir[3 * (size_t)i + 2]};
        Hit h;
        trace<false, false>(sc, o, d, exclude ? exclude[i] : kNone, kNone, 0.0f, stk, h, nn, nt);
        out[i] = make_float4(h.t, h.u, h.v, __uint_as_float(h.id));
    }
}

hipError_t launch_rays(const DevScene &sc, const float *src, const float *dir, const uint32_t *exclude,
                       float4 *out, uint32_t n, uint32_t *fetch, uint32_t grid, uint32_t *ovf_ref, float *ovf_t,
                       int64_t ovf_stride, hipStream_t s) {
    hipLaunchKernelGGL(ray_kernel, dim3(grid), dim3(256), 0, s, sc, src, dir, exclude, out, n, fetch, ovf_ref,
                       ovf_t, ovf_stride);
    return hipGetLastError();
}

// Zeroes p[0, n) and, if q is not null, q[0, m): one launch for both.
__global__ void reset_u32_kernel(uint32_t *p, uint32_t n, uint32_t *q, uint32_t m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
    if (q && i < m) q[i] = 0u;
}

hipError_t launch_reset(uint32_t *p, uint32_t n, hipStream_t s, uint32_t *q, uint32_t m) {
    const uint32_t k = n > m ? n : m;
    hipLaunchKernelGGL(reset_u32_kernel, dim3((k + 255) / 256), dim3(256), 0, s, p, n, q, m);
    return hipGetLastError();
}

// Occupancy of a kernel does not change within a process (one architecture): asked once.
int aov_blocks_per_cu(bool count) {
    static const int nb_count = [] {
        int nb = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, aov_kernel<true>, 256, 0);
        return nb > 0 ? nb : 1;
    }();
    static const int nb_plain = [] {
        int nb = 0;
        (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, aov_kernel<false>, 256, 0);
        return nb > 0 ? nb : 1;
    }();
    return count ? nb_count : nb_plain;
}

hipError_t launch_aov(const DevScene &sc, const rt_camera &cam, const AovDev &out, uint32_t passes,
                      uint32_t *counters, uint32_t *next_counters, uint32_t grid, uint32_t *ovf_ref, float *ovf_t,
                      int64_t ovf_stride, unsigned long long *stat, unsigned long long *next_stat, bool count,
                      hipStream_t s) {
    const uint32_t tiles_x = (cam.width + 7) / 8, tiles_y = (cam.height + 7) / 8;
    const uint32_t ntiles = tiles_x * tiles_y;
    if (count)
        hipLaunchKernelGGL(aov_kernel<true>, dim3(grid), dim3(256), 0, s, sc, cam, out, passes, counters,
                           next_counters, tiles_x, ntiles, ovf_ref, ovf_t, ovf_stride, stat, next_stat);
    else
        hipLaunchKernelGGL(aov_kernel<false>, dim3(grid), dim3(256), 0, s, sc, cam, out, passes, counters,
                           next_counters, tiles_x, ntiles, ovf_ref, ovf_t, ovf_stride, stat, next_stat);
    return hipGetLastError();
}

}  // namespace rtamd
