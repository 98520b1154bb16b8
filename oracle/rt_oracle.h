/*
 * rt_oracle.h -- CPU restatement of RenderToy's path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function restates the reference algorithm with the same IEEE-754
 * single-precision operation order (built with -ffp-contract=off).  Sources
 * followed (paths relative to the RenderToy tree):
 *   camera rays      src/renderer.cpp:103-124,253-259, src/object.cpp:42-64
 *   triangle cache   src/object.cpp:83-105,180-243
 *   Moller-Trumbore  src/object.cpp:107-167
 *   octree BVH       src/bvh.cpp:9-341, include/RenderToy/bvh.h:54-63
 *   AOV passes       src/renderer.cpp:20-54,131-229,391-427
 *   Disney BSDF      src/material.cpp:12-295, src/rtmath.cpp:169-353
 *   integrator       src/renderer.cpp:236-384, src/world.cpp:5-42,
 *                    include/RenderToy/surfacepoint.h:20-45
 *   RNG              src/rtmath.cpp:355-373 (std::mt19937, default seed)
 *
 * Pinning of each function is listed in the header comment of rt_oracle.cpp.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Material record: 24 floats, same field order as PrincipledBSDF
 * (include/RenderToy/material.h:12-77):
 *  0-2 base_color  3-5 emission  6 roughness  7 metallic  8 spec_trans
 *  9 ior  10 anisotropic  11 subsurface  12 specular_tint  13 sheen
 *  14 sheen_tint  15 clearcoat  16 clearcoat_roughness  17 at_distance
 *  18-20 extinction  21-23 unused */
#define ORC_MAT_FLOATS 24

typedef struct orc_scene orc_scene;

/* verts: [n][3][3] world-space vertices (Triangle::vert_w before caching),
 * norms: [n][3][3] per-vertex normals (may be NULL when no mesh is smooth),
 * tri_mesh: [n] mesh index; mesh_mat/mesh_smooth: [n_meshes];
 * mats: [n_mats][24]; emitters: [n_emit] triangle indices in
 * World::emissive_triangles order. */
orc_scene *orc_scene_create(uint32_t n, const float *verts, const float *norms,
                            const uint32_t *tri_mesh, uint32_t n_meshes,
                            const uint32_t *mesh_mat, const uint8_t *mesh_smooth,
                            uint32_t n_mats, const float *mats,
                            uint32_t n_emit, const uint32_t *emitters,
                            const float *sky, const float *ground);
void orc_scene_destroy(orc_scene *s);

/* Triangle cache as Triangle::UpdateCache computes it:
 * out[n][13] = v0(3) e1(3) e2(3) normal(3) area(1). */
void orc_tri_cache(const orc_scene *s, float *out);

/* Octree build statistics: nodes, leaves, max depth, root box (6 floats). */
void orc_octree_stats(orc_scene *s, int64_t *nodes, int64_t *leaves, int64_t *max_leaf, float *root_box);

/* Renderer::PrepareScreenSpace */
void orc_prepare_screen_space(float gate_x, float gate_y, float focal, uint32_t W, uint32_t H,
                              float *top, float *right);

/* Camera rays for pixels (x, y) with y=0 the bottom row; out_src/out_dir [W*H][3]. */
void orc_camera_rays(const float *o2w, float top, float right, uint32_t W, uint32_t H,
                     float *out_src, float *out_dir);

/* Closest hit.  mode 0 = brute force over World::triangles (bvh.cpp:268-286),
 * mode 1 = octree best-first traversal (bvh.cpp:288-339).  exclude may be NULL
 * (no exclusion) else per-ray triangle index (0xFFFFFFFF = none).
 * out_id = 0xFFFFFFFF on miss; t/u/v are written for hits only. */
void orc_intersect(orc_scene *s, int mode, uint64_t nrays, const float *src, const float *dir,
                   const uint32_t *exclude, uint32_t *out_id, float *out_t, float *out_u,
                   float *out_v, int nthreads);

/* AOV passes on the pixels listed in pix (NULL = all W*H, in order).
 * Outputs are per listed pixel: id (u32), t, normal pass rgb, depth pass,
 * albedo rgb, intersect-test mask.  Any output pointer may be NULL. */
void orc_render_aov(orc_scene *s, int mode, const float *o2w, float top, float right,
                    uint32_t W, uint32_t H, float near_clip, float far_clip,
                    uint64_t npix, const uint32_t *pix,
                    uint32_t *out_id, float *out_t, float *out_normal, float *out_depth,
                    float *out_albedo, float *out_mask, int nthreads);

/* Disney BSDF pieces (material.cpp:143-295). */
void orc_bsdf_eval(const float *mat, float eta, const float *V, const float *N, const float *L,
                   float *f_out, float *pdf_out);
int orc_bsdf_sample(const float *mat, float eta, const float *ffnormal, const float *in_dir,
                    float r1, float r2, float *out_dir, float *color, float *pdf);

/* Philox4x32-10 (Salmon et al., SC'11) block function. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Path tracer, PathTracingRenderer::Render restated.
 * rng_mode 0: one std::mt19937 (default seed 5489) shared in the reference's
 *             single-thread draw order, after `warmup` discarded Random::Float draws.
 * rng_mode 1: counter-based Philox stream keyed by seed, the stream the GPU
 *             path tracer uses (pixel, sample, depth addressing).
 * Renders the pixel rectangle [x0,x1)x[y0,y1) of a W x H frame; out is
 * [(y1-y0)*(x1-x0)][3], accumulated as buffer += L/spp in sample order.
 * max_depth = number of path vertices (the reference hardwires 5).
 * nthreads > 1 is allowed only with rng_mode 1.
 * counts (may be NULL) receives {extension rays, shadow rays whose contribution
 * is not identically zero}; returns their sum. */
int64_t orc_render_pt(orc_scene *s, int mode, const float *o2w, float top, float right,
                      uint32_t W, uint32_t H, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
                      uint32_t spp, uint32_t max_depth, int rng_mode, uint64_t seed,
                      uint64_t warmup, float *out, int nthreads, int64_t *counts);

/* Per-sample radiance for listed (pixel, sample) pairs in Philox mode, for
 * per-path comparisons: out[n][3]. */
void orc_pt_samples(orc_scene *s, int mode, const float *o2w, float top, float right,
                    uint32_t W, uint32_t H, uint64_t n, const uint32_t *pixel,
                    const uint32_t *sample, uint32_t max_depth, uint64_t seed, float *out,
                    int nthreads);

#ifdef __cplusplus
}
#endif

#endif
