/*
 * ref_driver.cpp -- C entry points over the RenderToy reference, compiled from
 * the reference's own sources (see oracle/Makefile) into oracle/_ref/libref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  Used to write golden fixtures (tests/golden/) and
 * as the CPU baseline.  It builds a RenderToy::World through the reference's
 * public API and calls its renderers / BVH / BSDF unchanged.
 *
 * Access: Triangle keeps its cached edges private (object.h:148-165) and
 * Renderer::PrepareScreenSpace is protected (renderer.h:111); this translation
 * unit opens them for reading only, without changing any layout.
 */
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <optional>
#include <random>
#include <sstream>
#include <unordered_map>
#include <vector>

#define private public
#define protected public
#include <RenderToy/rendertoy.h>
#undef private
#undef protected

using namespace RenderToy;

/* Array allocations made inside libref get a zero-filled guard band on both
 * sides (the library is linked -Bsymbolic, so the reference's own new[]
 * resolves here).  Image's reflect edge mode reads up to a row past an image
 * buffer (compositor.h:26-40: coordinate n maps to n), which is undefined
 * behaviour; with the guard band those reads return black, the value the
 * oracle defines for them.  Nothing else in the reference reads outside its
 * arrays, so its results are unchanged. */
static constexpr std::size_t kGuard = 1u << 20;   // bytes on each side

void *operator new[](std::size_t n) {
    char *p = static_cast<char *>(std::malloc(n + 2 * kGuard));
    if (!p) throw std::bad_alloc();
    std::memset(p, 0, kGuard);
    std::memset(p + kGuard + n, 0, kGuard);
    return p + kGuard;
}
void operator delete[](void *p) noexcept {
    if (p) std::free(static_cast<char *>(p) - kGuard);
}
void operator delete[](void *p, std::size_t) noexcept { operator delete[](p); }

namespace {
struct RefWorld {
    World world;
    std::vector<Mesh *> meshes;
    std::vector<PrincipledBSDF *> mats;
    std::unordered_map<const Triangle *, uint32_t> ids;
    RenderContext *rc = nullptr;
    uint32_t W = 0, H = 0;
    std::mt19937 replica;   /* tracks the reference's global engine (rtmath.cpp:355) */
};

PrincipledBSDF *make_mat(const float *m) {
    return new PrincipledBSDF(Vector3f(m[0], m[1], m[2]), Vector3f(m[3], m[4], m[5]), m[6], m[7], m[8], m[9], m[10],
                              m[11], m[12], m[13], m[14], m[15], m[16], m[17], Vector3f(m[18], m[19], m[20]));
}

void ensure_rc(RefWorld *w, uint32_t W, uint32_t H) {
    if (w->rc && w->W == W && w->H == H) return;
    delete w->rc;
    w->rc = new RenderContext(&w->world, FormatSettings(SizeN(W, H), Vector2f(1.0f, 1.0f)));
    w->W = W;
    w->H = H;
}
}  // namespace

extern "C" {

void *ref_world_create(uint32_t n, const float *verts, const float *norms, const uint32_t *tri_mesh,
                       uint32_t n_meshes, const uint32_t *mesh_mat, const uint8_t *mesh_smooth, uint32_t n_mats,
                       const float *mats, const float *sky, const float *ground) {
    RefWorld *w = new RefWorld();
    for (uint32_t i = 0; i < n_mats; ++i) w->mats.push_back(make_mat(mats + 24 * i));
    for (uint32_t i = 0; i < n_meshes; ++i) {
        Mesh *m = new Mesh();
        m->tex = w->mats[mesh_mat[i]];
        m->smooth = mesh_smooth[i] != 0;
        w->meshes.push_back(m);
        w->world.meshes.push_back(m);
    }
    for (uint32_t i = 0; i < n; ++i) {
        std::array<Vector3f, 3> v, nn;
        std::array<Vector2f, 3> uv;
        for (int k = 0; k < 3; ++k) {
            v[k] = Vector3f(verts[i * 9 + k * 3], verts[i * 9 + k * 3 + 1], verts[i * 9 + k * 3 + 2]);
            nn[k] = norms ? Vector3f(norms[i * 9 + k * 3], norms[i * 9 + k * 3 + 1], norms[i * 9 + k * 3 + 2])
                          : Vector3f::O;
            uv[k] = Vector2f::O;
        }
        Mesh *m = w->meshes[tri_mesh[i]];
        Triangle *t = new Triangle(v, nn, uv, m);
        w->world.triangles.push_back(t);
        m->tris.push_back(t);
        w->ids[t] = i;
    }
    if (sky) w->world.sky_emission = Vector3f(sky[0], sky[1], sky[2]);
    if (ground) w->world.ground_reflection = Vector3f(ground[0], ground[1], ground[2]);
    w->world.PrepareDirectLightSampling();
    return w;
}

uint32_t ref_world_emitters(void *wp, uint32_t *out, uint32_t cap) {
    RefWorld *w = (RefWorld *)wp;
    uint32_t k = 0;
    for (const Triangle *t : w->world.emissive_triangles) {
        if (k < cap) out[k] = w->ids[t];
        ++k;
    }
    return k;
}

/* out[n][13]: vert_w[0], v0v1_w, v0v2_w, normal, area (object.cpp:224-243) */
void ref_tri_cache(void *wp, float *out) {
    RefWorld *w = (RefWorld *)wp;
    for (size_t i = 0; i < w->world.triangles.size(); ++i) {
        const Triangle *t = w->world.triangles[i];
        float *o = out + 13 * i;
        for (int k = 0; k < 3; ++k) {
            o[k] = t->vert_w[0][k];
            o[3 + k] = t->v0v1_w[k];
            o[6 + k] = t->v0v2_w[k];
            o[9 + k] = t->normal[k];
        }
        o[12] = t->area;
    }
}

/* BVH::Intersect over explicit rays (bvh.cpp:266-341). */
void ref_intersect(void *wp, uint64_t nrays, const float *src, const float *dir, const uint32_t *exclude,
                   uint32_t *out_id, float *out_t, float *out_u, float *out_v) {
    RefWorld *w = (RefWorld *)wp;
    ensure_rc(w, 1, 1);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < (int64_t)nrays; ++i) {
        Ray r(Vector3f(src[3 * i], src[3 * i + 1], src[3 * i + 2]), Vector3f(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]));
        const Triangle *ex = nullptr;
        if (exclude && exclude[i] != 0xFFFFFFFFu) ex = w->world.triangles[exclude[i]];
        Vector3f pos;
        float t = 0, u = 0, v = 0;
        const Triangle *h = w->rc->bvh->Intersect(r, pos, t, u, v, ex);
        out_id[i] = h ? w->ids.at(h) : 0xFFFFFFFFu;
        if (h) {
            out_t[i] = t;
            out_u[i] = u;
            out_v[i] = v;
        }
    }
}

/* Camera::SetO2W + Renderer::PrepareScreenSpace (renderer.cpp:103-124). */
void ref_set_camera(void *wp, const float *o2w, float focal, float gate_x, float gate_y, float near_clip,
                    float far_clip) {
    RefWorld *w = (RefWorld *)wp;
    w->world.cameras.clear();
    w->world.cameras.push_back(Camera(Matrix4x4f::I, focal, Vector2f(gate_x, gate_y), near_clip, far_clip));
    Matrix4x4f M;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) M[i][j] = o2w[i * 4 + j];
    w->world.cameras[0].SetO2W(M);
}

void ref_screen_space(void *wp, uint32_t W, uint32_t H, float *top, float *right) {
    RefWorld *w = (RefWorld *)wp;
    ensure_rc(w, W, H);
    NormalRenderer r(w->rc);
    r.PrepareScreenSpace(&w->world.cameras[0], *top, *right);
}

/* kind: 0 normal, 1 depth, 2 albedo, 3 intersect-test, 4 path tracing (spp).
 * Returns the milliseconds spent in Render(). */
double ref_render(void *wp, int kind, uint32_t W, uint32_t H, uint32_t spp, uint64_t warmup, float *out) {
    RefWorld *w = (RefWorld *)wp;
    ensure_rc(w, W, H);
    std::memset((void *)w->rc->buffer, 0, sizeof(Vector3f) * (size_t)W * H);
    for (uint64_t i = 0; i < warmup; ++i) (void)Random::Float();
    Renderer *r = nullptr;
    switch (kind) {
    case 0: r = new NormalRenderer(w->rc); break;
    case 1: r = new DepthBufferRenderer(w->rc); break;
    case 2: r = new AlbedoRenderer(w->rc); break;
    case 3: r = new IntersectTestRenderer(w->rc); break;
    default: r = new PathTracingRenderer(w->rc, (int)spp); break;
    }
    auto t0 = std::chrono::steady_clock::now();
    r->Render();
    auto t1 = std::chrono::steady_clock::now();
    delete r;
    for (size_t i = 0; i < (size_t)W * H; ++i)
        for (int k = 0; k < 3; ++k) out[3 * i + k] = w->rc->buffer[i][k];
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

/* PrincipledBSDF::Eval (material.cpp:143-204). */
void ref_bsdf_eval(const float *mat, float eta, const float *V, const float *N, const float *L, float *f_out,
                   float *pdf_out) {
    PrincipledBSDF *m = make_mat(mat);
    RayState st;
    st.eta = eta;
    float pdf = 0;
    Vector3f f = m->Eval(st, Vector3f(V[0], V[1], V[2]), Vector3f(N[0], N[1], N[2]), Vector3f(L[0], L[1], L[2]), pdf);
    for (int k = 0; k < 3; ++k) f_out[k] = f[k];
    *pdf_out = pdf;
    delete m;
}

/* PrincipledBSDF::Sample (material.cpp:206-295).  The two uniforms it draws
 * from the global engine are predicted by a replica engine and returned in r. */
int ref_bsdf_sample(void *wp, const float *mat, float eta, const float *ffnormal, const float *in_dir, float *r,
                    float *out_dir, float *color, float *pdf) {
    RefWorld *w = (RefWorld *)wp;
    PrincipledBSDF *m = make_mat(mat);
    std::uniform_real_distribution<float> d(0.0f, 1.0f);
    r[0] = d(w->replica);
    r[1] = d(w->replica);
    RayState st;
    st.eta = eta;
    st.ffnormal = Vector3f(ffnormal[0], ffnormal[1], ffnormal[2]);
    Vector3f od, c;
    float p = 0;
    bool ok = m->Sample(Vector3f(in_dir[0], in_dir[1], in_dir[2]), od, c, p, st);
    for (int k = 0; k < 3; ++k) {
        out_dir[k] = od[k];
        color[k] = c[k];
    }
    *pdf = p;
    delete m;
    return ok ? 1 : 0;
}

/* Matrix helpers used by the demos: Translation * RotationEulerXYZ (rtmath.cpp:134-167). */
void ref_translation_rotation(const float *t, const float *euler, float *out) {
    Matrix4x4f M = AffineTransformation::Translation(Vector3f(t[0], t[1], t[2])) *
                   AffineTransformation::RotationEulerXYZ(Vector3f(euler[0], euler[1], euler[2]));
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[i * 4 + j] = M[i][j];
}

void ref_world_destroy(void *wp) {
    RefWorld *w = (RefWorld *)wp;
    delete w->rc;
    w->rc = nullptr;
    /* World owns raw pointers and never frees them (world.h:13-41); leak as the reference does. */
    delete w;
}


/* ---- Image (include/RenderToy/compositor.h, src/compositor.cpp) ---------- */

static void img_in(Image &img, const float *src) {
    const std::size_t n = img.resolution.Area();
    for (std::size_t i = 0; i < n; ++i) img.GetBuffer()[i] = Vector3f(src[3 * i], src[3 * i + 1], src[3 * i + 2]);
}
static void img_out(const Image &img, float *dst) {
    const std::size_t n = img.resolution.Area();
    for (std::size_t i = 0; i < n; ++i) {
        const Vector3f &v = img.GetBuffer()[i];
        dst[3 * i] = v.x();
        dst[3 * i + 1] = v.y();
        dst[3 * i + 2] = v.z();
    }
}

/* UnitizedGaussianKernel<float, _O> weights; orient 0 All, 1 X, 2 Y. */
void ref_gaussian_weights(uint32_t size, float sigma, int orient, float *out) {
    if (orient == 0) {
        UnitizedGaussianKernel<float, Orientation::All> k(size, sigma);
        for (uint32_t i = 0; i < size; ++i)
            for (uint32_t j = 0; j < size; ++j) out[i * size + j] = k.kernel_mat[i][j];
    } else if (orient == 1) {
        UnitizedGaussianKernel<float, Orientation::X> k(size, sigma);
        for (uint32_t j = 0; j < size; ++j) out[j] = k.kernel_mat[0][j];
    } else {
        UnitizedGaussianKernel<float, Orientation::Y> k(size, sigma);
        for (uint32_t i = 0; i < size; ++i) out[i] = k.kernel_mat[i][0];
    }
}

/* Image::Convolute<_Normalize, _Length, _Abs> with a rows x cols kernel;
 * flags: 1 normalise, 2 length, 4 abs. */
void ref_image_convolve(float *data, uint32_t W, uint32_t H, uint32_t rows, uint32_t cols, const float *k,
                        int flags) {
    Image img(SizeN(W, H));
    img_in(img, data);
    IConvolutionKernel<float> ker(rows, cols);
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t j = 0; j < cols; ++j) ker.kernel_mat[i][j] = k[i * cols + j];
    switch (flags & 7) {
    case 0: img.Convolute<false, false, false>(ker); break;
    case 1: img.Convolute<true, false, false>(ker); break;
    case 2: img.Convolute<false, true, false>(ker); break;
    case 3: img.Convolute<true, true, false>(ker); break;
    case 4: img.Convolute<false, false, true>(ker); break;
    case 5: img.Convolute<true, false, true>(ker); break;
    case 6: img.Convolute<false, true, true>(ker); break;
    default: img.Convolute<true, true, true>(ker); break;
    }
    img_out(img, data);
}

/* op: 0 GaussianBlur(size, a), 1 Bloom(size, a, b), 2 Tonemap(a), 3
 * GammaCorrection(a), 4 GreyScale<BT709>, 5 Binarize<BT709>(a), 6/7/8
 * EdgeDetection<X/Y/All>, 9 GreyScale<BT601>, 10 Tonemap<kAverage>(a). */
void ref_image_apply(int op, float *data, uint32_t W, uint32_t H, uint32_t size, float a, float b) {
    Image img(SizeN(W, H));
    img_in(img, data);
    switch (op) {
    case 0: img.GaussianBlur(size, a); break;
    case 1: img.Bloom(size, a, b); break;
    case 2: img.Tonemap(a); break;
    case 3: img.GammaCorrection(a); break;
    case 4: img.GreyScale<Convert::ColorStandard::kITURBT709>(); break;
    case 5: img.Binarize<Convert::ColorStandard::kITURBT709>(a); break;
    case 6: img.EdgeDetection<Orientation::X>(); break;
    case 7: img.EdgeDetection<Orientation::Y>(); break;
    case 8: img.EdgeDetection<Orientation::All>(); break;
    case 9: img.GreyScale<Convert::ColorStandard::kITURBT601>(); break;
    default: img.Tonemap<Convert::ColorStandard::kAverage>(a); break;
    }
    img_out(img, data);
}

/* Image::operator+ - * / ; op 0..3. */
void ref_image_arith(const float *a, const float *b, float *out, uint32_t W, uint32_t H, int op) {
    Image ia(SizeN(W, H)), ib(SizeN(W, H));
    img_in(ia, a);
    img_in(ib, b);
    Image r = op == 0 ? ia + ib : op == 1 ? ia - ib : op == 2 ? ia * ib : ia / ib;
    img_out(r, out);
}


/* PPMExporter / BMPExporter / ASCIIExporter (src/exporter.cpp); format 0 BMP,
 * 1 PPM, 2 ASCII.  Returns the byte count (written to out when it fits). */
uint64_t ref_image_export(const float *data, uint32_t W, uint32_t H, int format, uint8_t *out, uint64_t cap) {
    Image img(SizeN(W, H));
    img_in(img, data);
    std::ostringstream os;
    if (format == 0) BMPExporter(img).Export(os);
    else if (format == 1) PPMExporter(img).Export(os);
    else ASCIIExporter(img).Export(os);
    const std::string s = os.str();
    if (out && s.size() <= cap) std::memcpy(out, s.data(), s.size());
    return s.size();
}


/* OBJModelImporter::Import (src/importer.cpp:29-146) into a fresh World:
 * world.triangles' object-space vertices and normals (9 floats each), the index
 * of each triangle's mesh in world.meshes (or 0xFFFFFFFF), each mesh's smooth
 * flag, and its name (names joined with '\n').  Returns the triangle count;
 * arrays are written when the capacities suffice. */
uint64_t ref_obj_import(const char *path, float *verts, float *norms, uint32_t *tri_mesh, uint64_t cap_tris,
                        uint8_t *smooth, char *names, uint64_t cap_meshes, uint64_t cap_names, uint64_t *n_meshes) {
    World w;
    OBJModelImporter::Import(w, path);
    const uint64_t n = w.triangles.size();
    *n_meshes = w.meshes.size();
    std::unordered_map<const Mesh *, uint32_t> mid;
    for (size_t m = 0; m < w.meshes.size(); ++m) mid[w.meshes[m]] = (uint32_t)m;
    if (n <= cap_tris)
        for (uint64_t i = 0; i < n; ++i) {
            const Triangle *t = w.triangles[i];
            for (int k = 0; k < 3; ++k)
                for (int c = 0; c < 3; ++c) {
                    verts[9 * i + 3 * k + c] = t->vert[k][c];
                    norms[9 * i + 3 * k + c] = t->norm[k][c];
                }
            auto it = mid.find(t->parent);
            tri_mesh[i] = it == mid.end() ? 0xFFFFFFFFu : it->second;
        }
    if (w.meshes.size() <= cap_meshes) {
        std::string all;
        for (size_t m = 0; m < w.meshes.size(); ++m) {
            smooth[m] = w.meshes[m]->smooth ? 1 : 0;
            all += w.meshes[m]->name;
            all += '\n';
        }
        if (all.size() < cap_names) std::memcpy(names, all.c_str(), all.size() + 1);
    }
    return n;
}

}  // extern "C"
