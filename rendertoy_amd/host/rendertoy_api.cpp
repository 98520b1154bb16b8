// rendertoy_api.cpp -- host side of the RenderToy C++ API (include/RenderToy/) on MI355X.
//
// Scene assembly (Geometry, Camera, Triangle, Mesh, World) keeps RenderToy's
// float semantics so the data handed to the device is what RenderToy's own
// renderers would see.  RenderContext flattens the World into an rt_scene once
// (where RenderToy builds its BVH, src/renderer.cpp:77-85), and every
// Renderer::Render() is one call into the C ABI (include/rendertoy_amd.h).
// Built with -ffp-contract=off.
#include <RenderToy/rendertoy.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rendertoy_amd.h"

namespace RenderToy {

namespace {
void check(int rc) {
    if (rc != RT_OK) throw Exception::DeviceException(rt_last_error());
}
}  // namespace

// ---------------------------------------------------------------- rtmath

namespace Convert {
Vector3f XYZToSRGB(const Vector3f &xyz) {
    // Linear sRGB (D65) from CIE XYZ, Lindbloom's matrix.
    const Matrix3x3f m = {{3.2404542f, -1.5371385f, -0.4985314f},
                          {-0.9692660f, 1.8760108f, 0.0415560f},
                          {0.0556434f, -0.2040259f, 1.0572252f}};
    return m * xyz;
}

Vector3f BlackBody(const float t) {
    // Krystek (1985): Planckian locus in CIE 1960 (u, v) as rational functions of T,
    // evaluated with double coefficients and stored as float.
    float tt = t * t;
    float u = (0.860117757 + 1.54118254e-4 * t + 1.28641212e-7 * tt) / (1.0 + 8.42420235e-4 * t + 7.08145163e-7 * tt);
    float v = (0.317398726 + 4.22806245e-5 * t + 4.20481691e-8 * tt) / (1.0 - 2.89741816e-5 * t + 1.61456053e-7 * tt);
    // (u, v) -> CIE xy -> XYZ with Y = 1
    Vector2f xy = Vector2f(3.0f * u, 2.0f * v) / (2.0 * u - 8.0 * v + 4.0);
    Vector3f xyz = Vector3f(xy.x() / xy.y(), 1.0f, (1.0f - xy.x() - xy.y()) / xy.y());
    Vector3f rgb = XYZToSRGB(xyz);
    return rgb / std::max(rgb.x(), std::max(rgb.y(), rgb.z()));
}
}  // namespace Convert

namespace AffineTransformation {
Matrix4x4f Translation(const Vector3f &origin) {
    Matrix4x4f m(Matrix4x4f::I);
    for (int i = 0; i < 3; ++i) m[i][3] = origin[i];
    return m;
}

Matrix4x4f RotationEulerXYZ(const Vector3f &e) {
    const float x = e.x(), y = e.y(), z = e.z();
    const Matrix4x4f Z = {{std::cos(z), -std::sin(z), 0.0f, 0.0f},
                          {std::sin(z), std::cos(z), 0.0f, 0.0f},
                          {0.0f, 0.0f, 1.0f, 0.0f},
                          {0.0f, 0.0f, 0.0f, 1.0f}};
    const Matrix4x4f Y = {{std::cos(y), 0.0f, -std::sin(y), 0.0f},
                          {0.0f, 1.0f, 0.0f, 0.0f},
                          {std::sin(y), 0.0f, std::cos(y), 0.0f},
                          {0.0f, 0.0f, 0.0f, 1.0f}};
    const Matrix4x4f X = {{1.0f, 0.0f, 0.0f, 0.0f},
                          {0.0f, std::cos(x), -std::sin(x), 0.0f},
                          {0.0f, std::sin(x), std::cos(x), 0.0f},
                          {0.0f, 0.0f, 0.0f, 1.0f}};
    return Z * Y * X;
}
}  // namespace AffineTransformation

// ---------------------------------------------------------------- objects

Geometry::Geometry() : object_to_world(Matrix4x4f::I), world_to_object(Matrix4x4f::I) {}

void Geometry::SetO2W(const Matrix4x4f &m) {
    object_to_world = m;
    // inverse of a rigid transform: [R^T | -R^T t]
    Matrix3x3f R = m.ComplementMinor(3, 3).Transposed();
    Vector3f t = {m[0][3], m[1][3], m[2][3]};
    Vector3f c = (-1.0f) * (R * t);
    world_to_object = Matrix4x4f::I;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) world_to_object[i][j] = R[i][j];
        world_to_object[i][3] = c[i];
    }
}

Vector3f Geometry::O2WTransform(const Vector3f &v) const {
    Vector4f r = object_to_world * Vector4f(v.x(), v.y(), v.z(), 1.0f);
    return Vector3f(r[0], r[1], r[2]);
}
Vector3f Geometry::W2OTransform(const Vector3f &v) const {
    Vector4f r = world_to_object * Vector4f(v.x(), v.y(), v.z(), 1.0f);
    return Vector3f(r[0], r[1], r[2]);
}
Ray Geometry::O2WTransform(const Ray &ray) const {
    Ray r(O2WTransform(ray.src), O2WTransform(ray.direction));
    r.direction -= r.src;
    r.direction.Normalize();
    return r;
}
Ray Geometry::W2OTransform(const Ray &ray) const {
    Ray r(W2OTransform(ray.src), W2OTransform(ray.direction));
    r.direction -= r.src;
    r.direction.Normalize();
    return r;
}

Camera::Camera(Matrix4x4f, float focal_length_, Vector2f gate_dimension_, float near_clipping_plane_,
               float far_clipping_plane_)
    : focal_length(focal_length_), gate_dimension(gate_dimension_), near_clipping_plane(near_clipping_plane_),
      far_clipping_plane(far_clipping_plane_) {
    Geometry::SetO2W(object_to_world);
}

Triangle::Triangle(const std::array<Vector3f, 3> &vert_, const std::array<Vector3f, 3> &norm_,
                   const std::array<Vector2f, 3> &uv_, Mesh *const parent_)
    : vert(vert_), norm(norm_), uv(uv_), parent(parent_), vert_w(vert_) {
    UpdateCache();
}

void Triangle::UpdateCache() {
    for (int k = 0; k < 3; ++k) vert_w[k] = parent ? parent->O2WTransform(vert[k]) : vert[k];
    v0v1_w = vert_w[1] - vert_w[0];
    v0v2_w = vert_w[2] - vert_w[0];
    area = v0v1_w.Cross(v0v2_w).Length() * 0.5f;
    normal = v0v1_w.Normalized().Cross(vert_w[2] - vert_w[1]).Normalized();
}

Vector3f Triangle::NormalC(const float u, const float v) const {
    if (parent && parent->smooth) return ((1 - u - v) * norm[0] + u * norm[1] + v * norm[2]).Normalized();
    return normal;
}

Polygon::Polygon(const std::size_t V_, const std::vector<Vector3f> &vert_, const std::vector<Vector3f> &norm_,
                 const std::vector<Vector2f> &uv_, Mesh *parent_)
    : V(V_), vert(vert_), norm(norm_), uv(uv_), parent(parent_) {}

std::vector<Triangle *> Polygon::ConvertToTriangle() const {
    std::vector<Triangle *> out;
    for (std::size_t b = 1; b + 1 < V; ++b)
        out.push_back(new Triangle({vert[0], vert[b], vert[b + 1]}, {norm[0], norm[b], norm[b + 1]},
                                   {uv[0], uv[b], uv[b + 1]}, parent));
    return out;
}

void Mesh::SetO2W(const Matrix4x4f &m) {
    Geometry::SetO2W(m);
    for (Triangle *t : tris) t->UpdateCache();
}

PrincipledBSDF::PrincipledBSDF(const Vector3f &base_color_, const Vector3f &emission_, const float roughness_,
                               const float metallic_, const float spec_trans_, const float ior_,
                               const float anisotropic_, const float subsurface_, const float specular_tint_,
                               const float sheen_, const float sheen_tint_, const float clearcoat_,
                               const float clearcoat_roughness_, const float at_distance_, const Vector3f &extinction_)
    : base_color(base_color_), emission(emission_), roughness(roughness_), metallic(metallic_),
      anisotropic(anisotropic_), subsurface(subsurface_), specular_tint(specular_tint_), sheen(sheen_),
      sheen_tint(sheen_tint_), clearcoat(clearcoat_), clearcoat_roughness(clearcoat_roughness_),
      spec_trans(spec_trans_), ior(ior_), at_distance(at_distance_), extinction(extinction_) {
    if (roughness < 0.001f) roughness = 0.001f;
}

void World::PrepareDirectLightSampling() {
    emissive_triangles.clear();
    for (Mesh *m : meshes)
        if (m->tex && m->tex->emission != Vector3f::O)
            for (Triangle *t : m->tris) emissive_triangles.push_back(t);
}

// ---------------------------------------------------------------- device scene

namespace {
rt_scene *flatten(World &world) {
    const size_t n = world.triangles.size();
    std::vector<float> v(n * 9), nrm(n * 9);
    std::vector<uint32_t> tri_mesh(n), mesh_mat, emit;
    std::vector<uint8_t> smooth;
    std::vector<rt_material> mats;
    std::unordered_map<const Mesh *, uint32_t> mesh_id;
    std::unordered_map<const PrincipledBSDF *, uint32_t> mat_id;
    std::unordered_map<const Triangle *, uint32_t> tri_id;
    static const PrincipledBSDF default_material;
    auto add_mesh = [&](const Mesh *m) {
        mesh_id[m] = (uint32_t)mesh_mat.size();
        const PrincipledBSDF *b = m->tex ? m->tex : &default_material;
        if (!mat_id.count(b)) {
            rt_material r;
            std::memset(&r, 0, sizeof(r));
            for (int c = 0; c < 3; ++c) {
                r.base_color[c] = b->base_color[c];
                r.emission[c] = b->emission[c];
                r.extinction[c] = b->extinction[c];
            }
            r.roughness = b->roughness;
            r.metallic = b->metallic;
            r.spec_trans = b->spec_trans;
            r.ior = b->ior;
            r.anisotropic = b->anisotropic;
            r.subsurface = b->subsurface;
            r.specular_tint = b->specular_tint;
            r.sheen = b->sheen;
            r.sheen_tint = b->sheen_tint;
            r.clearcoat = b->clearcoat;
            r.clearcoat_roughness = b->clearcoat_roughness;
            r.at_distance = b->at_distance;
            mat_id[b] = (uint32_t)mats.size();
            mats.push_back(r);
        }
        mesh_mat.push_back(mat_id[b]);
        smooth.push_back(m->smooth ? 1 : 0);
    };
    for (Mesh *m : world.meshes) add_mesh(m);
    for (size_t i = 0; i < n; ++i) {
        const Triangle *t = world.triangles[i];
        if (!t->parent) throw Exception::IRenderToyException("triangle without a parent mesh");
        if (!mesh_id.count(t->parent)) add_mesh(t->parent);
        tri_id[t] = (uint32_t)i;
        const auto &w = t->WorldVertices();
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) {
                v[9 * i + 3 * k + c] = w[k][c];
                nrm[9 * i + 3 * k + c] = t->norm[k][c];
            }
        tri_mesh[i] = mesh_id[t->parent];
    }
    for (const Triangle *t : world.EmissiveTriangles()) {
        auto it = tri_id.find(t);
        if (it != tri_id.end()) emit.push_back(it->second);
    }
    rt_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.num_triangles = (uint32_t)n;
    d.vertices = v.data();
    d.normals = nrm.data();
    d.tri_mesh = tri_mesh.data();
    d.num_meshes = (uint32_t)mesh_mat.size();
    d.mesh_material = mesh_mat.data();
    d.mesh_smooth = smooth.data();
    d.num_materials = (uint32_t)mats.size();
    d.materials = mats.data();
    d.num_emitters = (uint32_t)emit.size();
    d.emitters = emit.data();
    for (int c = 0; c < 3; ++c) {
        d.sky_emission[c] = world.sky_emission[c];
        d.ground_reflection[c] = world.ground_reflection[c];
    }
    rt_scene *s = nullptr;
    check(rt_scene_create(&d, &s));
    return s;
}
}  // namespace

RenderContext::RenderContext(World *world_, FormatSettings fs) : world(world_), format_settings(fs) {
    buffer = new Vector3f[format_settings.resolution.Area()];
    if (world) {
        try {
            device_scene = flatten(*world);
        } catch (...) {
            delete[] buffer;
            throw;
        }
    }
}

RenderContext::~RenderContext() {
    delete[] buffer;
    if (device_scene) rt_scene_destroy(device_scene);
}

// ---------------------------------------------------------------- renderers

void Renderer::PrepareScreenSpace(Camera *cam, float &top, float &right) {
    top = (Convert::InchToMM(cam->gate_dimension.y()) / 2.0f) / cam->focal_length;
    right = (Convert::InchToMM(cam->gate_dimension.x()) / 2.0f) / cam->focal_length;
    float gate_aspr = cam->gate_dimension.x() / cam->gate_dimension.y();
    float film_aspr = render_context->format_settings.resolution.AspectRatio();
    float xscale = 1.0f, yscale = 1.0f;
    if (gate_aspr >= film_aspr) xscale = film_aspr / gate_aspr;
    else yscale = gate_aspr / film_aspr;
    right *= xscale;
    top *= yscale;
}

namespace {
Camera *camera_of(RenderContext *rc) {
    if (!rc || !rc->world || !rc->device_scene) throw Exception::IRenderToyException("RenderContext has no world");
    if (rc->camera_id < 0 || (size_t)rc->camera_id >= rc->world->cameras.size())
        throw Exception::IRenderToyException("camera_id out of range");
    return &rc->world->cameras[rc->camera_id];
}

rt_camera fill_camera(const RenderContext *rc, const Camera *cam, float top, float right, float near_clip,
                      float far_clip) {
    rt_camera c;
    std::memset(&c, 0, sizeof(c));
    const Matrix4x4f &M = cam->GetO2W();
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) c.o2w[4 * i + j] = M[i][j];
    c.top = top;
    c.right = right;
    c.near_clip = near_clip;
    c.far_clip = far_clip;
    c.width = (uint32_t)rc->format_settings.resolution.width;
    c.height = (uint32_t)rc->format_settings.resolution.height;
    return c;
}

// Single-channel pass -> Vector3f(v) per pixel, as RenderToy writes depth and masks.
void expand(RenderContext *rc, const std::vector<float> &v) {
    for (size_t i = 0; i < v.size(); ++i) rc->buffer[i] = Vector3f(v[i]);
}
}  // namespace

#define RT_CAMERA(near_clip, far_clip)                                                       \
    Camera *cam = camera_of(render_context);                                                 \
    float top, right;                                                                        \
    PrepareScreenSpace(cam, top, right);                                                     \
    rt_camera c = fill_camera(render_context, cam, top, right, (near_clip), (far_clip))

TestRenderer::TestRenderer(const SizeN &res) : Renderer(new RenderContext(nullptr, FormatSettings(res, Vector2f::O))) {}
TestRenderer::~TestRenderer() { delete render_context; }
void TestRenderer::Render() {
    const auto &r = render_context->format_settings.resolution;
    for (size_t y = 0; y < r.height; ++y)
        for (size_t x = 0; x < r.width; ++x)
            (*render_context)(x, y) = Vector3f(float(x) / float(r.width), 1.0f, 1.0f);
}

void IntersectTestRenderer::Render() {
    RT_CAMERA(camera_of(render_context)->near_clipping_plane, camera_of(render_context)->far_clipping_plane);
    std::vector<float> mask((size_t)c.width * c.height);
    rt_aov_out o;
    std::memset(&o, 0, sizeof(o));
    o.mask = mask.data();
    check(rt_render_aov(render_context->device_scene, &c, RT_AOV_MASK, &o));
    expand(render_context, mask);
}

DepthBufferRenderer::DepthBufferRenderer(RenderContext *rc) : Renderer(rc) {
    near = rc->world->cameras.at(rc->camera_id).near_clipping_plane;
    far = rc->world->cameras.at(rc->camera_id).far_clipping_plane;
}

void DepthBufferRenderer::Render() {
    RT_CAMERA(near, far);
    std::vector<float> depth((size_t)c.width * c.height);
    rt_aov_out o;
    std::memset(&o, 0, sizeof(o));
    o.depth = depth.data();
    check(rt_render_aov(render_context->device_scene, &c, RT_AOV_DEPTH, &o));
    expand(render_context, depth);
}

void NormalRenderer::Render() {
    RT_CAMERA(camera_of(render_context)->near_clipping_plane, camera_of(render_context)->far_clipping_plane);
    rt_aov_out o;
    std::memset(&o, 0, sizeof(o));
    o.normal = reinterpret_cast<float *>(render_context->buffer);
    check(rt_render_aov(render_context->device_scene, &c, RT_AOV_NORMAL, &o));
}

void AlbedoRenderer::Render() {
    RT_CAMERA(camera_of(render_context)->near_clipping_plane, camera_of(render_context)->far_clipping_plane);
    rt_aov_out o;
    std::memset(&o, 0, sizeof(o));
    o.albedo = reinterpret_cast<float *>(render_context->buffer);
    check(rt_render_aov(render_context->device_scene, &c, RT_AOV_ALBEDO, &o));
}

void PathTracingRenderer::Render() {
    RT_CAMERA(camera_of(render_context)->near_clipping_plane, camera_of(render_context)->far_clipping_plane);
    if (iteration_count <= 0) return;
    rt_pt_params p;
    std::memset(&p, 0, sizeof(p));
    p.spp = (uint32_t)iteration_count;
    p.max_depth = max_depth;
    p.seed = seed;
    std::vector<Vector3f> frame((size_t)c.width * c.height);
    check(rt_render_pt(render_context->device_scene, &c, &p, reinterpret_cast<float *>(frame.data())));
    // RenderToy accumulates into the buffer (buffer += L/spp).
    for (size_t i = 0; i < frame.size(); ++i) render_context->buffer[i] += frame[i];
}

// ---------------------------------------------------------------- OBJ import

void OBJModelImporter::Import(World &world, const std::string &path) {
    std::ifstream fs(path);
    if (!fs) return;   // as RenderToy: a missing file leaves the world unchanged
    std::vector<Vector3f> vert{Vector3f::O}, norm{Vector3f::O};
    std::vector<Vector2f> uv{Vector2f::O};
    Mesh *cur = nullptr;
    auto flush = [&]() {
        if (cur) world.meshes.push_back(cur);
    };
    std::string line;
    while (std::getline(fs, line)) {
        std::istringstream ls(line);
        std::string id;
        if (!(ls >> id)) continue;
        if (id == "o") {
            flush();
            cur = new Mesh();
            ls >> cur->name;
        } else if (id == "v") {
            Vector3f p;
            ls >> p;
            vert.push_back(p);
        } else if (id == "vn") {
            Vector3f p;
            ls >> p;
            norm.push_back(p);
        } else if (id == "vt") {
            Vector2f p;
            ls >> p;
            uv.push_back(p);
        } else if (id == "s") {
            std::string a;
            ls >> a;
            if (!cur) cur = new Mesh();
            cur->smooth = (a != "0" && a != "off");
        } else if (id == "f") {
            if (!cur) cur = new Mesh();
            std::vector<Vector3f> pv, pn;
            std::vector<Vector2f> pt;
            std::string tok;
            while (ls >> tok) {
                int idx[3] = {0, 0, 0};
                bool has[3] = {false, false, false};
                size_t k = 0, start = 0;
                for (size_t i = 0; i <= tok.size() && k < 3; ++i) {
                    if (i == tok.size() || tok[i] == '/') {
                        if (i > start) {
                            try {
                                idx[k] = std::stoi(tok.substr(start, i - start));
                                has[k] = true;
                            } catch (...) {
                            }
                        }
                        ++k;
                        start = i + 1;
                    }
                }
                auto pick3 = [](const std::vector<Vector3f> &a, bool h, int i) {
                    return (h && i > 0 && (size_t)i < a.size()) ? a[i] : Vector3f::O;
                };
                pv.push_back(pick3(vert, has[0], idx[0]));
                pn.push_back(pick3(norm, has[2], idx[2]));
                pt.push_back((has[1] && idx[1] > 0 && (size_t)idx[1] < uv.size()) ? uv[idx[1]] : Vector2f::O);
            }
            Polygon poly(pv.size(), pv, pn, pt, cur);
            auto tris = poly.ConvertToTriangle();
            world.triangles.insert(world.triangles.end(), tris.begin(), tris.end());
            cur->tris.insert(cur->tris.begin(), tris.begin(), tris.end());
        }
    }
    flush();
}

// ---------------------------------------------------------------- Image

namespace {
// A host Image's pixels on the device for one rt_image_* call.
struct DevImage {
    rt_image *h = nullptr;
    explicit DevImage(const Image &img) {
        check(rt_image_create((uint32_t)img.resolution.width, (uint32_t)img.resolution.height, &h));
        const int rc = rt_image_upload(h, reinterpret_cast<const float *>(img.GetBuffer()));
        if (rc != RT_OK) {
            rt_image_destroy(h);
            check(rc);
        }
    }
    void Download(Image &img) { check(rt_image_download(h, reinterpret_cast<float *>(img.GetBuffer()))); }
    ~DevImage() { rt_image_destroy(h); }
};

void require_same(const SizeN &a, const SizeN &b, const char *what) {
    if (a != b) throw Exception::ImageSizeNotMatchException(what);
}
}  // namespace

Image::Image(const SizeN &res) : buffer(new Vector3f[res.Area()]), resolution(res) {}
Image::Image(const RenderContext *const rc) : buffer(new Vector3f[rc->format_settings.resolution.Area()]),
                                              resolution(rc->format_settings.resolution) {
    std::copy(rc->buffer, rc->buffer + resolution.Area(), buffer);
}
Image::Image(const Image &o) : buffer(new Vector3f[o.resolution.Area()]), resolution(o.resolution) {
    std::copy(o.buffer, o.buffer + resolution.Area(), buffer);
}
Image::~Image() { delete[] buffer; }

const void Image::operator=(const Image &o) {
    require_same(resolution, o.resolution, "The = operator of two images requires two images have the same resolution.");
    std::copy(o.buffer, o.buffer + resolution.Area(), buffer);
}

#define RT_IMAGE_OP(OP, NAME)                                                                              \
    const Image Image::operator OP(const Image &o) {                                                     \
        require_same(resolution, o.resolution,                                                            \
                     "The " #OP " operator of two images requires two images have the same resolution."); \
        Image r(resolution);                                                                              \
        for (size_t i = 0; i < resolution.Area(); ++i) r.buffer[i] = buffer[i] OP o.buffer[i];            \
        return r;                                                                                         \
    }
RT_IMAGE_OP(+, add)
RT_IMAGE_OP(-, sub)
RT_IMAGE_OP(*, mul)
RT_IMAGE_OP(/, div)
#undef RT_IMAGE_OP

const Image Image::Extract(const std::function<bool(const Vector3f &)> &keep, const Vector3f &def) {
    Image r(resolution);
    for (size_t i = 0; i < resolution.Area(); ++i) r.buffer[i] = keep(buffer[i]) ? buffer[i] : def;
    return r;
}

const void Image::Transform(const std::function<void(Vector3f &)> &f) {
    for (size_t i = 0; i < resolution.Area(); ++i) f(buffer[i]);
}

void Image::DeviceConvolute(const std::vector<float> &w, std::size_t rows, std::size_t cols, unsigned flags) {
    DevImage d(*this);
    check(rt_image_convolve(d.h, (uint32_t)rows, (uint32_t)cols, w.data(), flags));
    d.Download(*this);
}

void Image::DeviceTonemap(float limit, int standard) {
    DevImage d(*this);
    check(rt_image_tonemap(d.h, limit, standard));
    d.Download(*this);
}

void Image::DeviceGreyScale(int standard) {
    DevImage d(*this);
    check(rt_image_greyscale(d.h, standard));
    d.Download(*this);
}

void Image::DeviceBinarize(int standard, float threshold) {
    DevImage d(*this);
    check(rt_image_binarize(d.h, standard, threshold));
    d.Download(*this);
}

void Image::DeviceEdgeDetection(int orientation) {
    DevImage d(*this);
    check(rt_image_edge_detect(d.h, orientation));
    d.Download(*this);
}

Image &Image::GaussianBlur(const std::size_t size, const float sigma) {
    DevImage d(*this);
    check(rt_image_gaussian_blur(d.h, (uint32_t)size, sigma));
    d.Download(*this);
    return *this;
}

Image &Image::Bloom(const std::size_t size, const float sigma, const float threshold) {
    DevImage d(*this);
    check(rt_image_bloom(d.h, (uint32_t)size, sigma, threshold));
    d.Download(*this);
    return *this;
}

Image &Image::GammaCorrection(const float gamma) {
    DevImage d(*this);
    check(rt_image_gamma(d.h, gamma));
    d.Download(*this);
    return *this;
}

// Writes go through the flat index y * width + x, kept inside the buffer: a
// point clamped to x = width lands on the next row's first pixel, as in RenderToy.
void Image::DrawLine(PointN p1, PointN p2, const Vector3f &color) {
    p1.SizeClamp(resolution);
    p2.SizeClamp(resolution);
    const long W = (long)resolution.width, n = (long)resolution.Area();
    auto put = [&](long y, long x) {
        const long i = y * W + x;
        if (i >= 0 && i < n) buffer[i] = color;
    };
    if (std::abs(p2.y - p1.y) < std::abs(p2.x - p1.x)) {   // x-major: step x, move y on error
        if (p1.x > p2.x) std::swap(p1, p2);
        const int ddy = std::abs(p2.y - p1.y), sy = p2.y - p1.y < 0 ? -1 : 1, ddx = p2.x - p1.x;
        int err = 0, y = p1.y;
        for (int x = p1.x; x < p2.x; ++x) {
            if (err > ddx) {
                y += sy;
                err -= 2 * ddx;
            }
            put(y, x);
            err += 2 * ddy;
        }
    } else {   // y-major
        if (p1.y > p2.y) std::swap(p1, p2);
        const int ddx = std::abs(p2.x - p1.x), sx = p2.x - p1.x < 0 ? -1 : 1, ddy = p2.y - p1.y;
        int err = 0, x = p1.x;
        for (int y = p1.y; y < p2.y; ++y) {
            if (err > ddy) {
                x += sx;
                err -= 2 * ddy;
            }
            put(y, x);
            err += 2 * ddx;
        }
    }
}

void Image::Fill(PointN p1, PointN p2, const Vector3f &color) {
    p1.SizeClamp(resolution);
    p2.SizeClamp(resolution);
    if (p1.x > p2.x) std::swap(p1.x, p2.x);
    if (p1.y > p2.y) std::swap(p1.y, p2.y);
    for (int x = p1.x; x < p2.x; ++x)
        for (int y = p1.y; y < p2.y; ++y) buffer[(size_t)y * resolution.width + x] = color;
}

// ---------------------------------------------------------------- exporters

namespace {
void export_with(const Image &img, int format, std::ostream &os) {
    DevImage d(img);
    size_t n = 0;
    check(rt_image_export(d.h, format, nullptr, 0, &n));
    std::vector<uint8_t> bytes(n);
    check(rt_image_export(d.h, format, bytes.data(), bytes.size(), &n));
    os.write(reinterpret_cast<const char *>(bytes.data()), (std::streamsize)n);
}
}  // namespace

void PPMExporter::Export(std::ostream &os) { export_with(image, RT_EXPORT_PPM, os); }
void BMPExporter::Export(std::ostream &os) { export_with(image, RT_EXPORT_BMP, os); }
void ASCIIExporter::Export(std::ostream &os) { export_with(image, RT_EXPORT_ASCII, os); }

}  // namespace RenderToy
