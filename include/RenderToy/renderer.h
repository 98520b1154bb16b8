// RenderToy/renderer.h -- RenderContext and the renderers (include/RenderToy/renderer.h
// in RenderToy).  Render() runs on the MI355X through the C ABI in rendertoy_amd.h;
// RenderContext::buffer is filled in the same layout (row-major, y = 0 at the bottom).
#ifndef RENDERTOY_AMD_RENDERER_H
#define RENDERTOY_AMD_RENDERER_H

#include <cstdint>

#include "rtmath.h"
#include "world.h"

struct rt_scene;

namespace RenderToy {

enum class RenderTech { kTest = 0, kIntersectTest, kDepthBuffer, kNormal, kPathTracing, kAlbedo, kLineEdge };
enum class ProjectionMode { kPerspective = 0, kParallel };

struct FormatSettings {
    SizeN resolution = SizeN(1920, 1080);
    Vector2f aspect = Vector2f(1.0f, 1.0f);
    FormatSettings() = default;
    FormatSettings(SizeN resolution_, Vector2f aspect_) : resolution(resolution_), aspect(aspect_) {}
};

// Owns the framebuffer and, where RenderToy builds its BVH, the device scene:
// the World is flattened and uploaded once, at construction.
struct RenderContext {
    World *world;
    Vector3f *buffer;
    FormatSettings format_settings;
    int camera_id = 0;
    rt_scene *device_scene = nullptr;

    RenderContext(World *world_, FormatSettings format_settings_);
    ~RenderContext();
    RenderContext(const RenderContext &) = delete;
    RenderContext &operator=(const RenderContext &) = delete;

    Vector3f &operator()(const std::size_t x, const std::size_t y) { return buffer[y * format_settings.resolution.width + x]; }
    const Vector3f &operator()(const std::size_t x, const std::size_t y) const {
        return buffer[y * format_settings.resolution.width + x];
    }
};

class Renderer {
  protected:
    void PrepareScreenSpace(Camera *cam, float &top, float &right);

  public:
    RenderContext *render_context;
    Renderer() = delete;
    Renderer(const Renderer &) = delete;
    Renderer(const Renderer &&) = delete;
    Renderer(RenderContext *render_context_) : render_context(render_context_) {}
    virtual ~Renderer() = default;
    virtual void Render() = 0;
};

class TestRenderer : public Renderer {
  public:
    TestRenderer(const SizeN &resolution_);
    ~TestRenderer() override;
    void Render() override final;
};

class IntersectTestRenderer : public Renderer {
  public:
    IntersectTestRenderer(RenderContext *rc) : Renderer(rc) {}
    void Render() override final;
};

class DepthBufferRenderer : public Renderer {
  public:
    float near, far;
    DepthBufferRenderer(RenderContext *rc);
    DepthBufferRenderer(RenderContext *rc, float near_, float far_) : Renderer(rc), near(near_), far(far_) {}
    void Render() override final;
};

class NormalRenderer : public Renderer {
  public:
    NormalRenderer(RenderContext *rc) : Renderer(rc) {}
    void Render() override final;
};

class PathTracingRenderer : public Renderer {
    int iteration_count;

  public:
    // max_depth counts path vertices (RenderToy stops after vertex 4, i.e. 5);
    // seed selects the counter-based random stream.
    uint32_t max_depth = 5;
    uint64_t seed = 0;
    PathTracingRenderer(RenderContext *rc, const int iteration_count_, uint32_t max_depth_ = 5, uint64_t seed_ = 0)
        : Renderer(rc), iteration_count(iteration_count_), max_depth(max_depth_), seed(seed_) {}
    void Render() override final;
};

class AlbedoRenderer : public Renderer {
  public:
    AlbedoRenderer(RenderContext *rc) : Renderer(rc) {}
    void Render() override final;
};

}  // namespace RenderToy

#endif
