// RenderToy/compositor.h -- Image post-processing used by RenderToy's demos
// (include/RenderToy/compositor.h in RenderToy); host side, outside the GPU path.
#ifndef RENDERTOY_AMD_COMPOSITOR_H
#define RENDERTOY_AMD_COMPOSITOR_H

#include <functional>

#include "renderer.h"
#include "rtmath.h"

namespace RenderToy {

class Image {
    Vector3f *buffer;

  public:
    const SizeN resolution;

    Image(const SizeN &resolution_);
    Image(const RenderContext *const render_context_);
    Image(const Image &image);
    ~Image();

    Vector3f &operator()(const std::size_t x, const std::size_t y) { return buffer[y * resolution.width + x]; }
    const Vector3f &operator()(const std::size_t x, const std::size_t y) const {
        return buffer[y * resolution.width + x];
    }
    Vector3f *GetBuffer() { return buffer; }
    const Vector3f *GetBuffer() const { return buffer; }

    [[nodiscard]] Image Extract(const std::function<bool(const Vector3f &)> &keep,
                                const Vector3f &default_color = Vector3f::O) const;
    void Transform(const std::function<void(Vector3f &)> &f);
    // Separable normalised Gaussian of `size` taps, reflected edges.
    Image &GaussianBlur(const std::size_t size, const float sigma);
    // Adds a blurred copy of the pixels whose luminance exceeds `threshold`.
    Image &Bloom(const std::size_t size, const float sigma, const float threshold);
    Image &Tonemap(const float limit = 1.5f);
    Image &GammaCorrection(const float gamma = 2.2f);
    Image operator+(const Image &img) const;
    Image &operator=(const Image &img);
};

}  // namespace RenderToy

#endif
