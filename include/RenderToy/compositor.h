// RenderToy/compositor.h -- RenderToy's Image (include/RenderToy/compositor.h in
// RenderToy).  The buffer stays on the host, as RenderToy's; convolutions,
// blur, bloom, tonemapping, gamma, greyscale, binarisation and edge detection
// run on the device through rt_image_* (bit-identical to RenderToy wherever it
// reads inside its buffer; see include/rendertoy_amd.h).  Pixel drawing,
// Extract/Transform with user callbacks and the arithmetic operators run here.
#ifndef RENDERTOY_AMD_COMPOSITOR_H
#define RENDERTOY_AMD_COMPOSITOR_H

#include <functional>

#include "renderer.h"
#include "rtmath.h"

namespace RenderToy {

enum class EdgeMode { kReflect = 0, kBlackPadding = 1, kWhitePadding = 2 };

class Image {
    Vector3f *buffer;

    // device-backed operations (rendertoy_api.cpp)
    void DeviceConvolute(const std::vector<float> &w, std::size_t rows, std::size_t cols, unsigned flags);
    void DeviceTonemap(float limit, int standard);
    void DeviceGreyScale(int standard);
    void DeviceBinarize(int standard, float threshold);
    void DeviceEdgeDetection(int orientation);

  public:
    enum class SampleMode { kNearestNeighbor = 0, kBilinearInterpolation = 1 };

    const SizeN resolution;

    Image(const SizeN &resolution_);
    Image(const RenderContext *const render_context_);
    Image(const Image &image);
    ~Image();

    Vector3f &operator()(const std::size_t x, const std::size_t y) { return buffer[y * resolution.width + x]; }
    const Vector3f &operator()(const std::size_t x, const std::size_t y) const {
        return buffer[y * resolution.width + x];
    }
    Vector3f *&GetBuffer() { return buffer; }
    const Vector3f *const &GetBuffer() const { return const_cast<const Vector3f *const &>(buffer); }

    [[nodiscard]] const Image Extract(const std::function<bool(const Vector3f &)> &pixel_filter,
                                      const Vector3f &default_color = Vector3f::O);
    const void Transform(const std::function<void(Vector3f &)> &pixel_processor);

    // Image::Convolute<_Normalize, _Length, _Abs> (compositor.h:107-146)
    template <bool _Normalize = false, bool _Length = false, bool _Abs = false>
    Image &Convolute(const IConvolutionKernel<float> &k) {
        std::vector<float> w;
        for (std::size_t i = 0; i < k.kernel_mat.row; ++i)
            for (std::size_t j = 0; j < k.kernel_mat.column; ++j) w.push_back(k.kernel_mat[i][j]);
        DeviceConvolute(w, k.kernel_mat.row, k.kernel_mat.column,
                        (_Normalize ? 1u : 0u) | (_Length ? 2u : 0u) | (_Abs ? 4u : 0u));
        return *this;
    }
    Image &GaussianBlur(const std::size_t size, const float sigma);
    template <Orientation _O>
    Image &EdgeDetection() {
        DeviceEdgeDetection(static_cast<int>(_O));
        return *this;
    }
    template <Convert::ColorStandard _CS>
    Image &GreyScale() {
        DeviceGreyScale(static_cast<int>(_CS));
        return *this;
    }
    template <Convert::ColorStandard _CS>
    Image &Binarize(const float threshold) {
        DeviceBinarize(static_cast<int>(_CS), threshold);
        return *this;
    }
    template <Convert::ColorStandard _CS = Convert::ColorStandard::kITURBT709>
    Image &Tonemap(const float limit = 1.5f) {
        DeviceTonemap(limit, static_cast<int>(_CS));
        return *this;
    }
    Image &GammaCorrection(const float gamma = 2.2f);

    const Image operator+(const Image &img);
    const Image operator-(const Image &img);
    const Image operator*(const Image &img);
    const Image operator/(const Image &img);
    const void operator=(const Image &img);

    Image &Bloom(const std::size_t size, const float sigma, const float threshold);

    // Bresenham lines and rectangle fills (compositor.cpp:163-252); the points
    // are clamped to [0, width] x [0, height] first, as RenderToy does.
    void DrawLine(PointN p1, PointN p2, const Vector3f &color = Vector3f::White);
    void Fill(PointN p1, PointN p2, const Vector3f &color = Vector3f::White);
};

}  // namespace RenderToy

#endif
