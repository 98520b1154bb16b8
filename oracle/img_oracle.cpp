// img_oracle.cpp -- CPU restatement of RenderToy's Image operations
// (include/RenderToy/compositor.h, src/compositor.cpp) and of the Gaussian
// kernel weights (include/RenderToy/rtmath.h:988-1060).
//
// TEST INFRASTRUCTURE ONLY: the checker for rendertoy_amd/csrc/image.hip.
// Images are float[height][width][3], row-major.  Single precision, one IEEE
// operation at a time, in the reference's order (built with -ffp-contract=off).
//
// Edge reads: Image::BufferGetWrapper<kReflect> (compositor.h:26-40) reflects
// an out-of-range coordinate once (-p, or 2n - p) and reads buffer[x*W + y].
// Coordinate n maps to n, so the reference reads one element past a row (the
// next row's first pixel) and, on the last row or column, past the buffer's
// end (undefined behaviour).  Here a flat index inside the buffer reads that
// pixel, as the reference does, and one outside it reads black.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct V3 {
    float x, y, z;
};

inline V3 load(const float *img, int64_t i) { return {img[3 * i], img[3 * i + 1], img[3 * i + 2]}; }

inline V3 reflect_read(const float *img, int W, int H, int px, int py) {
    // compositor.h:29-38: px is the row, py the column
    if (px < 0) px = -px;
    if (py < 0) py = -py;
    if (py >= W) py = 2 * W - py;
    if (px >= H) px = 2 * H - px;
    const int64_t flat = (int64_t)px * W + py;
    if (flat < 0 || flat >= (int64_t)W * H) return {0.0f, 0.0f, 0.0f};
    return load(img, flat);
}

// Convert::Luma (rtmath.h:1140-1166)
inline float luma(const V3 &v, int standard) {
    switch (standard) {
    case 1: return 0.299f * v.x + 0.587f * v.y + 0.114f * v.z;
    case 2: return 0.2627f * v.x + 0.6780f * v.y + 0.0593f * v.z;
    case 3: return 0.2122f * v.x + 0.7013f * v.y + 0.0865f * v.z;
    case 4: return (v.x + v.y + v.z) / 3.0f;
    default: return 0.2126f * v.x + 0.7152f * v.y + 0.0722f * v.z;
    }
}

// Image::Convolute<_Normalize, _Length, _Abs> (compositor.h:107-146): out = sum
// over taps in row-major order of tmp(i - rows/2 + ii, j - cols/2 + jj) * k[ii][jj].
void convolve(float *img, int W, int H, int rows, int cols, const float *k, int flags) {
    std::vector<float> tmp(img, img + (size_t)W * H * 3);
    const int r2 = rows >> 1, c2 = cols >> 1;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < H; ++i)
        for (int j = 0; j < W; ++j) {
            V3 acc{0.0f, 0.0f, 0.0f};
            for (int ii = 0; ii < rows; ++ii)
                for (int jj = 0; jj < cols; ++jj) {
                    const V3 v = reflect_read(tmp.data(), W, H, i - r2 + ii, j - c2 + jj);
                    const float w = k[ii * cols + jj];
                    acc.x = acc.x + v.x * w;
                    acc.y = acc.y + v.y * w;
                    acc.z = acc.z + v.z * w;
                }
            if (flags & 4) acc = {std::abs(acc.x), std::abs(acc.y), std::abs(acc.z)};
            if (flags & 2) {
                const float l = std::sqrt(0.0f + acc.x * acc.x + acc.y * acc.y + acc.z * acc.z);
                acc = {l, l, l};
            }
            if (flags & 1) {
                const float l = std::sqrt(0.0f + acc.x * acc.x + acc.y * acc.y + acc.z * acc.z);
                acc = {acc.x / l, acc.y / l, acc.z / l};
            }
            float *o = img + 3 * ((size_t)i * W + j);
            o[0] = acc.x;
            o[1] = acc.y;
            o[2] = acc.z;
        }
}

}  // namespace

extern "C" {

// UnitizedGaussianKernel (rtmath.h:988-1060): orient 0 = size x size, 1 = X
// (1 x size), 2 = Y (size x 1).  Weights exp(-d^2 / (2 sigma sigma)) / sum.
void orc_gaussian_weights(uint32_t size, float sigma, int orient, float *out) {
    const float orig = float(size >> 1);
    float sum = 0.0f;
    if (orient == 0) {
        for (uint32_t i = 0; i < size; ++i) {
            const float y2 = std::pow(float(i) - orig, 2.0f);
            for (uint32_t j = 0; j < size; ++j) {
                const float x2 = std::pow(float(j) - orig, 2.0f);
                const float g = std::exp(-(x2 + y2) / (2.0f * sigma * sigma));
                sum += g;
                out[i * size + j] = g;
            }
        }
        for (uint32_t i = 0; i < size * size; ++i) out[i] /= sum;
        return;
    }
    for (uint32_t j = 0; j < size; ++j) {
        const float d2 = std::pow(float(j) - orig, 2.0f);
        const float g = std::exp(-(d2) / (2.0f * sigma * sigma));
        sum += g;
        out[j] = g;
    }
    for (uint32_t j = 0; j < size; ++j) out[j] /= sum;
}

// flags: 1 normalise, 2 length, 4 abs (Convolute's template switches).
void orc_image_convolve(float *img, uint32_t W, uint32_t H, uint32_t rows, uint32_t cols, const float *k,
                        int flags) {
    convolve(img, (int)W, (int)H, (int)rows, (int)cols, k, flags);
}

// Image::GaussianBlur (compositor.cpp:42-49): X pass, then Y pass.
void orc_image_gaussian_blur(float *img, uint32_t W, uint32_t H, uint32_t size, float sigma) {
    std::vector<float> k(size);
    orc_gaussian_weights(size, sigma, 1, k.data());
    convolve(img, (int)W, (int)H, 1, (int)size, k.data(), 0);
    orc_gaussian_weights(size, sigma, 2, k.data());
    convolve(img, (int)W, (int)H, (int)size, 1, k.data(), 0);
}

// Image::Bloom (compositor.cpp:70-78): pixels with BT.709 luma above the
// threshold, blurred, added back.
void orc_image_bloom(float *img, uint32_t W, uint32_t H, uint32_t size, float sigma, float threshold) {
    const size_t n = (size_t)W * H;
    std::vector<float> layer(n * 3);
    for (size_t i = 0; i < n; ++i) {
        const V3 v = load(img, (int64_t)i);
        const bool keep = luma(v, 0) > threshold;
        layer[3 * i] = keep ? v.x : 0.0f;
        layer[3 * i + 1] = keep ? v.y : 0.0f;
        layer[3 * i + 2] = keep ? v.z : 0.0f;
    }
    orc_image_gaussian_blur(layer.data(), W, H, size, sigma);
    for (size_t i = 0; i < 3 * n; ++i) img[i] = img[i] + layer[i];
}

// Image::Tonemap (compositor.h:221-228, Convert::Tonemap rtmath.h:1172-1177):
// v * 1 / (1 + luma / limit).
void orc_image_tonemap(float *img, uint32_t W, uint32_t H, float limit, int standard) {
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const V3 v = load(img, (int64_t)i);
        const float s = 1.0f + luma(v, standard) / limit;
        img[3 * i] = (v.x * 1.0f) / s;
        img[3 * i + 1] = (v.y * 1.0f) / s;
        img[3 * i + 2] = (v.z * 1.0f) / s;
    }
}

// Image::GammaCorrection (compositor.cpp:95-101): pow(v, 1 / gamma) per channel.
void orc_image_gamma(float *img, uint32_t W, uint32_t H, float gamma) {
    const float ratio = 1.0f / gamma;
    for (size_t i = 0; i < (size_t)W * H * 3; ++i) img[i] = std::pow(img[i], ratio);
}

// Image::GreyScale / Binarize (compositor.h:184-216).
void orc_image_greyscale(float *img, uint32_t W, uint32_t H, int standard) {
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const float l = luma(load(img, (int64_t)i), standard);
        img[3 * i] = img[3 * i + 1] = img[3 * i + 2] = l;
    }
}

void orc_image_binarize(float *img, uint32_t W, uint32_t H, int standard, float threshold) {
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        const float b = luma(load(img, (int64_t)i), standard) > threshold ? 1.0f : 0.0f;
        img[3 * i] = img[3 * i + 1] = img[3 * i + 2] = b;
    }
}

// Image::EdgeDetection<_O> (compositor.h:154-180) with SobelKernel (rtmath.h:1080-1090):
// orient 1 = X, 2 = Y, 0 = All (|X result + Y result| as a length).
void orc_image_edge_detect(float *img, uint32_t W, uint32_t H, int orient) {
    static const float kx[9] = {-1, 0, 1, -2, 0, 2, -1, 0, 1};
    static const float ky[9] = {1, 2, 1, 0, 0, 0, -1, -2, -1};
    if (orient == 1 || orient == 2) {
        convolve(img, (int)W, (int)H, 3, 3, orient == 1 ? kx : ky, 0);
        return;
    }
    std::vector<float> cpy(img, img + (size_t)W * H * 3);
    convolve(cpy.data(), (int)W, (int)H, 3, 3, kx, 0);
    convolve(img, (int)W, (int)H, 3, 3, ky, 0);
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        V3 v = load(img, (int64_t)i);
        const V3 c = load(cpy.data(), (int64_t)i);
        v = {v.x + c.x, v.y + c.y, v.z + c.z};
        const float l = std::sqrt(0.0f + v.x * v.x + v.y * v.y + v.z * v.z);
        img[3 * i] = img[3 * i + 1] = img[3 * i + 2] = l;
    }
}

// Image::operator+ - * / (compositor.cpp:103-161): op 0..3, element-wise.
void orc_image_arith(const float *a, const float *b, float *out, uint32_t W, uint32_t H, int op) {
    for (size_t i = 0; i < (size_t)W * H * 3; ++i) {
        switch (op) {
        case 0: out[i] = a[i] + b[i]; break;
        case 1: out[i] = a[i] - b[i]; break;
        case 2: out[i] = a[i] * b[i]; break;
        default: out[i] = a[i] / b[i]; break;
        }
    }
}

}  // extern "C"
