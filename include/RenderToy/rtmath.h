// RenderToy/rtmath.h -- vectors, matrices and unit helpers of the RenderToy API
// (include/RenderToy/rtmath.h in RenderToy), host side.
//
// Arithmetic keeps RenderToy's operation order: Dot accumulates from zero in
// component order, Length is sqrt(Length2), Normalize divides by the length,
// Matrix*Vector is row dots and Matrix*Matrix accumulates k = 0..N-1 from zero.
// Cameras built with these types therefore produce the same float bits as in
// RenderToy, and those bits reach the GPU unchanged.
#ifndef RENDERTOY_AMD_RTMATH_H
#define RENDERTOY_AMD_RTMATH_H

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <functional>
#include <initializer_list>
#include <istream>
#include <type_traits>
#include <utility>
#include <vector>

namespace RenderToy {

template <typename T = float>
inline constexpr T kPi = T(3.14159265358979323846L);
inline constexpr float kFloatEpsilon = 1e-7f;
inline constexpr float kFloatInfinity = 1e7f;

enum class Orientation { X = 0, Y = 1, All = 2 };

template <typename T, std::size_t N>
struct Vector;

template <typename T, std::size_t N>
struct VectorConstants {};

template <typename T, std::size_t N>
struct Vector : public VectorConstants<T, N> {
    std::array<T, N> arr{};

    constexpr Vector() = default;
    template <typename... A, std::enable_if_t<(sizeof...(A) == N && N > 1), bool> = true>
    constexpr Vector(A &&...v) : arr{static_cast<T>(std::forward<A>(v))...} {}
    Vector(const T &v) { arr.fill(v); }

    constexpr T &operator[](std::size_t i) { return arr[i]; }
    constexpr const T &operator[](std::size_t i) const { return arr[i]; }
    constexpr T &x() { return arr[0]; }
    constexpr const T &x() const { return arr[0]; }
    template <std::size_t M = N, std::enable_if_t<(M > 1), bool> = true>
    constexpr T &y() { return arr[1]; }
    template <std::size_t M = N, std::enable_if_t<(M > 1), bool> = true>
    constexpr const T &y() const { return arr[1]; }
    template <std::size_t M = N, std::enable_if_t<(M > 2), bool> = true>
    constexpr T &z() { return arr[2]; }
    template <std::size_t M = N, std::enable_if_t<(M > 2), bool> = true>
    constexpr const T &z() const { return arr[2]; }
    template <std::size_t M = N, std::enable_if_t<(M > 3), bool> = true>
    constexpr T &w() { return arr[3]; }
    template <std::size_t M = N, std::enable_if_t<(M > 3), bool> = true>
    constexpr const T &w() const { return arr[3]; }

    Vector operator+(const Vector &b) const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = arr[i] + b.arr[i]; return r; }
    Vector operator-(const Vector &b) const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = arr[i] - b.arr[i]; return r; }
    Vector operator-() const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = -arr[i]; return r; }
    Vector operator*(const T s) const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = arr[i] * s; return r; }
    Vector operator*(const Vector &b) const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = arr[i] * b.arr[i]; return r; }
    Vector operator/(const T s) const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = arr[i] / s; return r; }
    Vector operator/(const Vector &b) const { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = arr[i] / b.arr[i]; return r; }
    Vector &operator+=(const Vector &b) { for (std::size_t i = 0; i < N; ++i) arr[i] += b.arr[i]; return *this; }
    Vector &operator-=(const Vector &b) { for (std::size_t i = 0; i < N; ++i) arr[i] -= b.arr[i]; return *this; }
    Vector &operator*=(const T s) { for (auto &e : arr) e *= s; return *this; }
    Vector &operator/=(const T s) { for (auto &e : arr) e /= s; return *this; }
    friend Vector operator*(const float s, const Vector &a) { Vector r(a); for (auto &e : r.arr) e *= s; return r; }

    // Component-wise |a-b| <= 1e-7, as RenderToy compares vectors.
    bool operator==(const Vector &b) const {
        for (std::size_t i = 0; i < N; ++i)
            if (std::abs(arr[i] - b.arr[i]) > kFloatEpsilon) return false;
        return true;
    }
    bool operator!=(const Vector &b) const { return !(*this == b); }

    static T Dot(const Vector &a, const Vector &b) { T r(0); for (std::size_t i = 0; i < N; ++i) r += a.arr[i] * b.arr[i]; return r; }
    T Dot(const Vector &b) const { return Dot(*this, b); }
    template <std::size_t M = N, std::enable_if_t<(M == 3), bool> = true>
    static Vector Cross(const Vector &a, const Vector &b) {
        return Vector(a.arr[1] * b.arr[2] - a.arr[2] * b.arr[1], a.arr[2] * b.arr[0] - a.arr[0] * b.arr[2],
                      a.arr[0] * b.arr[1] - a.arr[1] * b.arr[0]);
    }
    template <std::size_t M = N, std::enable_if_t<(M == 3), bool> = true>
    Vector Cross(const Vector &b) const { return Cross(*this, b); }
    T Length2() const { T r(0); for (T a : arr) r += a * a; return r; }
    T Length() const { return std::sqrt(Length2()); }
    void Normalize() { T l = Length(); for (T &a : arr) a /= l; }
    Vector Normalized() const { Vector r(*this); r.Normalize(); return r; }
    bool IsZero() const { for (T a : arr) if (std::abs(a) > kFloatEpsilon) return false; return true; }
    static Vector Abs(const Vector &a) { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = std::abs(a.arr[i]); return r; }
    static Vector Pow(const Vector &a, const Vector &b) { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = std::pow(a.arr[i], b.arr[i]); return r; }
    static Vector Log(const Vector &a) { Vector r; for (std::size_t i = 0; i < N; ++i) r.arr[i] = std::log(a.arr[i]); return r; }
    friend std::istream &operator>>(std::istream &is, Vector &a) { for (auto &e : a.arr) is >> e; return is; }
};

template <typename T> using Vector2 = Vector<T, 2>;
template <typename T> using Vector3 = Vector<T, 3>;
template <typename T> using Vector4 = Vector<T, 4>;
using Vector2f = Vector2<float>;
using Vector3f = Vector3<float>;
using Vector4f = Vector4<float>;
using Vector2d = Vector2<double>;
using Vector3d = Vector3<double>;
using Vector4d = Vector4<double>;

template <typename T>
struct VectorConstants<T, 2> { static const Vector<T, 2> O, X, Y; };
template <typename T>
struct VectorConstants<T, 3> { static const Vector<T, 3> O, X, Y, Z, White; };
template <typename T>
struct VectorConstants<T, 4> { static const Vector<T, 4> O, X, Y, Z, W; };
template <typename T> inline const Vector<T, 2> VectorConstants<T, 2>::O{T(0), T(0)};
template <typename T> inline const Vector<T, 2> VectorConstants<T, 2>::X{T(1), T(0)};
template <typename T> inline const Vector<T, 2> VectorConstants<T, 2>::Y{T(0), T(1)};
template <typename T> inline const Vector<T, 3> VectorConstants<T, 3>::O{T(0), T(0), T(0)};
template <typename T> inline const Vector<T, 3> VectorConstants<T, 3>::X{T(1), T(0), T(0)};
template <typename T> inline const Vector<T, 3> VectorConstants<T, 3>::Y{T(0), T(1), T(0)};
template <typename T> inline const Vector<T, 3> VectorConstants<T, 3>::Z{T(0), T(0), T(1)};
template <typename T> inline const Vector<T, 3> VectorConstants<T, 3>::White{T(1), T(1), T(1)};
template <typename T> inline const Vector<T, 4> VectorConstants<T, 4>::O{T(0), T(0), T(0), T(0)};
template <typename T> inline const Vector<T, 4> VectorConstants<T, 4>::X{T(1), T(0), T(0), T(0)};
template <typename T> inline const Vector<T, 4> VectorConstants<T, 4>::Y{T(0), T(1), T(0), T(0)};
template <typename T> inline const Vector<T, 4> VectorConstants<T, 4>::Z{T(0), T(0), T(1), T(0)};
template <typename T> inline const Vector<T, 4> VectorConstants<T, 4>::W{T(0), T(0), T(0), T(1)};

template <typename T, std::size_t N>
struct MatrixConstants {};

template <typename T, std::size_t N>
struct Matrix : public MatrixConstants<T, N> {
    using VectorType = Vector<T, N>;
    std::array<VectorType, N> row{};

    constexpr Matrix() = default;
    Matrix(std::initializer_list<VectorType> rows) {
        std::size_t i = 0;
        for (const auto &r : rows)
            if (i < N) row[i++] = r;
    }

    VectorType &operator[](std::size_t i) { return row[i]; }
    const VectorType &operator[](std::size_t i) const { return row[i]; }
    VectorType operator*(const VectorType &v) const {
        VectorType r;
        for (std::size_t i = 0; i < N; ++i) r[i] = row[i].Dot(v);
        return r;
    }
    Matrix operator*(const Matrix &b) const {
        Matrix r;
        for (std::size_t i = 0; i < N; ++i)
            for (std::size_t j = 0; j < N; ++j)
                for (std::size_t k = 0; k < N; ++k) r.row[i][j] += row[i][k] * b.row[k][j];
        return r;
    }
    Matrix operator+(const Matrix &b) const { Matrix r; for (std::size_t i = 0; i < N; ++i) r.row[i] = row[i] + b.row[i]; return r; }
    Matrix operator-(const Matrix &b) const { Matrix r; for (std::size_t i = 0; i < N; ++i) r.row[i] = row[i] - b.row[i]; return r; }
    bool operator==(const Matrix &b) const { for (std::size_t i = 0; i < N; ++i) if (row[i] != b.row[i]) return false; return true; }
    bool operator!=(const Matrix &b) const { return !(*this == b); }
    Matrix Transposed() const {
        Matrix r;
        for (std::size_t i = 0; i < N; ++i)
            for (std::size_t j = 0; j < N; ++j) r.row[i][j] = row[j][i];
        return r;
    }
    void Transpose() { *this = Transposed(); }
    template <std::size_t M = N, std::enable_if_t<(M >= 3), bool> = true>
    Matrix<T, N - 1> ComplementMinor(std::size_t i, std::size_t j) const {
        Matrix<T, N - 1> r;
        for (std::size_t a = 0, ra = 0; a < N; ++a) {
            if (a == i) continue;
            for (std::size_t b = 0, rb = 0; b < N; ++b) {
                if (b == j) continue;
                r[ra][rb++] = row[a][b];
            }
            ++ra;
        }
        return r;
    }
};

using Matrix2x2f = Matrix<float, 2>;
using Matrix3x3f = Matrix<float, 3>;
using Matrix4x4f = Matrix<float, 4>;

template <typename T, std::size_t N>
Matrix<T, N> IdentityMatrix() {
    Matrix<T, N> m;
    for (std::size_t i = 0; i < N; ++i) m[i][i] = T(1);
    return m;
}
template <typename T>
struct MatrixConstants<T, 2> { static const Matrix<T, 2> I, O; };
template <typename T>
struct MatrixConstants<T, 3> { static const Matrix<T, 3> I, O; };
template <typename T>
struct MatrixConstants<T, 4> { static const Matrix<T, 4> I, O; };
template <typename T> inline const Matrix<T, 2> MatrixConstants<T, 2>::I = IdentityMatrix<T, 2>();
template <typename T> inline const Matrix<T, 2> MatrixConstants<T, 2>::O{};
template <typename T> inline const Matrix<T, 3> MatrixConstants<T, 3>::I = IdentityMatrix<T, 3>();
template <typename T> inline const Matrix<T, 3> MatrixConstants<T, 3>::O{};
template <typename T> inline const Matrix<T, 4> MatrixConstants<T, 4>::I = IdentityMatrix<T, 4>();
template <typename T> inline const Matrix<T, 4> MatrixConstants<T, 4>::O{};

struct SizeN;

struct PointN {
    int x, y;
    PointN() : x(0), y(0) {}
    PointN(int x_, int y_) : x(x_), y(y_) {}
    PointN operator+(const PointN &p) const { return PointN(x + p.x, y + p.y); }
    // clamps to [0, width] x [0, height] (inclusive), as RenderToy's PointN::SizeClamp
    void SizeClamp(const SizeN &size);
};

struct SizeN {
    std::size_t width, height;
    SizeN() : width(0), height(0) {}
    SizeN(std::size_t w, std::size_t h) : width(w), height(h) {}
    std::size_t Area() const { return width * height; }
    float AspectRatio() const { return float(width) / float(height); }
    PointN Center() const { return PointN(static_cast<int>(width) >> 1, static_cast<int>(height) >> 1); }
    bool operator==(const SizeN &a) const { return width == a.width && height == a.height; }
    bool operator!=(const SizeN &a) const { return !(*this == a); }
};

inline void PointN::SizeClamp(const SizeN &size) {
    x = std::clamp(x, 0, static_cast<int>(size.width));
    y = std::clamp(y, 0, static_cast<int>(size.height));
}

// A rows x columns matrix of weights, indexed [row][column].
template <typename T>
struct GeneralizedMatrix {
    std::size_t row, column;
    std::vector<std::vector<T>> data;
    GeneralizedMatrix(std::size_t r = 2, std::size_t c = 2, T v = T(0)) : row(r), column(c), data(r, std::vector<T>(c, v)) {}
    GeneralizedMatrix(std::initializer_list<std::vector<T>> rows)
        : row(rows.size()), column(rows.begin()->size()), data(rows) {}
    std::vector<T> &operator[](std::size_t r) { return data[r]; }
    const std::vector<T> &operator[](std::size_t r) const { return data[r]; }
};

// Convolution kernels of Image::Convolute (RenderToy rtmath.h:975-1090).
template <typename T>
struct IConvolutionKernel {
    IConvolutionKernel() = default;
    IConvolutionKernel(std::size_t r, std::size_t c) : kernel_mat(r, c) {}
    IConvolutionKernel(const GeneralizedMatrix<T> &m) : kernel_mat(m) {}
    GeneralizedMatrix<T> kernel_mat;
};

// exp(-d^2 / (2 sigma sigma)) normalised to sum 1: size x size (All), 1 x size (X)
// or size x 1 (Y), centred at size / 2, in float with pow/exp as RenderToy.
template <typename T, Orientation O = Orientation::All>
struct UnitizedGaussianKernel : public IConvolutionKernel<T> {
    UnitizedGaussianKernel(const std::size_t size, const T sigma)
        : IConvolutionKernel<T>(O == Orientation::X ? 1 : size, O == Orientation::Y ? 1 : size) {
        const T orig = T(size >> 1);
        T sum = T(0);
        auto &m = this->kernel_mat;
        if constexpr (O == Orientation::All) {
            for (std::size_t i = 0; i < size; ++i) {
                const T y2 = std::pow(T(i) - orig, T(2));
                for (std::size_t j = 0; j < size; ++j) {
                    const T x2 = std::pow(T(j) - orig, T(2));
                    const T g = std::exp(-(x2 + y2) / (T(2) * sigma * sigma));
                    sum += g;
                    m[i][j] = g;
                }
            }
        } else {
            for (std::size_t j = 0; j < size; ++j) {
                const T d2 = std::pow(T(j) - orig, T(2));
                const T g = std::exp(-(d2) / (T(2) * sigma * sigma));
                sum += g;
                (O == Orientation::X ? m[0][j] : m[j][0]) = g;
            }
        }
        for (auto &r : m.data)
            for (auto &v : r) v /= sum;
    }
};

template <typename T, Orientation O>
struct SobelKernel : public IConvolutionKernel<T> {
    SobelKernel()
        : IConvolutionKernel<T>(O == Orientation::X ? GeneralizedMatrix<T>({{T(-1), T(0), T(1)}, {T(-2), T(0), T(2)},
                                                                             {T(-1), T(0), T(1)}})
                                                    : GeneralizedMatrix<T>({{T(1), T(2), T(1)}, {T(0), T(0), T(0)},
                                                                             {T(-1), T(-2), T(-1)}})) {}
};

template <typename T, Orientation O>
struct PrewittKernel : public IConvolutionKernel<T> {
    PrewittKernel()
        : IConvolutionKernel<T>(O == Orientation::X ? GeneralizedMatrix<T>({{T(-1), T(0), T(1)}, {T(-1), T(0), T(1)},
                                                                             {T(-1), T(0), T(1)}})
                                                    : GeneralizedMatrix<T>({{T(1), T(1), T(1)}, {T(0), T(0), T(0)},
                                                                             {T(-1), T(-1), T(-1)}})) {}
};

// x*(1-a) + y*a
template <typename T, typename C>
T Lerp(const T x, const T y, const C a) { return x * (C(1) - a) + y * a; }

namespace Convert {
template <typename P = float, typename R = float>
constexpr R InchToMM(const P inch) { return R(inch) * R(25.4); }
template <typename P = int, typename R = float>
constexpr R DegreeToRadians(const P deg) { return R(deg) * kPi<R> / R(180); }

enum class ColorStandard { kITURBT709 = 0, kITURBT601 = 1, kITURBT2020 = 2, kSMPTE240M = 3, kAverage = 4 };

template <ColorStandard CS = ColorStandard::kITURBT709>
constexpr float Luma(const Vector3f &c) {
    if constexpr (CS == ColorStandard::kITURBT709) return 0.2126f * c.x() + 0.7152f * c.y() + 0.0722f * c.z();
    else if constexpr (CS == ColorStandard::kITURBT601) return 0.299f * c.x() + 0.587f * c.y() + 0.114f * c.z();
    else if constexpr (CS == ColorStandard::kITURBT2020) return 0.2627f * c.x() + 0.6780f * c.y() + 0.0593f * c.z();
    else if constexpr (CS == ColorStandard::kSMPTE240M) return 0.2122f * c.x() + 0.7013f * c.y() + 0.0865f * c.z();
    else return (c.x() + c.y() + c.z()) / 3.0f;
}
// Reinhard operator on luminance.
template <ColorStandard CS = ColorStandard::kITURBT709>
Vector3f Tonemap(const Vector3f &c, const float limit) { return c * 1.0f / (1.0f + Luma<CS>(c) / limit); }

// Planckian-locus colour of a black body at t kelvin (Krystek's rational
// approximation in CIE 1960 UCS), normalised to a maximum component of 1.
Vector3f BlackBody(const float t);
Vector3f XYZToSRGB(const Vector3f &xyz);
}  // namespace Convert

namespace AffineTransformation {
Matrix4x4f Translation(const Vector3f &origin);
// Intrinsic XYZ Euler angles (radians): Rz * Ry * Rx.
Matrix4x4f RotationEulerXYZ(const Vector3f &euler_xyz);
}  // namespace AffineTransformation

}  // namespace RenderToy

#endif
