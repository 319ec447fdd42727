// Small fixed-size vector and matrix types of the host classes (include/tsdf,
// include/tracking, include/meshing).
//
// With Eigen on the include path (the reference builds with Eigen 3.4) they are Eigen's
// own types, so code written against the reference's classes (Eigen::Matrix4f poses,
// Eigen::Vector3f origins and vertices, r.pose.block<3, 3>(0, 0), ...) compiles
// unchanged. Without Eigen (define AZU_NO_EIGEN to force this) they are minimal
// stand-ins with the same storage (column-major matrices, packed xyz vectors) and the
// subset of Eigen's interface the host classes and their callers use: element access
// (v.x(), v(i), v[i], M(r, c)), data(), Zero(), Identity(), vector arithmetic and the
// 4x4 product.
#pragma once

#if !defined(AZU_NO_EIGEN) && defined(__has_include)
#if __has_include(<Eigen/Core>)
#define AZU_WITH_EIGEN 1
#endif
#endif

#ifdef AZU_WITH_EIGEN
#include <Eigen/Core>

namespace kfusion {
using Vec3f = Eigen::Vector3f;
using Vec3i = Eigen::Vector3i;
using Mat4f = Eigen::Matrix4f;
using Mat6f = Eigen::Matrix<float, 6, 6>;
}  // namespace kfusion

#else

namespace kfusion {

template <class T>
struct Vec3 {
    T v[3] = {T(0), T(0), T(0)};
    constexpr Vec3() = default;
    constexpr Vec3(T x, T y, T z) : v{x, y, z} {}
    static constexpr Vec3 Zero() { return Vec3(); }
    T& x() { return v[0]; }
    T& y() { return v[1]; }
    T& z() { return v[2]; }
    constexpr T x() const { return v[0]; }
    constexpr T y() const { return v[1]; }
    constexpr T z() const { return v[2]; }
    T& operator()(int i) { return v[i]; }
    constexpr T operator()(int i) const { return v[i]; }
    T& operator[](int i) { return v[i]; }
    constexpr T operator[](int i) const { return v[i]; }
    T* data() { return v; }
    const T* data() const { return v; }
    friend Vec3 operator+(const Vec3& a, const Vec3& b) { return {a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]}; }
    friend Vec3 operator-(const Vec3& a, const Vec3& b) { return {a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]}; }
    friend Vec3 operator*(const Vec3& a, T s) { return {a.v[0] * s, a.v[1] * s, a.v[2] * s}; }
    friend Vec3 operator*(T s, const Vec3& a) { return {a.v[0] * s, a.v[1] * s, a.v[2] * s}; }
    friend Vec3 operator/(const Vec3& a, T s) { return {a.v[0] / s, a.v[1] / s, a.v[2] / s}; }
    friend bool operator==(const Vec3& a, const Vec3& b) {
        return a.v[0] == b.v[0] && a.v[1] == b.v[1] && a.v[2] == b.v[2];
    }
};
using Vec3f = Vec3<float>;
using Vec3i = Vec3<int>;
static_assert(sizeof(Vec3f) == 12, "packed xyz, as Eigen::Vector3f");

// Column-major R x C float matrix (Eigen's default storage).
template <int R, int C>
struct Mat {
    float m[R * C] = {};
    static Mat Zero() { return Mat(); }
    static Mat Identity() {
        Mat a;
        for (int i = 0; i < (R < C ? R : C); ++i) a.m[i * R + i] = 1.0f;
        return a;
    }
    float& operator()(int r, int c) { return m[c * R + r]; }
    float operator()(int r, int c) const { return m[c * R + r]; }
    float* data() { return m; }
    const float* data() const { return m; }
    friend bool operator==(const Mat& a, const Mat& b) {
        for (int i = 0; i < R * C; ++i)
            if (a.m[i] != b.m[i]) return false;
        return true;
    }
};
using Mat4f = Mat<4, 4>;
using Mat6f = Mat<6, 6>;

inline Mat4f operator*(const Mat4f& a, const Mat4f& b) {
    Mat4f c;
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            float s = 0.0f;
            for (int k = 0; k < 4; ++k) s += a(i, k) * b(k, j);
            c(i, j) = s;
        }
    return c;
}

}  // namespace kfusion

#endif
