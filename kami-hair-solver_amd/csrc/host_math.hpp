// FP64 vector and frame helpers of the rod model build. They also compile
// for the device (the root targets of uploaded frames are computed there
// with the same operations).
#pragma once

#include <algorithm>
#include <cmath>

#if defined(__HIP__)
#define KHS_HD __host__ __device__
#else
#define KHS_HD
#endif

namespace khs {

struct Vec3 {
    double x = 0.0, y = 0.0, z = 0.0;
    KHS_HD double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};

KHS_HD inline Vec3 operator+(Vec3 a, Vec3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
KHS_HD inline Vec3 operator-(Vec3 a, Vec3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
KHS_HD inline Vec3 operator*(double s, Vec3 a) { return {s * a.x, s * a.y, s * a.z}; }
KHS_HD inline Vec3 operator*(Vec3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
KHS_HD inline Vec3 over(Vec3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
KHS_HD inline double dot(Vec3 a, Vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
KHS_HD inline Vec3 cross(Vec3 a, Vec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
KHS_HD inline double length2(Vec3 a) { return dot(a, a); }
KHS_HD inline double length(Vec3 a) { return std::sqrt(dot(a, a)); }
KHS_HD inline bool is_finite(Vec3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
KHS_HD inline Vec3 unit(Vec3 a)
{
    const double n2 = length2(a);
    return n2 > 0.0 ? over(a, std::sqrt(n2)) : a;
}

// Row-major 3x3; the columns of a rod frame are its directors d1, d2, d3.
struct Mat3 {
    double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    KHS_HD double operator()(int r, int c) const { return m[3 * r + c]; }
    KHS_HD double &operator()(int r, int c) { return m[3 * r + c]; }
    KHS_HD Vec3 column(int c) const { return {m[c], m[3 + c], m[6 + c]}; }
};

KHS_HD inline Mat3 with_columns(Vec3 a, Vec3 b, Vec3 c)
{
    Mat3 r;
    const Vec3 cols[3] = {a, b, c};
    for (int k = 0; k < 3; ++k) {
        r(0, k) = cols[k].x;
        r(1, k) = cols[k].y;
        r(2, k) = cols[k].z;
    }
    return r;
}

KHS_HD inline Mat3 transposed(const Mat3 &a)
{
    Mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = a(j, i);
    return r;
}

KHS_HD inline Mat3 product(const Mat3 &a, const Mat3 &b)
{
    Mat3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}

KHS_HD inline Vec3 apply(const Mat3 &a, Vec3 v)
{
    return {a(0, 0) * v.x + a(0, 1) * v.y + a(0, 2) * v.z, a(1, 0) * v.x + a(1, 1) * v.y + a(1, 2) * v.z,
            a(2, 0) * v.x + a(2, 1) * v.y + a(2, 2) * v.z};
}

// 2 tan(theta/2) axis of a rotation from its skew part and trace
// (reference src/math.hpp:14-24).
KHS_HD inline Vec3 cayley_vector(const Mat3 &q)
{
    const Vec3 v{(q(2, 1) - q(1, 2)) * 0.5, (q(0, 2) - q(2, 0)) * 0.5, (q(1, 0) - q(0, 1)) * 0.5};
    const double c = (q(0, 0) + q(1, 1) + q(2, 2) - 1.0) * 0.5;
    return (2.0 / (1.0 + c + 1.0e-12)) * v;
}

// Orthonormal frame with d3 along the tangent and d1 from the hint
// (reference src/math.hpp:26-42).
KHS_HD inline Mat3 tangent_frame(Vec3 tangent, Vec3 hint)
{
    const Vec3 d3 = unit(tangent);
    Vec3 d1 = hint - d3 * dot(hint, d3);
    if (length2(d1) < 1.0e-16) {
        const Vec3 axis = std::fabs(d3.x) < 0.8 ? Vec3{1, 0, 0} : Vec3{0, 1, 0};
        d1 = axis - d3 * dot(axis, d3);
    }
    d1 = unit(d1);
    const Vec3 d2 = unit(cross(d3, d1));
    d1 = unit(cross(d2, d3));
    return with_columns(d1, d2, d3);
}

// A unit vector orthogonal to s, as Eigen 3.4's unitOrthogonal picks it.
KHS_HD inline Vec3 orthogonal_unit(Vec3 s)
{
    const double eps = 1.0e-12;
    if (!(std::fabs(s.x) <= std::fabs(s.z) * eps) || !(std::fabs(s.y) <= std::fabs(s.z) * eps)) {
        const double inv = 1.0 / std::sqrt(s.x * s.x + s.y * s.y);
        return {-s.y * inv, s.x * inv, 0.0};
    }
    const double inv = 1.0 / std::sqrt(s.y * s.y + s.z * s.z);
    return {0.0, -s.z * inv, s.y * inv};
}

// Minimal rotation of a director from one tangent to the next
// (reference src/math.hpp:44-57).
KHS_HD inline Vec3 transport(Vec3 director, Vec3 from, Vec3 to)
{
    const Vec3 a = unit(from), b = unit(to);
    const Vec3 axis = cross(a, b);
    const double s = length(axis);
    const double c = std::clamp(dot(a, b), -1.0, 1.0);
    if (s < 1.0e-12) {
        if (c > 0.0) return director;
        const Vec3 n = orthogonal_unit(a);
        return 2.0 * n * dot(n, director) - director;
    }
    const Vec3 n = over(axis, s);
    return director * c + cross(n, director) * s + n * dot(n, director) * (1.0 - c);
}

// Axis-angle vector of a rotation through its quaternion (largest-pivot
// extraction, angle 2 atan2(|v|, |w|)), as Eigen 3.4's AngleAxisd computes
// it (reference src/solver.cpp:19-26).
KHS_HD inline Vec3 axis_angle(const Mat3 &r)
{
    double q[4];  // x y z w
    const double tr = r(0, 0) + r(1, 1) + r(2, 2);
    if (tr > 0.0) {
        double t = std::sqrt(tr + 1.0);
        q[3] = 0.5 * t;
        t = 0.5 / t;
        q[0] = (r(2, 1) - r(1, 2)) * t;
        q[1] = (r(0, 2) - r(2, 0)) * t;
        q[2] = (r(1, 0) - r(0, 1)) * t;
    } else {
        int i = r(1, 1) > r(0, 0) ? 1 : 0;
        if (r(2, 2) > r(i, i)) i = 2;
        const int j = (i + 1) % 3, k = (j + 1) % 3;
        double t = std::sqrt(r(i, i) - r(j, j) - r(k, k) + 1.0);
        q[i] = 0.5 * t;
        t = 0.5 / t;
        q[3] = (r(k, j) - r(j, k)) * t;
        q[j] = (r(j, i) + r(i, j)) * t;
        q[k] = (r(k, i) + r(i, k)) * t;
    }
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    if (n == 0.0) return {};
    const double angle = 2.0 * std::atan2(n, std::fabs(q[3]));
    if (q[3] < 0.0) n = -n;
    const Vec3 axis{q[0] / n, q[1] / n, q[2] / n};
    if (!std::isfinite(angle) || !is_finite(axis)) return {};
    return axis * (angle > 3.14159265358979323846 ? angle - 2.0 * 3.14159265358979323846 : angle);
}

}  // namespace khs
