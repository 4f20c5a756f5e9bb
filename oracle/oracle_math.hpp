// Oracle math: FP64 vectors, 3x3 matrices, forward-mode dual numbers and
// the SO(3) maps of the reference hair step.
//
// TEST INFRASTRUCTURE ONLY. The oracle is the CPU restatement the HIP
// library is checked against; product code never includes or links it.
//
// Every formula follows the cited reference line; the code is written from
// the formulas, not from the reference's source.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace oracle {

struct V3 {
    double x = 0.0, y = 0.0, z = 0.0;
};

inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
// Division by a scalar multiplies by its reciprocal, as the device code of
// the reference does (cuda_backend.cu:42).
inline V3 div_recip(V3 a, double s) { return a * (1.0 / s); }
// Componentwise division, as Eigen's operator/ does on the host side.
inline V3 div_each(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm2(V3 a) { return dot(a, a); }
inline double norm(V3 a) { return std::sqrt(norm2(a)); }
inline V3 vmin(V3 a, V3 b) { return {std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)}; }
inline bool finite3(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline double comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline void set_comp(V3 &a, int i, double v) { (i == 0 ? a.x : (i == 1 ? a.y : a.z)) = v; }

// Host-side normalization (Eigen normalized(): divide by sqrt(squaredNorm)).
inline V3 normalized(V3 a)
{
    const double n2 = norm2(a);
    if (n2 > 0.0) return div_each(a, std::sqrt(n2));
    return a;
}

// Row-major 3x3 matrix, m[3*row + col].
struct M3 {
    double m[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double &operator()(int r, int c) { return m[3 * r + c]; }
    double operator()(int r, int c) const { return m[3 * r + c]; }
    V3 col(int c) const { return {m[c], m[3 + c], m[6 + c]}; }
};

inline M3 transpose(const M3 &a)
{
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r(i, j) = a(j, i);
    return r;
}

inline M3 matmul(const M3 &a, const M3 &b)
{
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a(i, k) * b(k, j);
            r(i, j) = s;
        }
    return r;
}

inline M3 from_columns(V3 c0, V3 c1, V3 c2)
{
    M3 r;
    r(0, 0) = c0.x; r(1, 0) = c0.y; r(2, 0) = c0.z;
    r(0, 1) = c1.x; r(1, 1) = c1.y; r(2, 1) = c1.z;
    r(0, 2) = c2.x; r(1, 2) = c2.y; r(2, 2) = c2.z;
    return r;
}

inline V3 matvec(const M3 &a, V3 v)
{
    return {a(0, 0) * v.x + a(0, 1) * v.y + a(0, 2) * v.z,
            a(1, 0) * v.x + a(1, 1) * v.y + a(1, 2) * v.z,
            a(2, 0) * v.x + a(2, 1) * v.y + a(2, 2) * v.z};
}

// ---------------------------------------------------------------------------
// Forward-mode dual number with N tangent directions. The reference
// differentiates its rod strains and the barrier with such numbers
// (cuda_backend.cu:59-96); the oracle keeps that method so that the GPU's
// hand-derived Jacobians are checked against automatic differentiation.
template <int N>
struct Dual {
    double v;
    double d[N];
    Dual(double value = 0.0) : v(value) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    static Dual var(double value, int index)
    {
        Dual r(value);
        r.d[index] = 1.0;
        return r;
    }
};

template <int N> Dual<N> operator+(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N> Dual<N> operator-(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N> Dual<N> operator-(const Dual<N> &a)
{
    Dual<N> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
// Product rule d(ab) = a'b + ab' (cuda_backend.cu:77-79).
template <int N> Dual<N> operator*(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
// Quotient rule via the value of the quotient (cuda_backend.cu:80-83).
template <int N> Dual<N> operator/(const Dual<N> &a, const Dual<N> &b)
{
    const double inv = 1.0 / b.v;
    Dual<N> r(a.v * inv);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * inv;
    return r;
}
// sqrt with the derivative guarded at zero (cuda_backend.cu:84-87).
template <int N> Dual<N> dsqrt(const Dual<N> &a)
{
    const double s = std::sqrt(a.v);
    Dual<N> r(s);
    const double f = 0.5 / std::fmax(s, 1.0e-300);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * f;
    return r;
}
template <int N> Dual<N> dsin(const Dual<N> &a)
{
    Dual<N> r(std::sin(a.v));
    const double c = std::cos(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = c * a.d[i];
    return r;
}
template <int N> Dual<N> dcos(const Dual<N> &a)
{
    Dual<N> r(std::cos(a.v));
    const double s = -std::sin(a.v);
    for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
    return r;
}
template <int N> Dual<N> dlog(const Dual<N> &a)
{
    Dual<N> r(std::log(a.v));
    const double inv = 1.0 / a.v;
    for (int i = 0; i < N; ++i) r.d[i] = inv * a.d[i];
    return r;
}

template <typename S> struct TV3 { S x, y, z; };
template <typename S> TV3<S> tadd(const TV3<S> &a, const TV3<S> &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename S> TV3<S> tsub(const TV3<S> &a, const TV3<S> &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename S> TV3<S> tscale(const TV3<S> &a, const S &s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename S> S tdot(const TV3<S> &a, const TV3<S> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

template <typename S> struct TM3 { S m[9]; };
template <typename S> TM3<S> tidentity()
{
    TM3<S> r;
    for (int i = 0; i < 9; ++i) r.m[i] = S(0.0);
    r.m[0] = S(1.0); r.m[4] = S(1.0); r.m[8] = S(1.0);
    return r;
}
template <typename S> TM3<S> tmadd(const TM3<S> &a, const TM3<S> &b)
{
    TM3<S> r;
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] + b.m[i];
    return r;
}
template <typename S> TM3<S> tmscale(const TM3<S> &a, const S &s)
{
    TM3<S> r;
    for (int i = 0; i < 9; ++i) r.m[i] = a.m[i] * s;
    return r;
}
// Row-by-column product with the running sum started at zero and the k
// terms added in order (cuda_backend.cu:114-117).
template <typename S> TM3<S> tmmul(const TM3<S> &a, const TM3<S> &b)
{
    TM3<S> r;
    for (int i = 0; i < 9; ++i) r.m[i] = S(0.0);
    for (int y = 0; y < 3; ++y)
        for (int x = 0; x < 3; ++x)
            for (int k = 0; k < 3; ++k) r.m[3 * y + x] = r.m[3 * y + x] + a.m[3 * y + k] * b.m[3 * k + x];
    return r;
}
template <typename S> TM3<S> ttranspose(const TM3<S> &a)
{
    TM3<S> r;
    for (int y = 0; y < 3; ++y)
        for (int x = 0; x < 3; ++x) r.m[3 * y + x] = a.m[3 * x + y];
    return r;
}
template <typename S> TV3<S> tmvec(const TM3<S> &a, const TV3<S> &b)
{
    return {a.m[0] * b.x + a.m[1] * b.y + a.m[2] * b.z,
            a.m[3] * b.x + a.m[4] * b.y + a.m[5] * b.z,
            a.m[6] * b.x + a.m[7] * b.y + a.m[8] * b.z};
}
template <typename S> TM3<S> tskew(const TV3<S> &w)
{
    TM3<S> r;
    for (int i = 0; i < 9; ++i) r.m[i] = S(0.0);
    r.m[1] = -w.z; r.m[2] = w.y;
    r.m[3] = w.z;  r.m[5] = -w.x;
    r.m[6] = -w.y; r.m[7] = w.x;
    return r;
}

// Rodrigues' formula exp([w]x) = I + a [w]x + b [w]x^2 with a = sin t / t and
// b = (1 - cos t) / t^2, t = |w|; below t^2 < 1e-10 both use their Taylor
// series to t^4 (cuda_backend.cu:130-136).
template <int N> TM3<Dual<N>> exp_so3(const TV3<Dual<N>> &w)
{
    typedef Dual<N> D;
    const D t2 = tdot(w, w);
    D a, b;
    if (t2.v < 1.0e-10) {
        a = D(1.0) - t2 / D(6.0) + t2 * t2 / D(120.0);
        b = D(0.5) - t2 / D(24.0) + t2 * t2 / D(720.0);
    } else {
        const D t = dsqrt(t2);
        a = dsin(t) / t;
        b = (D(1.0) - dcos(t)) / t2;
    }
    const TM3<D> W = tskew(w);
    return tmadd(tidentity<D>(), tmadd(tmscale(W, a), tmscale(tmmul(W, W), b)));
}

// Cayley-type rotation vector 2 v / (1 + c + 1e-12) with v the axial vector
// of the skew part and c = (tr Q - 1) / 2 (cuda_backend.cu:138-143,
// math.hpp:14-24).
template <int N> TV3<Dual<N>> cayley(const TM3<Dual<N>> &q)
{
    typedef Dual<N> D;
    TV3<D> v{(q.m[7] - q.m[5]) * D(0.5), (q.m[2] - q.m[6]) * D(0.5), (q.m[3] - q.m[1]) * D(0.5)};
    const D c = (q.m[0] + q.m[4] + q.m[8] - D(1.0)) * D(0.5);
    return tscale(v, D(2.0) / (D(1.0) + c + D(1.0e-12)));
}

// Plain double version used by the rest-state build (math.hpp:14-24).
inline V3 cayley_rotation_vector(const M3 &q)
{
    V3 v{(q(2, 1) - q(1, 2)) * 0.5, (q(0, 2) - q(2, 0)) * 0.5, (q(1, 0) - q(0, 1)) * 0.5};
    const double c = (q(0, 0) + q(1, 1) + q(2, 2) - 1.0) * 0.5;
    return (2.0 / (1.0 + c + 1.0e-12)) * v;
}

// Frame whose third column is the tangent and first column the hint made
// orthogonal to it, falling back to the x (or y) axis when the hint is
// parallel (math.hpp:26-42).
inline M3 frame_from_tangent(V3 tangent, V3 hint)
{
    const V3 d3 = normalized(tangent);
    V3 d1 = hint - d3 * dot(hint, d3);
    if (norm2(d1) < 1.0e-16) {
        const V3 axis = std::fabs(d3.x) < 0.8 ? v3(1, 0, 0) : v3(0, 1, 0);
        d1 = axis - d3 * dot(axis, d3);
    }
    d1 = normalized(d1);
    const V3 d2 = normalized(cross(d3, d1));
    d1 = normalized(cross(d2, d3));
    return from_columns(d1, d2, d3);
}

// Eigen 3.4.0 MatrixBase::unitOrthogonal for 3-vectors (published
// algorithm): rotate in the xy plane unless x and y are both much smaller
// than z (precision 1e-12), else in the yz plane.
inline V3 unit_orthogonal(V3 s)
{
    const double prec = 1.0e-12;
    const bool x_small = std::fabs(s.x) <= std::fabs(s.z) * prec;
    const bool y_small = std::fabs(s.y) <= std::fabs(s.z) * prec;
    if (!x_small || !y_small) {
        const double inv = 1.0 / std::sqrt(s.x * s.x + s.y * s.y);
        return {-s.y * inv, s.x * inv, 0.0};
    }
    const double inv = 1.0 / std::sqrt(s.y * s.y + s.z * s.z);
    return {0.0, -s.z * inv, s.y * inv};
}

// Rotates a director by the minimal rotation taking one tangent to the
// next; antiparallel tangents reflect about a unit orthogonal
// (math.hpp:44-57).
inline V3 parallel_transport(V3 director, V3 from, V3 to)
{
    const V3 a = normalized(from);
    const V3 b = normalized(to);
    const V3 axis = cross(a, b);
    const double s = norm(axis);
    const double c = std::clamp(dot(a, b), -1.0, 1.0);
    if (s < 1.0e-12) {
        if (c > 0.0) return director;
        const V3 n = unit_orthogonal(a);
        return 2.0 * n * dot(n, director) - director;
    }
    const V3 n = div_each(axis, s);
    return director * c + cross(n, director) * s + n * dot(n, director) * (1.0 - c);
}

// Rotation vector (axis * angle) of a rotation matrix, restating Eigen
// 3.4.0's AngleAxisd(Matrix3d): matrix -> quaternion by the trace/largest
// diagonal method, then angle = 2 atan2(|vec|, |w|) with the axis sign taken
// from w; a non-finite result gives zero (solver.cpp:19-26).
inline V3 rotation_vector(const M3 &r)
{
    double q[4];  // x, y, z, w
    double t = r(0, 0) + r(1, 1) + r(2, 2);
    if (t > 0.0) {
        t = std::sqrt(t + 1.0);
        q[3] = 0.5 * t;
        t = 0.5 / t;
        q[0] = (r(2, 1) - r(1, 2)) * t;
        q[1] = (r(0, 2) - r(2, 0)) * t;
        q[2] = (r(1, 0) - r(0, 1)) * t;
    } else {
        int i = 0;
        if (r(1, 1) > r(0, 0)) i = 1;
        if (r(2, 2) > r(i, i)) i = 2;
        const int j = (i + 1) % 3;
        const int k = (j + 1) % 3;
        t = std::sqrt(r(i, i) - r(j, j) - r(k, k) + 1.0);
        q[i] = 0.5 * t;
        t = 0.5 / t;
        q[3] = (r(k, j) - r(j, k)) * t;
        q[j] = (r(j, i) + r(i, j)) * t;
        q[k] = (r(k, i) + r(i, k)) * t;
    }
    double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
    double angle = 0.0;
    V3 axis{1.0, 0.0, 0.0};
    if (n != 0.0) {
        angle = 2.0 * std::atan2(n, std::fabs(q[3]));
        if (q[3] < 0.0) n = -n;
        axis = {q[0] / n, q[1] / n, q[2] / n};
    }
    if (!std::isfinite(angle) || !finite3(axis)) return V3{};
    if (angle > 3.14159265358979323846) angle -= 2.0 * 3.14159265358979323846;
    return axis * angle;
}

}  // namespace oracle
