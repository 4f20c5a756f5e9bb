// Device FP64 helpers for the hair step on gfx950: vectors, 3x3 matrices,
// closest-point geometry and the analytic Cosserat strain Jacobians.
//
// All routines are written for one lane; callers tile elements and nodes
// over a wavefront. Operation order in value computations follows the
// reference's device formulas so that, with contraction disabled, values
// round the same way the reference's do.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#else
// Host build of the same routines, used by the CPU tests of the device math.
#include <cmath>
#include <cstring>
#define __device__
#define __host__
#define __forceinline__ inline
using std::cos;
using std::fabs;
using std::fmax;
using std::fmin;
using std::sin;
using std::sqrt;
inline double __longlong_as_double(long long v)
{
    double d;
    std::memcpy(&d, &v, sizeof d);
    return d;
}
#endif

namespace khs {
namespace dev {

struct V3 {
    double x, y, z;
};

__device__ __forceinline__ V3 mk(double x, double y, double z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double norm(V3 a) { return sqrt(dot(a, a)); }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z)}; }
__device__ __forceinline__ V3 load3(const double *p) { return {p[0], p[1], p[2]}; }
__device__ __forceinline__ void store3(double *p, V3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }

// Padded box overlap (reference cuda_backend.cu:225-227).
__device__ __forceinline__ bool overlap(V3 al, V3 au, V3 bl, V3 bu, double p)
{
    return al.x <= bu.x + p && bl.x <= au.x + p && al.y <= bu.y + p && bl.y <= au.y + p && al.z <= bu.z + p &&
           bl.z <= au.z + p;
}

struct Closest {
    double distance, t;
    V3 bary;
};

// Closest point of triangle abc to p by Voronoi regions (Ericson 5.1.5).
__device__ inline V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c, V3 &bary)
{
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { bary = mk(1, 0, 0); return a; }
    const V3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { bary = mk(0, 1, 0); return b; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        bary = mk(1 - v, v, 0);
        return add(a, mul(ab, v));
    }
    const V3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { bary = mk(0, 0, 1); return c; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        bary = mk(1 - w, 0, w);
        return add(a, mul(ac, w));
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary = mk(0, 1 - w, w);
        return add(b, mul(sub(c, b), w));
    }
    const double inv = 1.0 / (va + vb + vc), v = vb * inv, w = vc * inv;
    bary = mk(1 - v - w, v, w);
    return add(add(a, mul(ab, v)), mul(ac, w));
}

// Clamped closest points of two segments (Ericson 5.1.9); keeps a strictly
// closer pair.
__device__ inline void closest_segments(V3 p1, V3 q1, V3 p2, V3 q2, double &best, double &bs, double &bt)
{
    const V3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
    const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double s = 0, t = 0;
    if (a <= 1e-14 && e <= 1e-14) {
    } else if (a <= 1e-14) {
        t = fmin(1.0, fmax(0.0, f / e));
    } else {
        const double c = dot(d1, r);
        if (e <= 1e-14) {
            s = fmin(1.0, fmax(0.0, -c / a));
        } else {
            const double b = dot(d1, d2), den = a * e - b * b;
            if (den != 0) s = fmin(1.0, fmax(0.0, (b * f - c * e) / den));
            t = (b * s + f) / e;
            if (t < 0) {
                t = 0;
                s = fmin(1.0, fmax(0.0, -c / a));
            } else if (t > 1) {
                t = 1;
                s = fmin(1.0, fmax(0.0, (b - c) / a));
            }
        }
    }
    const double d = norm(sub(add(p1, mul(d1, s)), add(p2, mul(d2, t))));
    if (d < best) { best = d; bs = s; bt = t; }
}

// Moller-Trumbore segment/triangle hit (reference geometry.hpp:89-110).
__device__ inline bool segment_hits_triangle(V3 p0, V3 p1, V3 a, V3 b, V3 c, double &t, V3 &bary)
{
    const V3 d = sub(p1, p0), e1 = sub(b, a), e2 = sub(c, a);
    const V3 h = mk(d.y * e2.z - d.z * e2.y, d.z * e2.x - d.x * e2.z, d.x * e2.y - d.y * e2.x);
    const double det = dot(e1, h);
    if (fabs(det) < 1.0e-14) return false;
    const double inv = 1.0 / det;
    const V3 s = sub(p0, a);
    const double u = inv * dot(s, h);
    if (u < 0.0 || u > 1.0) return false;
    const V3 q = mk(s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x);
    const double v = inv * dot(d, q);
    if (v < 0.0 || u + v > 1.0) return false;
    t = inv * dot(e2, q);
    if (t < 0.0 || t > 1.0) return false;
    bary = mk(1.0 - u - v, u, v);
    return true;
}

// Segment/triangle distance: both endpoints against the face, the segment
// against the three edges (reference cuda_backend.cu:217-223); with
// `hit_test` an intersection returns distance 0 first (host variant,
// reference geometry.hpp:112-149).
__device__ inline Closest closest_segment_triangle(V3 p0, V3 p1, V3 a, V3 b, V3 c, bool hit_test = false)
{
    Closest out{1.0e300, 0.0, mk(1, 0, 0)};
    if (hit_test) {
        double t;
        V3 bary;
        if (segment_hits_triangle(p0, p1, a, b, c, t, bary)) return Closest{0.0, t, bary};
        out.distance = __longlong_as_double(0x7ff0000000000000LL);
    }
    V3 bary;
    V3 cp = closest_point_triangle(p0, a, b, c, bary);
    double d = norm(sub(p0, cp));
    if (d < out.distance) out = Closest{d, 0.0, bary};
    cp = closest_point_triangle(p1, a, b, c, bary);
    d = norm(sub(p1, cp));
    if (d < out.distance) out = Closest{d, 1.0, bary};
    const V3 tv[3] = {a, b, c};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        double best = out.distance, s = 0, t = 0;
        closest_segments(p0, p1, tv[e], tv[(e + 1) % 3], best, s, t);
        if (best < out.distance) {
            out.distance = best;
            out.t = s;
            if (e == 0) out.bary = mk(1 - t, t, 0);
            else if (e == 1) out.bary = mk(0, 1 - t, t);
            else out.bary = mk(t, 0, 1 - t);
        }
    }
    return out;
}

// ------------------------------------------------------------------ SO(3)
// Row-major 3x3 products with the k terms summed from zero in order.
__device__ __forceinline__ void mat_mul(const double *a, const double *b, double *r)
{
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s = s + a[3 * y + k] * b[3 * k + x];
            r[3 * y + x] = s;
        }
}

// r = a^T b
__device__ __forceinline__ void mat_tmul(const double *a, const double *b, double *r)
{
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s = s + a[3 * k + y] * b[3 * k + x];
            r[3 * y + x] = s;
        }
}

// exp([w]x) = I + a W + b W^2, a = sin t / t, b = (1 - cos t) / t^2, with
// the t^4 Taylor series below t^2 < 1e-10 (reference cuda_backend.cu:130-136).
__device__ inline void exp_so3(V3 w, double *E)
{
    const double t2 = dot(w, w);
    double a, b;
    // Quotients are taken as a * (1 / b), the rounding of the reference's
    // dual-number division.
    if (t2 < 1.0e-10) {
        a = 1.0 - t2 * (1.0 / 6.0) + (t2 * t2) * (1.0 / 120.0);
        b = 0.5 - t2 * (1.0 / 24.0) + (t2 * t2) * (1.0 / 720.0);
    } else {
        const double t = sqrt(t2);
        a = sin(t) * (1.0 / t);
        b = (1.0 - cos(t)) * (1.0 / t2);
    }
    const double W[9] = {0.0, -w.z, w.y, w.z, 0.0, -w.x, -w.y, w.x, 0.0};
    double WW[9];
    mat_mul(W, W, WW);
#pragma unroll
    for (int k = 0; k < 9; ++k) E[k] = ((k % 4 == 0) ? 1.0 : 0.0) + (W[k] * a + WW[k] * b);
}

// Right Jacobian of the exponential, exp(u + du) = exp(u) exp(Jr(u) du):
// Jr = I - alpha [u]x + beta [u]x^2 with alpha = (1 - cos t) / t^2 and
// beta = (t - sin t) / t^3 (series below t^2 < 1e-2).
__device__ inline void right_jacobian(V3 u, double *J)
{
    const double t2 = dot(u, u);
    double al, be;
    if (t2 < 1.0e-2) {
        al = 0.5 - t2 * (1.0 / 24.0 - t2 * (1.0 / 720.0 - t2 * (1.0 / 40320.0 - t2 / 3628800.0)));
        be = 1.0 / 6.0 - t2 * (1.0 / 120.0 - t2 * (1.0 / 5040.0 - t2 * (1.0 / 362880.0 - t2 / 39916800.0)));
    } else {
        const double t = sqrt(t2);
        al = (1.0 - cos(t)) / t2;
        be = (t - sin(t)) / (t2 * t);
    }
    const double W[9] = {0.0, -u.z, u.y, u.z, 0.0, -u.x, -u.y, u.x, 0.0};
    double WW[9];
    mat_mul(W, W, WW);
#pragma unroll
    for (int k = 0; k < 9; ++k) J[k] = ((k % 4 == 0) ? 1.0 : 0.0) - al * W[k] + be * WW[k];
}

// Strains of one element and their partial derivatives.
//   gamma = 1/2 (Ri^T + Rj^T) (xj - xi) / L0 - gamma0
//   kappa = cayley(Ri^T Rj) / L0 - kappa0,  Rk = exp(uk) Rbar_k
// (reference cuda_backend.cu:234-245). Derivatives, with
// Rk(u + du) = Rk exp([Mk du]x), Mk = Rbar_k^T Jr(uk):
//   dgamma/dxj = -dgamma/dxi = (Ri^T + Rj^T) / (2 L0)
//   dgamma/duk = 1/2 [Rk^T t]x Mk
//   dkappa/duj =  (1/L0) ((trQ I - Q^T) / s + 2 v v^T / s^2) Mj
//   dkappa/dui = -(1/L0) ((trQ I - Q) / s + 2 v v^T / s^2) Mi
// with Q = Ri^T Rj, v its axial vector, s = 1 + (trQ - 1)/2 + 1e-12.
struct ElementJacobian {
    double strain[6];
    double Sx[9];  // dgamma/dxj (row = strain, column = coordinate)
    double Si[9], Sj[9], Ki[9], Kj[9];
};

// A node's rotation Rk = exp(uk) Rbar_k and, with the Jacobian,
// Mk = Rbar_k^T Jr(uk): the per-node half of an element's strains, which the
// two elements sharing the node use alike.
struct NodeRotation {
    double R[9];
    double M[9];
};

template <bool WITH_JACOBIAN>
__device__ inline void node_rotation(const double *qk, const double *Fk, NodeRotation &out)
{
    const V3 u = mk(qk[3], qk[4], qk[5]);
    double E[9];
    exp_so3(u, E);
    mat_mul(E, Fk, out.R);
    if (!WITH_JACOBIAN) return;
    double Jr[9];
    right_jacobian(u, Jr);
    mat_tmul(Fk, Jr, out.M);
}

// The element's strains (and Jacobian) from its nodes' positions and
// rotations.
template <bool WITH_JACOBIAN>
__device__ inline void element_strains_from(const double *qi, const double *qj, const NodeRotation &ni,
                                            const NodeRotation &nj, const double *rest, ElementJacobian &out)
{
    const double *Ri = ni.R, *Rj = nj.R;
    const double inv_L = 1.0 / rest[0];
    const V3 t = mk((qj[0] - qi[0]) * inv_L, (qj[1] - qi[1]) * inv_L, (qj[2] - qi[2]) * inv_L);
    double A2[9];  // Ri^T + Rj^T
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) A2[3 * y + x] = Ri[3 * x + y] + Rj[3 * x + y];
    out.strain[0] = (A2[0] * t.x + A2[1] * t.y + A2[2] * t.z) * 0.5 - rest[1];
    out.strain[1] = (A2[3] * t.x + A2[4] * t.y + A2[5] * t.z) * 0.5 - rest[2];
    out.strain[2] = (A2[6] * t.x + A2[7] * t.y + A2[8] * t.z) * 0.5 - rest[3];
    double Q[9];
    mat_tmul(Ri, Rj, Q);
    const V3 v = mk((Q[7] - Q[5]) * 0.5, (Q[2] - Q[6]) * 0.5, (Q[3] - Q[1]) * 0.5);
    const double c = (Q[0] + Q[4] + Q[8] - 1.0) * 0.5;
    const double s = 1.0 + c + 1.0e-12;
    const double f = 2.0 * (1.0 / s);
    out.strain[3] = (v.x * f) * inv_L - rest[4];
    out.strain[4] = (v.y * f) * inv_L - rest[5];
    out.strain[5] = (v.z * f) * inv_L - rest[6];
    if (!WITH_JACOBIAN) return;

#pragma unroll
    for (int k = 0; k < 9; ++k) out.Sx[k] = (A2[k] * inv_L) * 0.5;

    const double *Mi = ni.M, *Mj = nj.M;
    // a_k = Rk^T t; dgamma/duk = 1/2 [a_k]x Mk
    const V3 ai = mk(Ri[0] * t.x + Ri[3] * t.y + Ri[6] * t.z, Ri[1] * t.x + Ri[4] * t.y + Ri[7] * t.z,
                     Ri[2] * t.x + Ri[5] * t.y + Ri[8] * t.z);
    const V3 aj = mk(Rj[0] * t.x + Rj[3] * t.y + Rj[6] * t.z, Rj[1] * t.x + Rj[4] * t.y + Rj[7] * t.z,
                     Rj[2] * t.x + Rj[5] * t.y + Rj[8] * t.z);
    const double Ai[9] = {0.0, -0.5 * ai.z, 0.5 * ai.y, 0.5 * ai.z, 0.0, -0.5 * ai.x, -0.5 * ai.y, 0.5 * ai.x, 0.0};
    const double Aj[9] = {0.0, -0.5 * aj.z, 0.5 * aj.y, 0.5 * aj.z, 0.0, -0.5 * aj.x, -0.5 * aj.y, 0.5 * aj.x, 0.0};
    mat_mul(Ai, Mi, out.Si);
    mat_mul(Aj, Mj, out.Sj);

    const double tr = Q[0] + Q[4] + Q[8];
    const double is = inv_L / s, vv = 2.0 * inv_L / (s * s);
    const double vc[3] = {v.x, v.y, v.z};
    double Bj[9], Bi[9];
#pragma unroll
    for (int y = 0; y < 3; ++y)
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double id = (x == y) ? tr : 0.0;
            const double outer = vv * vc[y] * vc[x];
            Bj[3 * y + x] = (id - Q[3 * x + y]) * is + outer;
            Bi[3 * y + x] = -((id - Q[3 * y + x]) * is + outer);
        }
    mat_mul(Bi, Mi, out.Ki);
    mat_mul(Bj, Mj, out.Kj);
}

template <bool WITH_JACOBIAN>
__device__ inline void element_strains(const double *qi, const double *qj, const double *Fi, const double *Fj,
                                       const double *rest, ElementJacobian &out)
{
    NodeRotation ni, nj;
    node_rotation<WITH_JACOBIAN>(qi, Fi, ni);
    node_rotation<WITH_JACOBIAN>(qj, Fj, nj);
    element_strains_from<WITH_JACOBIAN>(qi, qj, ni, nj, rest, out);
}

}  // namespace dev
}  // namespace khs
