// Root targets of an uploaded frame of original curve points, computed on
// the device (reference solver.cpp:416-447, rotation_vector :19-26): node
// positions by the binding lerp of the points, and for the fixed nodes the
// rotation vector of the frame transported along the new tangents against
// the reference frame. The math is the model build's (host_math.hpp,
// compiled for both sides), so both give the same operations in the same
// order.
#pragma once

#include "host_math.hpp"
#include "kernels.hpp"

namespace khs {
namespace dev {

// One lane per node.
__global__ void k_root_positions(uint32_t node_count, const uint32_t *bind_left, const uint32_t *bind_right,
                                 const double *bind_t, const double *points, double *pos)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= node_count) return;
    const double *a = points + 3 * size_t(bind_left[n]), *b = points + 3 * size_t(bind_right[n]);
    const double t = bind_t[n];
    const Vec3 p = (1.0 - t) * Vec3{a[0], a[1], a[2]} + t * Vec3{b[0], b[1], b[2]};
    pos[3 * size_t(n)] = p.x;
    pos[3 * size_t(n) + 1] = p.y;
    pos[3 * size_t(n) + 2] = p.z;
}

__device__ inline Mat3 load_frame(const double *frame, uint32_t n)
{
    Mat3 r;
#pragma unroll
    for (int e = 0; e < 9; ++e) r.m[e] = frame[9 * size_t(n) + e];
    return r;
}

// One lane per strand: every tangent of the strand is checked (a degenerate
// one fails the frame, as on the host); the frame is transported from the
// root up to the last fixed node, and the fixed nodes get their rotation
// vectors. The step reads targets of fixed nodes only; the others get zero.
__global__ void k_root_rotations(RodView m, const double *pos, double *rot, unsigned *degenerate)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m.strand_count) return;
    const uint32_t f = m.strand_first[s], c = m.strand_count_of[s], fixed = m.strand_fixed[s];
    auto at = [&](uint32_t k) {
        const double *p = pos + 3 * size_t(f + k);
        return Vec3{p[0], p[1], p[2]};
    };
    Mat3 frame;
    Vec3 previous;
    bool ok = true;
    for (uint32_t k = 0; k < c; ++k) {
        const uint32_t lo = k == 0 ? 0 : k - 1, hi = k + 1 == c ? k : k + 1;
        const Vec3 t = at(hi) - at(lo);
        if (length2(t) < 1.0e-24) {
            ok = false;
            break;
        }
        const Vec3 tangent = unit(t);
        Vec3 r{};
        if (k < fixed) {
            const Mat3 ref = load_frame(m.frame, f + k);
            frame = k == 0 ? tangent_frame(tangent, ref.column(0))
                           : tangent_frame(tangent, transport(frame.column(0), previous, tangent));
            r = axis_angle(product(frame, transposed(ref)));
        }
        double *o = rot + 3 * size_t(f + k);
        o[0] = r.x;
        o[1] = r.y;
        o[2] = r.z;
        previous = tangent;
    }
    if (!ok) atomicOr(degenerate, 1u);
}

}  // namespace dev
}  // namespace khs
