// The rod model built on the device from the uploaded curve points: strand
// discretization with zero-length merges and the virtual extension
// (reference src/solver.cpp:227-318), parallel-transported frames and rest
// strains (:320-356), lumped masses (:358-373). One lane per strand walks
// its points in order, so every node, frame and rest value is formed by the
// same operations in the same order as the reference's sequential build
// (helpers of host_math.hpp, compiled without contraction). The strand
// layout (node counts and their scan) is the only part that crosses to the
// host: S counts down, S first-node indices up.
#pragma once

#include "host_math.hpp"
#include "kernels.hpp"
#include "root_kernels.hpp"

namespace khs {
namespace dev {

// Per strand output of the counting pass.
struct StrandShape {
    uint32_t nodes;            // internal nodes, extension included
    uint32_t real_nodes;       // of which not virtual
    uint32_t merged;           // zero-length segments merged
    uint32_t extension_nodes;  // virtual extension nodes
    double extension_length;   // rest length added by the extension
    uint32_t degenerate;       // 1: all points at one position
    uint32_t pad;
};

// Walks one strand's segments; `emit(previous, target, p, pieces)` is called
// for each segment longer than the merge threshold, `merge(p)` for each
// merged point. Returns the strand length and the tip segment.
struct StrandWalk {
    double length = 0.0, tip_length = 0.0;
    uint32_t segments = 0, tip_left = 0, tip_right = 0;
    Vec3 tip_direction;
};

template <class Emit, class Merge>
__device__ inline StrandWalk walk_strand(const double *points, uint32_t begin, uint32_t end, double hmax, Emit emit,
                                         Merge merge)
{
    constexpr double zero = 1.0e-12;
    auto point = [&](uint32_t p) {
        const double *x = points + 3 * size_t(p);
        return Vec3{x[0], x[1], x[2]};
    };
    StrandWalk w;
    w.tip_left = w.tip_right = begin;
    Vec3 previous = point(begin);
    for (uint32_t p = begin + 1; p < end; ++p) {
        const Vec3 target = point(p);
        const double len = length(target - previous);
        if (len <= zero) {
            merge(p);
            previous = target;
            continue;
        }
        const uint32_t pieces = max(1u, uint32_t(ceil(len / hmax)));
        emit(previous, target, p, pieces);
        previous = target;
        w.length += len;
        w.tip_left = p - 1;
        w.tip_right = p;
        w.tip_length = len;
        w.tip_direction = over(target - point(p - 1), len);
        ++w.segments;
    }
    return w;
}

__device__ inline uint32_t extension_pieces(double minimum_length, double strand_length, double hmax, double *extra)
{
    constexpr double zero = 1.0e-12;
    if (!(minimum_length > strand_length + zero)) return 0;
    *extra = minimum_length - strand_length;
    return max(1u, uint32_t(ceil(*extra / hmax)));
}

// Pass 1, lane per strand: node counts and build statistics.
__global__ void k_model_count(uint32_t strands, const double *points, const uint32_t *offsets, double hmax,
                              double minimum_length, StrandShape *shape)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= strands) return;
    StrandShape out{};
    out.nodes = 1;
    const StrandWalk w = walk_strand(
        points, offsets[s], offsets[s + 1], hmax, [&](Vec3, Vec3, uint32_t, uint32_t pieces) { out.nodes += pieces; },
        [&](uint32_t) { ++out.merged; });
    out.real_nodes = out.nodes;
    out.degenerate = (w.segments == 0 || out.nodes < 2) ? 1u : 0u;
    if (!out.degenerate) {
        double extra = 0.0;
        out.extension_nodes = extension_pieces(minimum_length, w.length, hmax, &extra);
        if (out.extension_nodes) out.extension_length = extra;
        out.nodes += out.extension_nodes;
    }
    shape[s] = out;
}

// Outputs of pass 2, all indexed by internal node.
struct ModelOut {
    double *q;          // 6 per node: position, zero rotation
    double *frame;      // 9 per node, row-major
    double *rest;       // 8 per node: L0, shear0, curvature0, 0
    double *mass, *rmass;
    uint32_t *flags, *strand_of, *bind_left, *bind_right;
    double *bind_t;
    uint32_t *mapping;  // original point -> internal node
    double *strand_max_rest;
};

// Pass 2, lane per strand: nodes, bindings, frames, rest strains and masses.
// `first` is the exclusive scan of the node counts of pass 1.
__global__ void k_model_nodes(uint32_t strands, const double *points, const uint32_t *offsets, double hmax,
                              double minimum_length, uint32_t fixed_root_nodes, const uint32_t *first_of,
                              double mass_per_length, double rotational_per_length, ModelOut o)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= strands) return;
    const uint32_t begin = offsets[s], end = offsets[s + 1], first = first_of[s];
    uint32_t n = first;
    auto push = [&](Vec3 x, uint32_t left, uint32_t right, double t, uint32_t flag) {
        double *q = o.q + 6 * size_t(n);
        q[0] = x.x;
        q[1] = x.y;
        q[2] = x.z;
        q[3] = q[4] = q[5] = 0.0;
        o.strand_of[n] = s;
        o.bind_left[n] = left;
        o.bind_right[n] = right;
        o.bind_t[n] = t;
        o.flags[n] = flag;
        ++n;
    };
    const double *p0 = points + 3 * size_t(begin);
    push(Vec3{p0[0], p0[1], p0[2]}, begin, begin, 0.0, 0u);
    o.mapping[begin] = first;
    const StrandWalk w = walk_strand(
        points, begin, end, hmax,
        [&](Vec3 previous, Vec3 target, uint32_t p, uint32_t pieces) {
            for (uint32_t d = 1; d <= pieces; ++d) {
                const double t = double(d) / pieces;
                push((1.0 - t) * previous + t * target, p - 1, p, t, 0u);
            }
            o.mapping[p] = n - 1;
        },
        [&](uint32_t p) { o.mapping[p] = n - 1; });
    double extra = 0.0;
    const uint32_t pieces = extension_pieces(minimum_length, w.length, hmax, &extra);
    if (pieces) {
        const double *tip = points + 3 * size_t(end - 1);
        for (uint32_t d = 1; d <= pieces; ++d) {
            const double dist = extra * double(d) / pieces;
            push(Vec3{tip[0], tip[1], tip[2]} + dist * w.tip_direction, w.tip_left, w.tip_right,
                 1.0 + dist / w.tip_length, kNodeVirtual);
        }
    }
    const uint32_t count = n - first, fixed = min(fixed_root_nodes, count);
    for (uint32_t k = 0; k < fixed; ++k) o.flags[first + k] |= kNodeFixed;

    // frames by parallel transport along central-difference tangents
    auto at = [&](uint32_t k) {
        const double *q = o.q + 6 * size_t(first + k);
        return Vec3{q[0], q[1], q[2]};
    };
    auto tangent = [&](uint32_t k) {
        const uint32_t lo = k == 0 ? 0 : k - 1, hi = k + 1 == count ? k : k + 1;
        return unit(at(hi) - at(lo));
    };
    auto store_frame = [&](uint32_t k, const Mat3 &r) {
#pragma unroll
        for (int e = 0; e < 9; ++e) o.frame[9 * size_t(first + k) + e] = r.m[e];
    };
    Vec3 t_prev = tangent(0);
    Mat3 frame = tangent_frame(t_prev, Vec3{1, 0, 0});
    store_frame(0, frame);
    for (uint32_t k = 1; k < count; ++k) {
        const Vec3 t = tangent(k);
        frame = tangent_frame(t, transport(frame.column(0), t_prev, t));
        store_frame(k, frame);
        t_prev = t;
    }

    // rest strains and lumped masses, element by element from the root
    double max_rest = 0.0, carried_mass = 0.0, carried_rot = 0.0;
    Mat3 ri = load_frame(o.frame, first);
    for (uint32_t k = 0; k + 1 < count; ++k) {
        const uint32_t i = first + k, j = i + 1;
        const Mat3 rj = load_frame(o.frame, j);
        uint32_t flag = o.flags[i] | kNodeHasElement;
        if (!(o.flags[j] & kNodeVirtual)) flag |= kNodeContact;
        o.flags[i] = flag;
        const Vec3 xi = at(k), xj = at(k + 1);
        const double L0 = length(xj - xi);
        const Vec3 t = over(xj - xi, L0);
        const Mat3 rit = transposed(ri), rjt = transposed(rj);
        Mat3 sum;
#pragma unroll
        for (int e = 0; e < 9; ++e) sum.m[e] = rit.m[e] + rjt.m[e];
        const Vec3 shear = 0.5 * apply(sum, t);
        const Vec3 curvature = over(cayley_vector(product(rit, rj)), L0);
        double *r = o.rest + 8 * size_t(i);
        r[0] = L0;
        r[1] = shear.x;
        r[2] = shear.y;
        r[3] = shear.z;
        r[4] = curvature.x;
        r[5] = curvature.y;
        r[6] = curvature.z;
        r[7] = 0.0;
        max_rest = max(max_rest, L0);
        const double m = mass_per_length * L0, mr = rotational_per_length * L0;
        o.mass[i] = (0.0 + carried_mass) + 0.5 * m;
        o.rmass[i] = (0.0 + carried_rot) + 0.5 * mr;
        carried_mass = 0.5 * m;
        carried_rot = 0.5 * mr;
        ri = rj;
    }
    o.mass[first + count - 1] = 0.0 + carried_mass;
    o.rmass[first + count - 1] = 0.0 + carried_rot;
    double *r = o.rest + 8 * size_t(first + count - 1);
#pragma unroll
    for (int e = 0; e < 8; ++e) r[e] = 0.0;
    o.strand_max_rest[s] = max_rest;
}

// Lumped masses again after a material change, lane per strand, from the
// rest lengths already on the device.
__global__ void k_model_masses(RodView m, const double *rest, double mass_per_length, double rotational_per_length,
                               const double *strand_density, double *mass, double *rmass)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m.strand_count) return;
    // per-strand materials: the strand's own mass and rotational mass per length
    if (strand_density) {
        mass_per_length = strand_density[2 * size_t(s)];
        rotational_per_length = strand_density[2 * size_t(s) + 1];
    }
    const uint32_t first = m.strand_first[s], count = m.strand_count_of[s];
    double carried_mass = 0.0, carried_rot = 0.0;
    for (uint32_t k = 0; k + 1 < count; ++k) {
        const uint32_t i = first + k;
        const double L0 = rest[8 * size_t(i)];
        const double a = mass_per_length * L0, b = rotational_per_length * L0;
        mass[i] = (0.0 + carried_mass) + 0.5 * a;
        rmass[i] = (0.0 + carried_rot) + 0.5 * b;
        carried_mass = 0.5 * a;
        carried_rot = 0.5 * b;
    }
    mass[first + count - 1] = 0.0 + carried_mass;
    rmass[first + count - 1] = 0.0 + carried_rot;
}

}  // namespace dev
}  // namespace khs
