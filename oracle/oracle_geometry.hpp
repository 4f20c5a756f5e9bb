// Oracle geometry: closest points, axis-aligned boxes and a median-split
// bounding volume hierarchy with an exact per-triangle box filter.
//
// TEST INFRASTRUCTURE ONLY (see oracle_math.hpp).
#pragma once

#include "oracle_math.hpp"

#include <limits>
#include <numeric>
#include <vector>

namespace oracle {

struct Tri {
    uint32_t i0 = 0, i1 = 0, i2 = 0;
};

struct Closest {
    double distance = 1.0e300;
    double t = 0.0;  // parameter on the hair segment
    V3 bary{1.0, 0.0, 0.0};
};

// Ericson, Real-Time Collision Detection 5.1.5: closest point of a triangle
// by Voronoi regions (cuda_backend.cu:200-209, geometry.hpp:15-51).
inline V3 closest_point_triangle(V3 p, V3 a, V3 b, V3 c, V3 &bary)
{
    const V3 ab = b - a, ac = c - a, ap = p - a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0 && d2 <= 0) { bary = v3(1, 0, 0); return a; }
    const V3 bp = p - b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0 && d4 <= d3) { bary = v3(0, 1, 0); return b; }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        bary = v3(1 - v, v, 0);
        return a + ab * v;
    }
    const V3 cp = p - c;
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0 && d5 <= d6) { bary = v3(0, 0, 1); return c; }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        bary = v3(1 - w, 0, w);
        return a + ac * w;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && (d4 - d3) >= 0 && (d5 - d6) >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        bary = v3(0, 1 - w, w);
        return b + (c - b) * w;
    }
    const double inv = 1.0 / (va + vb + vc), v = vb * inv, w = vc * inv;
    bary = v3(1 - v - w, v, w);
    return a + ab * v + ac * w;
}

// Ericson 5.1.9: closest points of two segments with clamping; keeps the
// pair when it is strictly closer than `best` (cuda_backend.cu:210-216,
// geometry.hpp:60-87).
inline void closest_segment_segment(V3 p1, V3 q1, V3 p2, V3 q2, double &best, double &bs, double &bt)
{
    const V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double s = 0, t = 0;
    if (a <= 1e-14 && e <= 1e-14) {
    } else if (a <= 1e-14) {
        t = std::fmin(1.0, std::fmax(0.0, f / e));
    } else {
        const double c = dot(d1, r);
        if (e <= 1e-14) {
            s = std::fmin(1.0, std::fmax(0.0, -c / a));
        } else {
            const double b = dot(d1, d2), den = a * e - b * b;
            if (den != 0) s = std::fmin(1.0, std::fmax(0.0, (b * f - c * e) / den));
            t = (b * s + f) / e;
            if (t < 0) {
                t = 0;
                s = std::fmin(1.0, std::fmax(0.0, -c / a));
            } else if (t > 1) {
                t = 1;
                s = std::fmin(1.0, std::fmax(0.0, (b - c) / a));
            }
        }
    }
    const double d = norm((p1 + d1 * s) - (p2 + d2 * t));
    if (d < best) { best = d; bs = s; bt = t; }
}

// Moller-Trumbore segment/triangle intersection (geometry.hpp:89-110).
inline bool segment_triangle_hit(V3 p0, V3 p1, V3 a, V3 b, V3 c, double &t, V3 &bary)
{
    const V3 d = p1 - p0, e1 = b - a, e2 = c - a, h = cross(d, e2);
    const double det = dot(e1, h);
    if (std::fabs(det) < 1.0e-14) return false;
    const double inv = 1.0 / det;
    const V3 s = p0 - a;
    const double u = inv * dot(s, h);
    if (u < 0.0 || u > 1.0) return false;
    const V3 q = cross(s, e1);
    const double v = inv * dot(d, q);
    if (v < 0.0 || u + v > 1.0) return false;
    t = inv * dot(e2, q);
    if (t < 0.0 || t > 1.0) return false;
    bary = v3(1.0 - u - v, u, v);
    return true;
}

// Segment/triangle distance as the minimum over both endpoints against the
// triangle and the segment against the three edges. The step's version has
// no intersection test (cuda_backend.cu:217-223); the build-time gap check
// adds one and reports distance 0 on a hit (geometry.hpp:112-149).
inline Closest closest_segment_triangle(V3 p0, V3 p1, V3 a, V3 b, V3 c, bool intersection_test)
{
    Closest out;
    if (intersection_test) {
        double t = 0;
        V3 bary;
        if (segment_triangle_hit(p0, p1, a, b, c, t, bary)) {
            out.distance = 0.0;
            out.t = t;
            out.bary = bary;
            return out;
        }
        out.distance = std::numeric_limits<double>::infinity();
    }
    V3 bary;
    V3 cp = closest_point_triangle(p0, a, b, c, bary);
    double d = norm(p0 - cp);
    if (d < out.distance) { out.distance = d; out.t = 0; out.bary = bary; }
    cp = closest_point_triangle(p1, a, b, c, bary);
    d = norm(p1 - cp);
    if (d < out.distance) { out.distance = d; out.t = 1; out.bary = bary; }
    const V3 tv[3] = {a, b, c};
    for (int e = 0; e < 3; ++e) {
        double best = out.distance, s = 0, t = 0;
        closest_segment_segment(p0, p1, tv[e], tv[(e + 1) % 3], best, s, t);
        if (best < out.distance) {
            out.distance = best;
            out.t = s;
            out.bary = V3{};
            if (e == 0) { out.bary.x = 1 - t; out.bary.y = t; }
            else if (e == 1) { out.bary.y = 1 - t; out.bary.z = t; }
            else { out.bary.z = 1 - t; out.bary.x = t; }
        }
    }
    return out;
}

// Box overlap with padding (cuda_backend.cu:225-227, geometry.hpp:151-156).
inline bool overlap(V3 al, V3 au, V3 bl, V3 bu, double p)
{
    return al.x <= bu.x + p && bl.x <= au.x + p && al.y <= bu.y + p && bl.y <= au.y + p &&
           al.z <= bu.z + p && bl.z <= au.z + p;
}

// Median-split hierarchy over triangle centroids with leaves of at most 8
// triangles (solver.cpp:36-88). Queries test each leaf triangle's own box,
// so the reported set is exactly {t : overlap(query box, box(t), padding)}
// whatever the tree shape; it is returned sorted by triangle index.
class Bvh {
public:
    struct Node {
        V3 lo, hi;
        uint32_t begin = 0, count = 0;
        int left = -1, right = -1;
    };

    void build(const std::vector<V3> &verts, const std::vector<Tri> &tris)
    {
        tris_ = &tris;
        order_.resize(tris.size());
        std::iota(order_.begin(), order_.end(), 0u);
        nodes_.clear();
        tri_lo_.assign(tris.size(), V3{});
        tri_hi_.assign(tris.size(), V3{});
        if (!tris.empty()) build_node(0, uint32_t(tris.size()), verts);
        refit(verts, nullptr);
    }

    // Recomputes triangle and node boxes; with `old` the boxes enclose both
    // positions (swept boxes, cuda_backend.cu:776-792).
    void refit(const std::vector<V3> &verts, const std::vector<V3> *old)
    {
        const auto &tris = *tris_;
        for (size_t k = 0; k < tris.size(); ++k) {
            const uint32_t ids[3] = {tris[k].i0, tris[k].i1, tris[k].i2};
            V3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
            for (int c = 0; c < 3; ++c) {
                lo = vmin(lo, verts[ids[c]]);
                hi = vmax(hi, verts[ids[c]]);
                if (old) {
                    lo = vmin(lo, (*old)[ids[c]]);
                    hi = vmax(hi, (*old)[ids[c]]);
                }
            }
            tri_lo_[k] = lo;
            tri_hi_[k] = hi;
        }
        // Children are created after their parent, so a reverse sweep
        // visits every child before its parent.
        for (size_t n = nodes_.size(); n-- > 0;) {
            Node &node = nodes_[n];
            if (node.left < 0) {
                V3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
                for (uint32_t k = node.begin; k < node.begin + node.count; ++k) {
                    lo = vmin(lo, tri_lo_[order_[k]]);
                    hi = vmax(hi, tri_hi_[order_[k]]);
                }
                node.lo = lo;
                node.hi = hi;
            } else {
                node.lo = vmin(nodes_[node.left].lo, nodes_[node.right].lo);
                node.hi = vmax(nodes_[node.left].hi, nodes_[node.right].hi);
            }
        }
    }

    void query(V3 lo, V3 hi, double padding, std::vector<uint32_t> &out) const
    {
        out.clear();
        if (nodes_.empty()) return;
        std::vector<int> stack{0};
        while (!stack.empty()) {
            const Node &n = nodes_[stack.back()];
            stack.pop_back();
            if (!overlap(lo, hi, n.lo, n.hi, padding)) continue;
            if (n.left >= 0) {
                stack.push_back(n.left);
                stack.push_back(n.right);
            } else {
                for (uint32_t k = n.begin; k < n.begin + n.count; ++k) {
                    const uint32_t t = order_[k];
                    if (overlap(lo, hi, tri_lo_[t], tri_hi_[t], padding)) out.push_back(t);
                }
            }
        }
        std::sort(out.begin(), out.end());
    }

    // Brute force over all triangle boxes (used by tests of the tree).
    void query_brute(V3 lo, V3 hi, double padding, std::vector<uint32_t> &out) const
    {
        out.clear();
        for (uint32_t t = 0; t < tri_lo_.size(); ++t)
            if (overlap(lo, hi, tri_lo_[t], tri_hi_[t], padding)) out.push_back(t);
    }

    bool empty() const { return nodes_.empty(); }

private:
    int build_node(uint32_t begin, uint32_t end, const std::vector<V3> &verts)
    {
        const auto &tris = *tris_;
        const int idx = int(nodes_.size());
        nodes_.push_back({});
        V3 cl{1e300, 1e300, 1e300}, cu{-1e300, -1e300, -1e300};
        for (uint32_t k = begin; k < end; ++k) {
            const Tri &t = tris[order_[k]];
            const V3 c = div_each(verts[t.i0] + verts[t.i1] + verts[t.i2], 3.0);
            cl = vmin(cl, c);
            cu = vmax(cu, c);
        }
        nodes_[idx].begin = begin;
        nodes_[idx].count = end - begin;
        if (end - begin <= 8) return idx;
        const V3 span = cu - cl;
        int axis = 0;
        if (span.y > comp(span, axis)) axis = 1;
        if (span.z > comp(span, axis)) axis = 2;
        const uint32_t mid = begin + (end - begin) / 2;
        std::nth_element(order_.begin() + begin, order_.begin() + mid, order_.begin() + end,
                         [&](uint32_t l, uint32_t r) {
                             const Tri &tl = tris[l], &tr = tris[r];
                             const double lc = (comp(verts[tl.i0], axis) + comp(verts[tl.i1], axis) + comp(verts[tl.i2], axis)) / 3.0;
                             const double rc = (comp(verts[tr.i0], axis) + comp(verts[tr.i1], axis) + comp(verts[tr.i2], axis)) / 3.0;
                             return lc < rc;
                         });
        const int left = build_node(begin, mid, verts);
        const int right = build_node(mid, end, verts);
        nodes_[idx].left = left;
        nodes_[idx].right = right;
        nodes_[idx].count = 0;
        return idx;
    }

    const std::vector<Tri> *tris_ = nullptr;
    std::vector<uint32_t> order_;
    std::vector<Node> nodes_;
    std::vector<V3> tri_lo_, tri_hi_;
};

}  // namespace oracle
