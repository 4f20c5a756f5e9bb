// Linear BVH over collider triangles, rebuilt on the device (Morton codes,
// radix sort, Karras 2012 hierarchy, bottom-up boxes), and its query.
//
// Each internal node stores the boxes of its two children, so a traversal
// step loads one node. A child that is a leaf is one triangle, and the box
// stored for it is that triangle's own box: a query therefore reports
// exactly the triangles whose box overlaps the padded query box, whatever
// the tree shape.
#pragma once

#include "device_math.hpp"

namespace khs {

struct alignas(16) BvhNode {
    double box[12];  // child 0 lo, hi; child 1 lo, hi
    int child[2];    // >= 0 internal node; < 0 leaf, triangle = -1 - child
    int range[2];    // first and last sorted leaf below the node
};

// Per-triangle record at the collider's current position: the triangle's
// box (lo, hi), its three vertices and one unused word, 128 bytes, one cache
// line. A query that has just tested a leaf's box reads the triangle's
// vertices from the line it already fetched, not through the index and
// vertex arrays (two more dependent loads).
constexpr int kTriRecord = 16;
constexpr int kRecordVertices = 6;  // offset of vertex 0

namespace dev {

constexpr int kBvhStack = 64;

// The Karras hierarchy over 30-bit codes with index tie-break has depth
// below 64, so the traversal stack never overflows.

// Calls f(triangle) for every triangle whose box overlaps [lo, hi] grown by
// pad, until f returns false. With leaf_box (the triangle records above,
// kTriRecord doubles each, box first) a leaf is tested against its record's
// box instead of the box stored in the tree, which must contain it: a tree
// refit with boxes swept over a collider motion then answers queries on the
// collider's current position exactly.
template <class F>
__device__ inline void bvh_visit(const BvhNode *nodes, int internal_count, const double *leaf_box, V3 lo, V3 hi,
                                 double pad, F &&f)
{
    int stack[kBvhStack];
    int top = 0, cur = 0;
    // A tree over T triangles has T - 1 internal nodes, each visited at most
    // once; the cap only bounds the loop should the tree be malformed.
    for (int visits = 0; visits < (1 << 24); ++visits) {
        const BvhNode &n = nodes[cur];
        int next = -1;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const V3 bl = mk(n.box[6 * c], n.box[6 * c + 1], n.box[6 * c + 2]);
            const V3 bu = mk(n.box[6 * c + 3], n.box[6 * c + 4], n.box[6 * c + 5]);
            if (!overlap(lo, hi, bl, bu, pad)) continue;
            const int ch = n.child[c];
            if (ch >= internal_count) continue;  // never taken by a well-formed tree
            if (ch < 0) {
                const uint32_t t = uint32_t(-1 - ch);
                if (leaf_box) {
                    const double *b = leaf_box + kTriRecord * size_t(t);
                    if (!overlap(lo, hi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), pad)) continue;
                }
                if (!f(t)) return;
            } else if (next < 0) {
                next = ch;
            } else if (top < kBvhStack) {
                stack[top++] = ch;
            }
        }
        if (next < 0) {
            if (top == 0) break;
            next = stack[--top];
        }
        cur = next;
    }
}

// Calls f(triangle) for every overlapping triangle (see bvh_visit).
template <class F>
__device__ inline void bvh_query(const BvhNode *nodes, int internal_count, const double *leaf_box, V3 lo, V3 hi,
                                 double pad, F &&f)
{
    bvh_visit(nodes, internal_count, leaf_box, lo, hi, pad, [&](uint32_t t) {
        f(t);
        return true;
    });
}

}  // namespace dev
}  // namespace khs
