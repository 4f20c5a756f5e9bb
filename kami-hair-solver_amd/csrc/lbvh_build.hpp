// Device construction of the collider LBVH: triangle boxes, Morton codes,
// Karras hierarchy over the sorted codes, bottom-up boxes, and a structural
// self-check. Compiles for the host too, when the includer first provides
// host stand-ins for the HIP thread indices and atomics (the CPU test of the
// construction runs each kernel thread by thread).
#pragma once

#include "kernels.hpp"
#include "lbvh.hpp"

namespace khs {
namespace dev {

__device__ __forceinline__ V3 vertex(const ColliderView &c, uint32_t i) { return load3(c.current + 3 * size_t(i)); }
__device__ __forceinline__ V3 vertex_prev(const ColliderView &c, uint32_t i)
{
    return load3(c.previous + 3 * size_t(i));
}
// Vertex k (0, 1, 2) of triangle t from its record (kTriRecord doubles per
// triangle, the collider's current position).
__device__ __forceinline__ V3 record_vertex(const double *records, uint32_t t, int k)
{
    return load3(records + kTriRecord * size_t(t) + kRecordVertices + 3 * k);
}
// Vertex k of triangle t at the current position: from the triangle's
// record when the view carries records, else through the index array (the
// same values).
__device__ __forceinline__ V3 triangle_vertex(const ColliderView &c, uint32_t t, int k)
{
    return c.leaf_box ? record_vertex(c.leaf_box, t, k) : vertex(c, c.triangles[3 * size_t(t) + k]);
}

// ------------------------------------------------------------------ LBVH
__device__ __forceinline__ uint32_t spread_bits(uint32_t v)
{
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}

// Triangle boxes (static, or swept over previous and current positions),
// the triangle records at the current position (when `records` is given)
// and per-block centroid bounds.
__global__ __launch_bounds__(256) void k_triangle_boxes(ColliderView col, int swept, double *tri_box,
                                                           double *records, double *bound_partials)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    __shared__ double lo_s[256][3], hi_s[256][3];
    V3 clo = mk(1e300, 1e300, 1e300), chi = mk(-1e300, -1e300, -1e300);
    if (t < col.triangle_count) {
        const uint32_t *t3 = col.triangles + 3 * size_t(t);
        V3 lo = mk(1e300, 1e300, 1e300), hi = mk(-1e300, -1e300, -1e300);
        V3 centroid = mk(0, 0, 0);
        V3 v[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = vertex(col, t3[c]);
            lo = vmin(lo, v[c]);
            hi = vmax(hi, v[c]);
            centroid = add(centroid, v[c]);
        }
        if (records) {
            double *r = records + kTriRecord * size_t(t);
            store3(r, lo);
            store3(r + 3, hi);
#pragma unroll
            for (int c = 0; c < 3; ++c) store3(r + kRecordVertices + 3 * c, v[c]);
            r[kTriRecord - 1] = 0.0;
        }
        if (swept) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const V3 o = vertex_prev(col, t3[c]);
                lo = vmin(lo, o);
                hi = vmax(hi, o);
            }
        }
        store3(tri_box + 6 * size_t(t), lo);
        store3(tri_box + 6 * size_t(t) + 3, hi);
        centroid = mul(centroid, 1.0 / 3.0);
        clo = centroid;
        chi = centroid;
    }
    lo_s[threadIdx.x][0] = clo.x; lo_s[threadIdx.x][1] = clo.y; lo_s[threadIdx.x][2] = clo.z;
    hi_s[threadIdx.x][0] = chi.x; hi_s[threadIdx.x][1] = chi.y; hi_s[threadIdx.x][2] = chi.z;
    __syncthreads();
    for (int s = 256 / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < 3; ++k) {
                lo_s[threadIdx.x][k] = fmin(lo_s[threadIdx.x][k], lo_s[threadIdx.x + s][k]);
                hi_s[threadIdx.x][k] = fmax(hi_s[threadIdx.x][k], hi_s[threadIdx.x + s][k]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            bound_partials[6 * size_t(blockIdx.x) + k] = lo_s[0][k];
            bound_partials[6 * size_t(blockIdx.x) + 3 + k] = hi_s[0][k];
        }
}

// Folds the per-block centroid bounds (one block of 256 threads).
__global__ __launch_bounds__(256) void k_bounds_final(const double *partials, uint32_t count, double *bounds)
{
    __shared__ double lo_s[256][3], hi_s[256][3];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i = threadIdx.x; i < count; i += 256)
        for (int k = 0; k < 3; ++k) {
            lo[k] = fmin(lo[k], partials[6 * size_t(i) + k]);
            hi[k] = fmax(hi[k], partials[6 * size_t(i) + 3 + k]);
        }
    for (int k = 0; k < 3; ++k) {
        lo_s[threadIdx.x][k] = lo[k];
        hi_s[threadIdx.x][k] = hi[k];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s)
            for (int k = 0; k < 3; ++k) {
                lo_s[threadIdx.x][k] = fmin(lo_s[threadIdx.x][k], lo_s[threadIdx.x + s][k]);
                hi_s[threadIdx.x][k] = fmax(hi_s[threadIdx.x][k], hi_s[threadIdx.x + s][k]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) {
            bounds[k] = lo_s[0][k];
            bounds[3 + k] = hi_s[0][k];
        }
}

// 30-bit Morton code of each triangle centroid in the centroid bounds.
__global__ void k_morton(ColliderView col, const double *bounds, uint32_t *keys, uint32_t *ids)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= col.triangle_count) return;
    const uint32_t *t3 = col.triangles + 3 * size_t(t);
    const V3 c = mul(add(add(vertex(col, t3[0]), vertex(col, t3[1])), vertex(col, t3[2])), 1.0 / 3.0);
    const double cc[3] = {c.x, c.y, c.z};
    uint32_t code = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double span = bounds[3 + k] - bounds[k];
        const double u = span > 0.0 ? (cc[k] - bounds[k]) / span : 0.5;
        const uint32_t b = uint32_t(fmin(fmax(u * 1024.0, 0.0), 1023.0));
        code |= spread_bits(b) << (2 - k);
    }
    keys[t] = code;
    ids[t] = t;
}

__device__ __forceinline__ int common_prefix(const uint32_t *keys, int count, int i, int j)
{
    if (j < 0 || j >= count) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz(uint32_t(i ^ j));
    return __clz(a ^ b);
}

// Karras 2012 hierarchy over sorted codes: internal node i spans a key range
// split at its highest differing bit. Leaves are stored as -1 - triangle.
__global__ void k_hierarchy(uint32_t count, const uint32_t *keys, const uint32_t *sorted_ids, BvhNode *nodes,
                            int *parent_internal, int *parent_leaf, unsigned *counters)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = int(count);
    if (n == 1) {
        if (i == 0) {
            nodes[0].child[0] = -1 - int(sorted_ids[0]);
            nodes[0].child[1] = -1 - int(sorted_ids[0]);
            nodes[0].range[0] = nodes[0].range[1] = 0;
            parent_leaf[0] = 0;
            counters[0] = 0;
        }
        return;
    }
    if (i >= n - 1) return;
    counters[i] = 0;
    const int d = (common_prefix(keys, n, i, i + 1) - common_prefix(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = common_prefix(keys, n, i, i - d);
    int lmax = 2;
    while (common_prefix(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (common_prefix(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = common_prefix(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) / 2;
        if (common_prefix(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    BvhNode &node = nodes[i];
    node.range[0] = lo;
    node.range[1] = hi;
    if (lo == gamma) {
        node.child[0] = -1 - int(sorted_ids[gamma]);
        parent_leaf[gamma] = i << 1;
    } else {
        node.child[0] = gamma;
        parent_internal[gamma] = i << 1;
    }
    if (hi == gamma + 1) {
        node.child[1] = -1 - int(sorted_ids[gamma + 1]);
        parent_leaf[gamma + 1] = (i << 1) | 1;
    } else {
        node.child[1] = gamma + 1;
        parent_internal[gamma + 1] = (i << 1) | 1;
    }
    if (i == 0) parent_internal[0] = -1;
}

__global__ void k_zero_counters(unsigned *counters, uint32_t count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) counters[i] = 0;
}

// Bottom-up boxes: each leaf writes its triangle box into its parent; the
// second child to arrive at a node writes the union into the grandparent.
// Workgroup w starts at sorted leaves [256 w, 256 w + 256). A node whose
// leaf range lies inside that chunk is reached only by threads of w, so its
// boxes and counter need workgroup-scope ordering only; agent scope, whose
// release writes back the XCD's L2, is paid only on the few nodes whose
// range crosses a chunk boundary.
__global__ void k_refit(uint32_t count, const uint32_t *sorted_ids, const double *tri_box, const int *parent_internal,
                        const int *parent_leaf, BvhNode *nodes, unsigned *counters)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    if (count == 1) {
        const double *b = tri_box + 6 * size_t(sorted_ids[0]);
        for (int c = 0; c < 6; ++c) nodes[0].box[c] = b[c];
        for (int c = 0; c < 3; ++c) {
            nodes[0].box[6 + c] = 1e300;
            nodes[0].box[9 + c] = -1e300;
        }
        return;
    }
    // The leaf box is read as six separate 64-bit loads. hipcc (ROCm 7.2,
    // gfx950) otherwise merges them into 128-bit loads and, in this loop,
    // reuses the upper halves of those register tuples while the values are
    // still live, storing garbage into every odd slot entry.
    double box[6];
    const volatile double *b = tri_box + 6 * size_t(sorted_ids[k]);
#pragma unroll
    for (int c = 0; c < 6; ++c) box[c] = b[c];
    int link = parent_leaf[k];
    // The hierarchy is at most 64 levels deep (30-bit codes, index tie-break).
    // Box slots are written and read with atomic stores and loads ordered by
    // the node counter's acquire-release, at the scope of the threads that
    // can meet at the node, so no stale cache line is ever read.
    const int chunk_lo = int(blockIdx.x * blockDim.x), chunk_hi = chunk_lo + int(blockDim.x);
    for (int level = 0; level < 128; ++level) {
        const int parent = link >> 1, side = link & 1;
        double *slot = nodes[parent].box + 6 * side;
        double both[12];
        const bool local = nodes[parent].range[0] >= chunk_lo && nodes[parent].range[1] < chunk_hi;
        if (local) {
#pragma unroll
            for (int c = 0; c < 6; ++c)
                __hip_atomic_store(slot + c, box[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (__hip_atomic_fetch_add(&counters[parent], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                return;
#pragma unroll
            for (int c = 0; c < 12; ++c)
                both[c] = __hip_atomic_load(nodes[parent].box + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c)
                __hip_atomic_store(slot + c, box[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(&counters[parent], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 0) return;
#pragma unroll
            for (int c = 0; c < 12; ++c)
                both[c] = __hip_atomic_load(nodes[parent].box + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            box[c] = fmin(both[c], both[6 + c]);
            box[3 + c] = fmax(both[3 + c], both[9 + c]);
        }
        link = parent_internal[parent];
        if (link < 0) return;
    }
}

// Two-phase bottom-up boxes. Workgroup w (kRefitChunk threads) starts at
// sorted leaves [kRefitChunk w, kRefitChunk (w + 1)); every node whose leaf
// range lies inside that chunk is reached only by the workgroup's threads,
// so phase 1 combines their boxes through LDS slots and counters. A node
// whose range crosses a chunk boundary receives its child boxes in global
// memory and an arrival on its global counter; the arrival that completes
// both children queues it. Phase 2, one workgroup after the kernel boundary,
// climbs from the queued nodes to the root. No acquire or release beyond
// the workgroup is needed. Chunks of 1024 leaves (100 KB of LDS, one
// workgroup per CU) leave phase 2 a quarter of the nodes that chunks of 256
// did; measured on the 450,360-triangle body, the pair of launches takes
// about 120 us less per substep.
constexpr int kRefitChunk = 1024;

// A node whose range lies inside a chunk has one of its range's ends as
// its index (Karras), so the chunk's nodes are among internal nodes
// [base, base + kRefitChunk): their ranges and parent links are read into
// LDS once, and the climb reads no global memory but the leaf boxes.
struct RefitLocal {
    double slot[kRefitChunk][12];  // child boxes of the chunk's nodes
    unsigned count[kRefitChunk];
    int first[kRefitChunk], last[kRefitChunk], parent[kRefitChunk];  // of internal nodes base + i
};

__device__ inline void refit_local_init(RefitLocal &s, uint32_t count, const int *parent_internal,
                                        const BvhNode *nodes)
{
    const uint32_t i = threadIdx.x, node = blockIdx.x * kRefitChunk + i;
    s.count[i] = 0;
    const bool inside = count > 1 && node < count - 1;
    s.first[i] = inside ? nodes[node].range[0] : -1;
    s.last[i] = inside ? nodes[node].range[1] : -1;
    s.parent[i] = inside ? parent_internal[node] : -1;
}

__device__ inline void refit_local_climb(RefitLocal &s, uint32_t count, const uint32_t *sorted_ids,
                                         const double *tri_box, const int *parent_internal, const int *parent_leaf,
                                         BvhNode *nodes, unsigned *counters, uint32_t *ready, unsigned *ready_count)
{
    const uint32_t base = blockIdx.x * kRefitChunk;
    const uint32_t k = base + threadIdx.x;
    if (k >= count) return;
    // six separate 64-bit loads (see k_refit)
    double box[6];
    const volatile double *b = tri_box + 6 * size_t(sorted_ids[k]);
#pragma unroll
    for (int c = 0; c < 6; ++c) box[c] = b[c];
    int link = parent_leaf[k];
    for (int level = 0; level < 128; ++level) {
        const int parent = link >> 1, side = link & 1;
        const int li = parent - int(base);
        const bool local = li >= 0 && li < kRefitChunk && s.first[li] >= int(base) &&
                           s.last[li] < int(base) + kRefitChunk;
        if (!local) {
#pragma unroll
            for (int c = 0; c < 6; ++c) nodes[parent].box[6 * side + c] = box[c];
            if (__hip_atomic_fetch_add(&counters[parent], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1)
                ready[__hip_atomic_fetch_add(ready_count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)] =
                    uint32_t(parent);
            return;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c)
            __hip_atomic_store(&s.slot[li][6 * side + c], box[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (__hip_atomic_fetch_add(&s.count[li], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0) return;
        double both[12];
#pragma unroll
        for (int c = 0; c < 12; ++c)
            both[c] = __hip_atomic_load(&s.slot[li][c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int c = 0; c < 12; ++c) nodes[parent].box[c] = both[c];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            box[c] = fmin(both[c], both[6 + c]);
            box[3 + c] = fmax(both[3 + c], both[9 + c]);
        }
        link = s.parent[li];
        if (link < 0) return;
    }
}

__global__ __launch_bounds__(kRefitChunk) void k_refit_local(uint32_t count, const uint32_t *sorted_ids,
                                                             const double *tri_box, const int *parent_internal,
                                                             const int *parent_leaf, BvhNode *nodes,
                                                             unsigned *counters, uint32_t *ready,
                                                             unsigned *ready_count)
{
    __shared__ RefitLocal s;
    refit_local_init(s, count, parent_internal, nodes);
    __syncthreads();
    refit_local_climb(s, count, sorted_ids, tri_box, parent_internal, parent_leaf, nodes, counters, ready,
                      ready_count);
}

// Phase 2 (one workgroup): from each queued node (both children's boxes in
// place) combine its box and climb, the second arrival at a node going on.
__global__ __launch_bounds__(1024) void k_refit_top(const int *parent_internal, BvhNode *nodes, unsigned *counters,
                                                    const uint32_t *ready, const unsigned *ready_count)
{
    const unsigned n = *ready_count;
    for (unsigned r = threadIdx.x; r < n; r += blockDim.x) {
        int node = int(ready[r]);
        double both[12];
#pragma unroll
        for (int c = 0; c < 12; ++c)
            both[c] = __hip_atomic_load(nodes[node].box + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        for (int level = 0; level < 128; ++level) {
            double box[6];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                box[c] = fmin(both[c], both[6 + c]);
                box[3 + c] = fmax(both[3 + c], both[9 + c]);
            }
            const int link = parent_internal[node];
            if (link < 0) break;
            const int parent = link >> 1, side = link & 1;
#pragma unroll
            for (int c = 0; c < 6; ++c)
                __hip_atomic_store(nodes[parent].box + 6 * side + c, box[c], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
            if (__hip_atomic_fetch_add(&counters[parent], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
                break;
#pragma unroll
            for (int c = 0; c < 12; ++c)
                both[c] = __hip_atomic_load(nodes[parent].box + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            node = parent;
        }
    }
}

// Tree check used by tests. Counts into out[]: 0 triangle boxes not inside
// the box stored for an ancestor, 1 broken parent chains, 2 out-of-order
// sort keys, 3 leaves whose own stored box differs from their triangle box,
// 4 stored child boxes with lo > hi, 5 deepest leaf.
__global__ void k_check_bvh(uint32_t count, const uint32_t *keys_sorted, const uint32_t *sorted_ids,
                            const double *tri_box, const int *parent_internal, const int *parent_leaf,
                            const BvhNode *nodes, unsigned long long *out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count || count < 2) return;
    if (k + 1 < count && keys_sorted[k] > keys_sorted[k + 1]) atomicAdd(out + 2, 1ull);
    if (k + 1 < count) {
        for (int side = 0; side < 2; ++side) {
            const double *bx = nodes[k].box + 6 * side;
            if (bx[0] > bx[3] || bx[1] > bx[4] || bx[2] > bx[5]) atomicAdd(out + 4, 1ull);
        }
    }
    const double *b = tri_box + 6 * size_t(sorted_ids[k]);
    int link = parent_leaf[k];
    int expected_child = -1 - int(sorted_ids[k]);
    for (int level = 0; level < 130; ++level) {
        if (link < 0 || (link >> 1) >= int(count - 1)) {
            atomicAdd(out + 1, 1ull);
            return;
        }
        const int parent = link >> 1, side = link & 1;
        if (nodes[parent].child[side] != expected_child) atomicAdd(out + 1, 1ull);
        const double *slot = nodes[parent].box + 6 * side;
        if (level == 0)
            for (int c = 0; c < 6; ++c)
                if (slot[c] != b[c]) {
                    atomicAdd(out + 3, 1ull);
                    break;
                }
        for (int c = 0; c < 3; ++c)
            if (b[c] < slot[c] || b[3 + c] > slot[3 + c]) {
                atomicAdd(out, 1ull);
                break;
            }
        if (parent == 0) {
            atomicMax(out + 5, (unsigned long long)(level + 1));
            return;
        }
        expected_child = parent;
        link = parent_internal[parent];
    }
    atomicAdd(out + 1, 1ull);
}

}  // namespace dev
}  // namespace khs
