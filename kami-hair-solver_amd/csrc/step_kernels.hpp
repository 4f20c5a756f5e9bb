// Device kernels of the incremental-potential hair step. Included by
// solver.hip only (one translation unit, no relocatable device code).
//
// Tiled kernels: wavefront w owns nodes [63 w, 63 w + 63); lane l handles the
// element whose left node is L = 63 w - 1 + l and, for l >= 1, owns node L.
// Lane 0 recomputes the last element of the previous tile, so every node
// receives its left element's contribution from the lane below it through a
// DPP lane shift, and its own element's from its own lane: no atomics, and a
// fixed summation order (inertia, left element, right element) that matches
// the oracle's node-then-element order.
#pragma once

#include "kernels.hpp"
#include "lbvh_build.hpp"
#include "model.hpp"

#include <type_traits>
#include <utility>

namespace khs {
namespace dev {

__device__ __forceinline__ double inf_double() { return __longlong_as_double(0x7ff0000000000000LL); }

// Value of the lane below (lane 0 receives 0), a DPP wavefront shift right.
__device__ __forceinline__ double from_lane_below(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, false);
    const int rhi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, false);
    return __hiloint2double(rhi, rlo);
}
// Value of the lane above (lane 63 receives 0), a DPP wavefront shift left.
__device__ __forceinline__ double from_lane_above(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int rlo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, false);
    const int rhi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, false);
    return __hiloint2double(rhi, rlo);
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ unsigned wave_or(unsigned v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v |= __shfl_xor(v, o, kWave);
    return v;
}

__device__ inline Partial empty_partial()
{
    Partial p;
#pragma unroll
    for (int k = 0; k < 8; ++k) p.sum[k] = 0.0;
    p.min_gap = inf_double();
    p.alpha = 1.0;
    p.max_translation = 0.0;
    p.candidates = p.active = 0;
    p.infeasible = p.failed = 0;
    return p;
}

// Block-wide fixed-order reduction of one Partial per thread; thread 0
// writes the result (also to `mirror`, host memory the host reads after
// its wait, when given).
template <int BLOCK = kBlock>
__device__ inline void block_reduce(Partial p, Partial *out, Partial *mirror = nullptr)
{
    __shared__ Partial waves[BLOCK / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < 8; ++k) p.sum[k] = wave_sum(p.sum[k]);
    p.min_gap = wave_min(p.min_gap);
    p.alpha = wave_min(p.alpha);
    p.max_translation = wave_max(p.max_translation);
    p.candidates = wave_sum_u64(p.candidates);
    p.active = wave_sum_u64(p.active);
    p.infeasible = wave_or(p.infeasible);
    p.failed = wave_or(p.failed);
    if (lane == 0) waves[wave] = p;
    __syncthreads();
    if (threadIdx.x == 0) {
        Partial r = waves[0];
        for (int w = 1; w < int(blockDim.x / kWave); ++w) {
            const Partial &o = waves[w];
            for (int k = 0; k < 8; ++k) r.sum[k] += o.sum[k];
            r.min_gap = fmin(r.min_gap, o.min_gap);
            r.alpha = fmin(r.alpha, o.alpha);
            r.max_translation = fmax(r.max_translation, o.max_translation);
            r.candidates += o.candidates;
            r.active += o.active;
            r.infeasible |= o.infeasible;
            r.failed |= o.failed;
        }
        *out = r;
        if (mirror) *mirror = r;
    }
}

// The Newton convergence test on the device, the same comparisons as the
// host's (cuda_backend.cu:662-672): stop = |g| <= abs_tol or |g| / initial
// <= rel_tol, with initial < 0 meaning this gradient is the substep's first.
// Work enqueued ahead for the next Newton iteration skips itself on stop.
struct ConvergenceTest {
    double abs_tol, rel_tol, initial;
    unsigned *stop;  // null: no test
};

__device__ inline void converged_test(const Partial &grad, const ConvergenceTest &test)
{
    const double residual = sqrt(grad.sum[kGradNorm2]);
    const double first = test.initial < 0.0 ? residual : test.initial;
    const double relative = residual / fmax(first, 1e-30);
    *test.stop = (residual <= test.abs_tol || relative <= test.rel_tol) ? 1u : 0u;
}

// Per-block results of the list passes, one word per block: the narrow
// phase's candidate count, the CCD pass's smallest step. Their sum and
// minimum do not depend on the order they are folded in.
struct ListPartials {
    const unsigned long long *candidates;  // or null
    const double *alpha;                   // or null
    uint32_t count;
};

// Folds per-block partials in index order (one block of kBlock threads),
// and the list passes' per-block words, then applies the convergence test to
// the result when one is given. Thread t folds partials t, t + kReduceBlock,
// ... in that order. A round requests up to kReduceBatch of its partials and
// kListBatch of its list words from memory before folding any, so their
// loads overlap instead of each waiting for the one before; one round covers
// the 1,280 partials and 3,072 list words of the step's passes (the
// evaluations leave one partial per 248 nodes, the swept step bound 3,072
// words), so the kernel waits on memory once.
constexpr int kReduceBlock = 256;
constexpr int kReduceBatch = 5;
constexpr int kListBatch = 12;
__global__ __launch_bounds__(kReduceBlock) void k_reduce(const Partial *partials, uint32_t count, Partial *out,
                                                         Partial *mirror, ConvergenceTest test, ListPartials lists)
{
    Partial acc = empty_partial();
    constexpr uint32_t kStride = kReduceBlock, kSpan = kReduceBlock * kReduceBatch,
                       kListSpan = kReduceBlock * kListBatch;
    for (uint32_t r = 0; r * kListSpan < lists.count || r * kSpan < count; ++r) {
        unsigned long long c[kListBatch];
        double a[kListBatch];
        Partial o[kReduceBatch];
        const uint32_t l0 = threadIdx.x + r * kListSpan, i0 = threadIdx.x + r * kSpan;
#pragma unroll
        for (int b = 0; b < kListBatch; ++b) {
            const uint32_t i = l0 + b * kStride;
            c[b] = (lists.candidates && i < lists.count) ? lists.candidates[i] : 0ull;
            a[b] = (lists.alpha && i < lists.count) ? lists.alpha[i] : 1.0;
        }
#pragma unroll
        for (int b = 0; b < kReduceBatch; ++b)
            o[b] = i0 + b * kStride < count ? partials[i0 + b * kStride] : empty_partial();
#pragma unroll
        for (int b = 0; b < kListBatch; ++b) {
            acc.candidates += c[b];
            acc.alpha = fmin(acc.alpha, a[b]);
        }
#pragma unroll
        for (int b = 0; b < kReduceBatch; ++b) {
            if (i0 + b * kStride >= count) break;
            for (int k = 0; k < 8; ++k) acc.sum[k] += o[b].sum[k];
            acc.min_gap = fmin(acc.min_gap, o[b].min_gap);
            acc.alpha = fmin(acc.alpha, o[b].alpha);
            acc.max_translation = fmax(acc.max_translation, o[b].max_translation);
            acc.candidates += o[b].candidates;
            acc.active += o[b].active;
            acc.infeasible |= o[b].infeasible;
            acc.failed |= o[b].failed;
        }
    }
    block_reduce<kReduceBlock>(acc, out, mirror);
    if (test.stop && threadIdx.x == 0) converged_test(*out, test);
}


// The contact radius terms and stiffness weights of the element whose left
// node is L: the material's (precomputed on the host) or, with per-strand
// materials, its strand's, formed in the same operation order.
__device__ __forceinline__ double radius_offset_of(const StepParams &p, size_t L)
{
    return p.material ? p.material[2 * L + 1] + p.collider_offset : p.radius_offset;
}
__device__ __forceinline__ double contact_pad_of(const StepParams &p, size_t L)
{
    return p.material ? (p.material[2 * L + 1] + p.collider_offset) + p.barrier : p.contact_pad;
}
__device__ __forceinline__ double sweep_target_of(const StepParams &p, size_t L)
{
    return p.material ? (p.material[2 * L + 1] + p.collider_offset) + p.minimum_gap : p.sweep_target;
}
constexpr double kPi = 3.1415926535897932384626433832795;
// kGA, kGA, EA, EI, EI, 2GI (cuda_backend.cu:905-910); MAT selects at
// compile time whether per-strand materials can be present, so the
// material-wide kernels keep their registers.
template <bool MAT>
__device__ __forceinline__ void weights_of(const StepParams &p, size_t L, double w[6])
{
    if (!MAT || !p.material) {
#pragma unroll
        for (int s = 0; s < 6; ++s) w[s] = p.weights[s];
        return;
    }
    const double E = p.material[2 * L], r = p.material[2 * L + 1];
    const double area = kPi * r * r, inertia = kPi * r * r * r * r / 4.0;
    const double G = E / (2.0 * (1.0 + p.poisson));
    w[0] = p.shear_correction * G * area;
    w[1] = p.shear_correction * G * area;
    w[2] = E * area;
    w[3] = E * inertia;
    w[4] = E * inertia;
    w[5] = 2.0 * G * inertia;
}

__device__ __forceinline__ int internal_count(const ColliderView &c)
{
    return c.triangle_count > 1 ? int(c.triangle_count) - 1 : 1;
}

// Closest candidate triangle of a segment from the hierarchy; ties in
// distance go to the lower triangle index.
__device__ inline bool closest_triangle(const ColliderView &c, V3 p0, V3 p1, double pad, Closest &best,
                                        uint32_t &best_tri, unsigned long long &count)
{
    best = Closest{1.0e300, 0.0, mk(1, 0, 0)};
    best_tri = kNoTriangle;
    bvh_query(c.bvh, internal_count(c), c.leaf_box, vmin(p0, p1), vmax(p0, p1), pad, [&](uint32_t t) {
        ++count;
        const uint32_t *tri = c.triangles + 3 * size_t(t);
        const Closest cand = closest_segment_triangle(p0, p1, vertex(c, tri[0]), vertex(c, tri[1]), vertex(c, tri[2]));
        if (cand.distance < best.distance || (cand.distance == best.distance && t < best_tri)) {
            best = cand;
            best_tri = t;
        }
    });
    return best_tri != kNoTriangle;
}

// Query points of element L in the three kinds of pass.
enum QueryMode { kAtState = 0, kAtTrial = 1, kSwept = 2 };

// Endpoints at q, or at alpha d + q_base in a line search (the expression
// the evaluation uses, so the same rounding).
template <int MODE>
__device__ __forceinline__ void element_points(uint32_t L, const double *q, const double *q_base, const double *dir,
                                               double alpha, V3 &p0, V3 &p1)
{
    if (MODE == kAtTrial) {
        const double *b0 = q_base + 6 * size_t(L), *d0 = dir + 6 * size_t(L);
        p0 = mk(alpha * d0[0] + b0[0], alpha * d0[1] + b0[1], alpha * d0[2] + b0[2]);
        p1 = mk(alpha * d0[6] + b0[6], alpha * d0[7] + b0[7], alpha * d0[8] + b0[8]);
    } else {
        p0 = load3(q + 6 * size_t(L));
        p1 = load3(q + 6 * size_t(L + 1));
    }
}

// Unpadded query box of element L: the segment, or for kSwept the segment
// at q and at q + d (the box the line-search bound queries).
template <int MODE>
__device__ __forceinline__ void query_box(uint32_t L, const double *q, const double *q_base, const double *dir,
                                          double alpha, V3 &lo, V3 &hi)
{
    V3 p0, p1;
    element_points<MODE>(L, q, q_base, dir, alpha, p0, p1);
    if (MODE == kSwept) {
        const V3 d0 = load3(dir + 6 * size_t(L)), d1 = load3(dir + 6 * size_t(L + 1));
        lo = vmin(vmin(p0, p1), vmin(add(p0, d0), add(p1, d1)));
        hi = vmax(vmax(p0, p1), vmax(add(p0, d0), add(p1, d1)));
    } else {
        lo = vmin(p0, p1);
        hi = vmax(p0, p1);
    }
}

// [lo - pad, hi + pad] inside the cached box with a margin far above the
// rounding of either box test, so every triangle passing the query's test
// passed the list's (false for NaN). The margin follows the cached box's own
// coordinate, so the test is monotone in the query box: a query box inside
// one that is covered is covered (each rounded subtraction and addition is
// monotone in the query coordinate).
__device__ __forceinline__ bool box_covered(const double *box, V3 lo, V3 hi, double pad)
{
    const double l[3] = {lo.x - pad, lo.y - pad, lo.z - pad}, h[3] = {hi.x + pad, hi.y + pad, hi.z + pad};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double ml = 1.0e-12 + 1.0e-9 * fabs(box[k]), mh = 1.0e-12 + 1.0e-9 * fabs(box[3 + k]);
        ok = ok && box[k] <= l[k] - ml && box[3 + k] >= h[k] + mh;
    }
    return ok;
}

// Where the skin holds more triangles than a list can (fine or sliver
// triangles, as where a mesh's rings meet at a pole), a list is gathered
// again for a box grown by only kThinSkin of the skin, so the element keeps
// a list (refreshed more often) instead of every query walking the
// hierarchy. thin_box shrinks a box grown by pad + skin to that.
constexpr double kThinSkin = 1.0 / 16.0;
__device__ __forceinline__ void thin_box(double skin, V3 &lo, V3 &hi)
{
    const double s = skin * (1.0 - kThinSkin);
    lo = mk(lo.x + s, lo.y + s, lo.z + s);
    hi = mk(hi.x - s, hi.y - s, hi.z - s);
}

// Before a query pass: every contact element whose query box (grown by
// pad, the largest pad of the passes) is not inside its cached box walks
// the hierarchy for that box grown by the skin and stores the overlapping
// triangles. With `full` (new collider boxes) every element does.
template <int MODE>
__global__ void k_contact_refresh(RodView m, ColliderView col, ContactCache cc, const double *q, const double *q_base,
                                  const double *dir, const double *alpha_at, double pad, int full)
{
    const double alpha = MODE == kAtTrial ? *alpha_at : 0.0;
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= m.node_count) return;
    const uint32_t fl = m.flags[L];
    if (!(fl & kNodeHasElement) || !(fl & kNodeContact)) return;
    V3 lo, hi;
    query_box<MODE>(L, q, q_base, dir, alpha, lo, hi);
    double *box = cc.box + 6 * size_t(L);
    if (!full && box_covered(box, lo, hi, pad)) return;
    const double grow = pad + cc.skin;
    V3 blo = mk(lo.x - grow, lo.y - grow, lo.z - grow), bhi = mk(hi.x + grow, hi.y + grow, hi.z + grow);
    uint32_t *tris = cc.tris + size_t(L) * kCacheCapacity;
    uint32_t n = 0;
    auto gather = [&] {
        n = 0;
        bvh_visit(col.bvh, internal_count(col), col.leaf_box, blo, bhi, 0.0, [&](uint32_t t) {
            if (n < kCacheCapacity) tris[n] = t;
            ++n;
            return n <= kCacheCapacity;  // an overflowed list is not used
        });
    };
    gather();
    if (n > kCacheCapacity) {
        thin_box(cc.skin, blo, bhi);
        gather();
    }
    if (n == 0) {
        // Nothing near: keep a box sixteen skins wider if it is empty too, so
        // hair moving through free space rarely gathers again.
        const double far = pad + 16.0 * cc.skin;
        const V3 flo = mk(lo.x - far, lo.y - far, lo.z - far), fhi = mk(hi.x + far, hi.y + far, hi.z + far);
        bool empty = true;
        bvh_visit(col.bvh, internal_count(col), col.leaf_box, flo, fhi, 0.0, [&](uint32_t) {
            empty = false;
            return false;
        });
        if (empty) {
            blo = flo;
            bhi = fhi;
        }
    }
    cc.count[L] = n <= kCacheCapacity ? n : kCacheOverflow;
    store3(box, blo);
    store3(box + 3, bhi);
    if (n > 0 && !(fl & kNodeFixed) && !cc.listed[L]) {
        cc.listed[L] = 1;
        cc.list[atomicAdd(cc.list_count, 1u)] = L;
    }
}

// Lanes per listed element in the list passes (one wavefront).
constexpr int kListLanes = kWave;

// Entries of a wavefront's breadth-first traversal queue (LDS).
constexpr uint32_t kGatherQueue = 1024;

// Wavefront-cooperative walk of the hierarchy for one box (pad 0, leaves
// tested against col.leaf_box): each step takes up to 64 queued nodes, one
// per lane, tests both children, and compacts the overlapping internal
// children onto the queue and the overlapping triangles into out (first
// `cap` kept). Returns the count of triangles found, or cap + 1 once it
// exceeds cap or the queue would overflow (the list is then not used).
// With stop_at_first it stops at the first triangle. All lanes call it with
// the same box; every quantity steering the loop is wave-uniform.
__device__ inline uint32_t wave_gather(const ColliderView &col, V3 lo, V3 hi, uint32_t *out, uint32_t cap,
                                       volatile int *queue, bool stop_at_first)
{
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned long long below = __lanemask_lt();
    const int ic = internal_count(col);
    uint32_t head = 0, tail = 1, n = 0;
    if (lane == 0) queue[0] = 0;
    // A tree of ic internal nodes queues each at most once; the bound only
    // ends the walk should the tree be malformed.
    while (head < tail && head <= uint32_t(ic) + 1u) {
        const uint32_t take = (tail - head) < uint32_t(kWave) ? (tail - head) : uint32_t(kWave);
        const int node = uint32_t(lane) < take ? queue[(head + lane) % kGatherQueue] : -1;
        head += take;
        bool push[2] = {false, false}, emit[2] = {false, false};
        int child[2] = {-1, -1};
        if (node >= 0) {
            const BvhNode &nd = col.bvh[node];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const V3 bl = mk(nd.box[6 * c], nd.box[6 * c + 1], nd.box[6 * c + 2]);
                const V3 bu = mk(nd.box[6 * c + 3], nd.box[6 * c + 4], nd.box[6 * c + 5]);
                const int ch = nd.child[c];
                child[c] = ch;
                if (!overlap(lo, hi, bl, bu, 0.0) || ch >= ic) continue;
                if (ch >= 0) {
                    push[c] = true;
                } else if (col.leaf_box) {
                    const double *b = col.leaf_box + kTriRecord * size_t(-1 - ch);
                    emit[c] = overlap(lo, hi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), 0.0);
                } else {
                    emit[c] = true;
                }
            }
        }
        const unsigned long long p0 = __ballot(push[0]), p1 = __ballot(push[1]);
        const uint32_t pushes = __popcll(p0) + __popcll(p1);
        if (tail - head + pushes > kGatherQueue) return cap + 1;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned long long pm = c == 0 ? p0 : p1;
            if (push[c]) queue[(tail + __popcll(pm & below)) % kGatherQueue] = child[c];
            tail += __popcll(pm);
            const unsigned long long em = __ballot(emit[c]);
            const uint32_t at = n + __popcll(em & below);
            if (emit[c] && at < cap) out[at] = uint32_t(-1 - child[c]);
            n += __popcll(em);
        }
        if (n > cap || (stop_at_first && n > 0)) return n > cap ? cap + 1 : n;
    }
    return n;
}

// Where a trial's step comes from: already in alpha_at (kAlphaStored),
// the CCD bound of the last reduction (kAlphaFromCcd), or a value the host
// chose (kAlphaGiven). The first kernel of a trial evaluation resolves it
// and stores it in alpha_at for the kernels after it.
enum AlphaSource { kAlphaStored = 0, kAlphaFromCcd = 1, kAlphaGiven = 2 };
struct TrialStep {
    const Partial *ccd;
    double value;
    int source;
    int has_collider;
};

// The CCD bound, scaled by 0.9 when it limits the step (cuda_backend.cu:623-648).
__device__ __forceinline__ double step_from_ccd(const Partial *ccd, int has_collider)
{
    double a = 1.0;
    if (has_collider) {
        a = ccd->alpha;
        if (a < 1.0) a *= 0.9;
    }
    return a;
}

__device__ __forceinline__ double resolve_step(const TrialStep &s, double *alpha_at)
{
    const double a = s.source == kAlphaFromCcd ? step_from_ccd(s.ccd, s.has_collider)
                                               : (s.source == kAlphaGiven ? s.value : *alpha_at);
    if (s.source != kAlphaStored && blockIdx.x == 0 && threadIdx.x == 0) *alpha_at = a;
    return a;
}

// Partial refresh: every contact element whose query box (grown by pad plus
// the rounding margin) left its cached box gets a new box (query box grown
// by pad + skin), and its wavefront gathers the triangles overlapping each
// such box in turn, breadth-first through an LDS queue; an empty result
// keeps a box fifteen skins wider still when that is empty too. Elements with
// a fixed left node take part only `with_roots` (the friction pass, the one
// reader of their lists), and keep lists but stay off the listed queue.
// The swept pass (boxes over a whole step, where many neighbouring elements
// leave their boxes at once) stores the new box and queues the element for
// the step-bound kernel after it (k_swept_ccd) instead, which spreads the
// gathers over its wavefronts.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_contact_check(RodView m, ColliderView col, ContactCache cc,
                                                          const double *q, const double *q_base, const double *dir,
                                                          double *alpha_at, TrialStep step, double pad,
                                                          int with_roots, const double *grad, Partial *partials)
{
    __shared__ int queues[kTilesPerBlock][kGatherQueue];
    volatile int *queue = queues[threadIdx.x / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const double alpha = MODE == kAtTrial ? resolve_step(step, alpha_at) : 0.0;
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    // (the swept check appends nothing to the listed queue)
    if (MODE == kSwept && L == 0) *cc.listed_snapshot = *cc.list_count;
    // every lane stays to the end: the gathers use the whole wavefront
    bool need = false;
    V3 nlo = mk(0, 0, 0), nhi = nlo;
    if (L < m.node_count && m.node_count >= 2) {
        // The flags, the element's state and its cached box are requested
        // together (the last node, which has no element, reads the element
        // before it and discards it).
        const uint32_t Lq = L + 1 < m.node_count ? L : m.node_count - 2;
        const uint32_t fl = m.flags[L];
        V3 lo, hi;
        query_box<MODE>(Lq, q, q_base, dir, alpha, lo, hi);
        double box[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) box[k] = cc.box[6 * size_t(Lq) + k];
        if ((fl & kNodeHasElement) && (fl & kNodeContact) && (!(fl & kNodeFixed) || with_roots)) {
            if (!box_covered(box, lo, hi, pad)) {
                const double grow = pad + cc.skin;
                need = true;
                nlo = mk(lo.x - grow, lo.y - grow, lo.z - grow);
                nhi = mk(hi.x + grow, hi.y + grow, hi.z + grow);
                if constexpr (MODE == kSwept) {
                    store3(cc.box + 6 * size_t(L), nlo);
                    store3(cc.box + 6 * size_t(L) + 3, nhi);
                    cc.pending_epoch[L] = cc.epoch;
                    cc.pending[atomicAdd(cc.pending_count, 1u)] = L;
                }
            }
        }
    }
    if constexpr (MODE == kSwept) {
        // The slope g.d of the step the swept pass bounds, a partial per
        // block (the step-bound kernel after it leaves the smallest step).
        Partial acc = empty_partial();
        if (L < m.node_count) {
            double gd = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) gd += grad[6 * size_t(L) + k] * dir[6 * size_t(L) + k];
            acc.sum[kSlope] = gd;
        }
        block_reduce(acc, partials + blockIdx.x);
        return;
    }
    for (unsigned long long mask = __ballot(need); mask; mask &= mask - 1) {
        const int src = __ffsll((long long)mask) - 1;
        const uint32_t E = uint32_t(__shfl(int(L), src, kWave));
        const V3 lo = mk(__shfl(nlo.x, src, kWave), __shfl(nlo.y, src, kWave), __shfl(nlo.z, src, kWave));
        const V3 hi = mk(__shfl(nhi.x, src, kWave), __shfl(nhi.y, src, kWave), __shfl(nhi.z, src, kWave));
        uint32_t *tris = cc.tris + size_t(E) * kCacheCapacity;
        V3 blo = lo, bhi = hi;
        uint32_t n = wave_gather(col, blo, bhi, tris, kCacheCapacity, queue, false);
        if (n > kCacheCapacity) {
            thin_box(cc.skin, blo, bhi);
            n = wave_gather(col, blo, bhi, tris, kCacheCapacity, queue, false);
        }
        if (n == 0) {
            const double far = 15.0 * cc.skin;
            const V3 flo = mk(lo.x - far, lo.y - far, lo.z - far), fhi = mk(hi.x + far, hi.y + far, hi.z + far);
            uint32_t none;
            if (wave_gather(col, flo, fhi, &none, 0, queue, true) == 0) {
                blo = flo;
                bhi = fhi;
            }
        }
        if (lane == 0) {
            double *box = cc.box + 6 * size_t(E);
            store3(box, blo);
            store3(box + 3, bhi);
            cc.count[E] = n <= kCacheCapacity ? n : kCacheOverflow;
            if (n > 0 && !(m.flags[E] & kNodeFixed) && !cc.listed[E]) {
                cc.listed[E] = 1;
                cc.list[atomicAdd(cc.list_count, 1u)] = E;
            }
        }
    }
}

// Minimum of (distance, triangle) in lexicographic order across a group of
// LANES lanes; NaN distances never enter (callers skip them).
template <int LANES = kListLanes>
__device__ __forceinline__ void group_closest(double &d, uint32_t &t)
{
#pragma unroll
    for (int o = LANES / 2; o >= 1; o >>= 1) {
        const double od = __shfl_xor(d, o, kWave);
        const uint32_t ot = __shfl_xor(t, o, kWave);
        if (od < d || (od == d && ot < t)) {
            d = od;
            t = ot;
        }
    }
}

// Wavefront-cooperative walk of the hierarchy for the box [lo, hi] grown by
// pad, with the same overlap tests as bvh_visit (so the same triangles):
// breadth-first through a ring of qcap LDS entries, f(t) called by the lane
// that reaches leaf t. False if the ring would overflow (the walk is then
// incomplete). All lanes call it with the same box; loop control is
// wave-uniform.
template <class F>
__device__ inline bool wave_walk(const ColliderView &col, V3 lo, V3 hi, double pad, volatile int *queue, uint32_t qcap,
                                 F &&f)
{
    const int lane = threadIdx.x & (kWave - 1);
    const unsigned long long below = __lanemask_lt();
    const int ic = internal_count(col);
    uint32_t head = 0, tail = 1;
    if (lane == 0) queue[0] = 0;
    while (head < tail && head <= uint32_t(ic) + 1u) {
        const uint32_t take = (tail - head) < uint32_t(kWave) ? (tail - head) : uint32_t(kWave);
        const int node = uint32_t(lane) < take ? queue[(head + lane) % qcap] : -1;
        head += take;
        bool push[2] = {false, false};
        int child[2] = {-1, -1};
        if (node >= 0) {
            const BvhNode &nd = col.bvh[node];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const V3 bl = mk(nd.box[6 * c], nd.box[6 * c + 1], nd.box[6 * c + 2]);
                const V3 bu = mk(nd.box[6 * c + 3], nd.box[6 * c + 4], nd.box[6 * c + 5]);
                const int ch = nd.child[c];
                child[c] = ch;
                if (!overlap(lo, hi, bl, bu, pad) || ch >= ic) continue;
                if (ch >= 0) {
                    push[c] = true;
                    continue;
                }
                const uint32_t t = uint32_t(-1 - ch);
                if (col.leaf_box) {
                    const double *b = col.leaf_box + kTriRecord * size_t(t);
                    if (!overlap(lo, hi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), pad)) continue;
                }
                f(t);
            }
        }
        const unsigned long long m0 = __ballot(push[0]), m1 = __ballot(push[1]);
        if (tail - head + __popcll(m0) + __popcll(m1) > qcap) return false;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned long long pm = c == 0 ? m0 : m1;
            if (push[c]) queue[(tail + __popcll(pm & below)) % qcap] = child[c];
            tail += __popcll(pm);
        }
        __builtin_amdgcn_wave_barrier();
    }
    return true;
}

// Closest triangle to a segment among those whose box overlaps the
// segment's box grown by pad (the set and tie rule of closest_triangle), by
// a cooperative walk: every lane gets the closest (distance, triangle) and
// the candidate count; false if the walk was incomplete.
__device__ inline bool wave_closest(const ColliderView &col, V3 p0, V3 p1, double pad, volatile int *queue,
                                    uint32_t qcap, double &best_d, uint32_t &best_t, uint32_t &count)
{
    double bd = 1.0e300;
    uint32_t bt = kNoTriangle, n = 0;
    const bool ok = wave_walk(col, vmin(p0, p1), vmax(p0, p1), pad, queue, qcap, [&](uint32_t t) {
        ++n;
        const double d = closest_segment_triangle(p0, p1, triangle_vertex(col, t, 0), triangle_vertex(col, t, 1),
                                                  triangle_vertex(col, t, 2))
                             .distance;
        if (d < bd || (d == bd && t < bt)) {
            bd = d;
            bt = t;
        }
    });
    group_closest<kWave>(bd, bt);
    uint32_t total = n;
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) total += uint32_t(__shfl_xor(int(total), o, kWave));
    best_d = bd;
    best_t = bt;
    count = total;
    return ok;
}

// Narrow phase, a group of kClosestLanes lanes per listed element: the
// group box-tests its list kClosestLanes entries at a time, packs the
// triangles whose leaf box overlaps the padded segment box into LDS, then
// spreads their distance computations over its lanes; the closest (ties to
// the lower index) goes to best_tri. The listed queue holds elements with a
// free left node, the ones the energy queries, and their candidates are
// counted; `extra` (the friction pass) appends the root elements (fixed left
// node, free right node), whose lists only friction reads. An element whose
// list overflowed is deferred: after its other elements the wavefront walks
// the hierarchy for it cooperatively (wave_closest, its groups' packing rows
// as the ring). One the wavefront cannot take (more than kOverflowSlots, or
// a ring overflow) is walked by one lane (closest_triangle). Every element
// of the pass thus gets its best_tri, and the consumers walk nothing.
// With the triangle the pass also leaves the element's closest pair as the
// evaluations use it (kPairWords per element: the distance, the segment
// parameter t and the closest point on the triangle, formed from the
// barycentric weights as the evaluation forms it), so they compute no
// segment-triangle distance of their own.
constexpr int kClosestLanes = 32;
static_assert(kClosestLanes <= int(kCacheCapacity), "a group's first list entries lie in the element's list");
constexpr uint32_t kOverflowSlots = 16;
constexpr int kPairWords = 5;
__device__ __forceinline__ void pair_words(const Closest &c, V3 A, V3 B, V3 C, double (&w)[kPairWords])
{
    const V3 bp = add(add(mul(A, c.bary.x), mul(B, c.bary.y)), mul(C, c.bary.z));
    w[0] = c.distance;
    w[1] = c.t;
    w[2] = bp.x;
    w[3] = bp.y;
    w[4] = bp.z;
}
__device__ __forceinline__ void store_pair(double *best_pair, uint32_t L, const double (&w)[kPairWords])
{
#pragma unroll
    for (int k = 0; k < kPairWords; ++k) best_pair[kPairWords * size_t(L) + k] = w[k];
}
// The closest pair of segment (p0, p1) and triangle t, into best_pair.
__device__ inline void store_pair_of(const ColliderView &col, double *best_pair, uint32_t L, V3 p0, V3 p1, uint32_t t)
{
    const V3 A = triangle_vertex(col, t, 0), B = triangle_vertex(col, t, 1), C = triangle_vertex(col, t, 2);
    double w[kPairWords];
    pair_words(closest_segment_triangle(p0, p1, A, B, C), A, B, C, w);
    store_pair(best_pair, L, w);
}
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_list_closest(const uint32_t *flags, ColliderView col, ContactCache cc,
                                                         StepParams sp, const double *q, const double *q_base,
                                                         const double *dir, double *alpha_at, TrialStep step,
                                                         uint32_t *best_tri, double *best_pair,
                                                         unsigned long long *block_candidates, const uint32_t *extra,
                                                         uint32_t extra_count, const unsigned *stop)
{
    if (stop && *stop) return;
    constexpr int G = kClosestLanes;
    __shared__ uint32_t packed[kBlock / G][kCacheCapacity];
    __shared__ uint32_t deferred[kBlock / kWave][kOverflowSlots];
    __shared__ uint32_t deferred_count[kBlock / kWave];
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) deferred_count[wave] = 0;
    __builtin_amdgcn_wave_barrier();
    // a trial's step (resolved here when no check ran before this pass)
    const double alpha = MODE == kAtTrial ? resolve_step(step, alpha_at) : 0.0;
    const int lane = threadIdx.x & (G - 1);
    const int slot = threadIdx.x / G;
    uint32_t *cand = packed[slot];
    const int shift = (threadIdx.x & (kWave - 1)) & ~(G - 1);  // the group's first lane in the wavefront
    const unsigned long long below = (1ull << lane) - 1;
    const uint32_t group = (blockIdx.x * blockDim.x + threadIdx.x) / G;
    const uint32_t groups = gridDim.x * blockDim.x / G;
    const uint32_t listed = *cc.list_count;
    unsigned long long candidates = 0;
    for (uint32_t j = group; j < listed + extra_count; j += groups) {
        const uint32_t L = j < listed ? cc.list[j] : extra[j - listed];
        // The element's count, flags, end points and first kClosestLanes
        // list entries depend on L alone: requested together, before the
        // count decides what happens to them.
        const uint32_t c = cc.count[L];
        const uint32_t fl = flags[L];
        V3 p0, p1;
        element_points<MODE>(L, q, q_base, dir, alpha, p0, p1);
        const uint32_t *tris = cc.tris + size_t(L) * kCacheCapacity;
        const uint32_t first = tris[lane];
        if (c == 0) continue;
        if (c == kCacheOverflow) {
            if (lane == 0) {
                const uint32_t at = atomicAdd(&deferred_count[wave], 1u);
                if (at < kOverflowSlots) {
                    deferred[wave][at] = L;
                } else {
                    Closest pair;
                    uint32_t tri;
                    unsigned long long n = 0;
                    closest_triangle(col, p0, p1, contact_pad_of(sp, L), pair, tri, n);
                    best_tri[L] = tri;
                    if (tri != kNoTriangle) store_pair_of(col, best_pair, L, p0, p1, tri);
                    if (!(fl & kNodeFixed)) candidates += n;
                }
            }
            continue;
        }
        const bool counted = !(fl & kNodeFixed);
        const double pad = contact_pad_of(sp, L);
        const V3 lo = vmin(p0, p1), hi = vmax(p0, p1);
        uint32_t n = 0;
        for (uint32_t k0 = 0; k0 < c; k0 += G) {
            const uint32_t k = k0 + lane;
            uint32_t t = 0;
            bool hit = false;
            if (k < c) {
                t = k0 == 0 ? first : tris[k];
                const double *b = cc.tri_box + kTriRecord * size_t(t);
                hit = overlap(lo, hi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), pad);
            }
            const unsigned long long mask = (__ballot(hit) >> shift) & ((G == 64) ? ~0ull : ((1ull << G) - 1));
            if (hit) cand[n + __popcll(mask & below)] = t;
            n += __popcll(mask);
        }
        __builtin_amdgcn_wave_barrier();
        double bd = 1.0e300;
        uint32_t bt = kNoTriangle;
        double bw[kPairWords] = {0, 0, 0, 0, 0};  // this lane's closest pair so far
        for (uint32_t i = lane; i < n; i += G) {
            const uint32_t t = cand[i];
            const V3 A = record_vertex(cc.tri_box, t, 0), B = record_vertex(cc.tri_box, t, 1),
                     C = record_vertex(cc.tri_box, t, 2);
            const Closest c = closest_segment_triangle(p0, p1, A, B, C);
            if (c.distance < bd || (c.distance == bd && t < bt)) {
                bd = c.distance;
                bt = t;
                pair_words(c, A, B, C, bw);
            }
        }
        __builtin_amdgcn_wave_barrier();
        const uint32_t own = bt;
        group_closest<G>(bd, bt);
        // the lane that found the closest triangle leaves it and its pair
        if (bt == kNoTriangle ? lane == 0 : own == bt) {
            best_tri[L] = bt;
            if (bt != kNoTriangle) store_pair(best_pair, L, bw);
        }
        if (lane == 0 && counted) candidates += n;
    }
    __builtin_amdgcn_wave_barrier();
    const uint32_t walks = min(deferred_count[wave], kOverflowSlots);
    volatile int *ring = reinterpret_cast<volatile int *>(&packed[wave * (kWave / G)][0]);
    for (uint32_t h = 0; h < walks; ++h) {
        const uint32_t L = deferred[wave][h];
        const double pad = contact_pad_of(sp, L);
        V3 p0, p1;
        element_points<MODE>(L, q, q_base, dir, alpha, p0, p1);
        double bd;
        uint32_t bt, n;
        const bool ok = wave_closest(col, p0, p1, pad, ring, (kWave / G) * kCacheCapacity, bd, bt, n);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            unsigned long long count = n;
            if (!ok) {
                Closest pair;
                count = 0;
                closest_triangle(col, p0, p1, pad, pair, bt, count);
            }
            best_tri[L] = bt;
            if (bt != kNoTriangle) store_pair_of(col, best_pair, L, p0, p1, bt);
            if (!(flags[L] & kNodeFixed)) candidates += count;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (block_candidates) {
        __shared__ unsigned long long wave_candidates[kBlock / kWave];
        candidates = wave_sum_u64(candidates);
        if ((threadIdx.x & (kWave - 1)) == 0) wave_candidates[wave] = candidates;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long total = 0;
            for (int w = 0; w < kBlock / kWave; ++w) total += wave_candidates[w];
            block_candidates[blockIdx.x] = total;
        }
    }
}

// Conservative advancement of one segment along the direction against one
// static triangle (cuda_backend.cu:989-1000): the largest safe step, 1 when
// the triangle does not limit it. Once alpha reaches `stop_at` every later
// return would be at least 0.9 alpha, so the loop returns alpha: callers
// that only take a minimum pass stop_at = (a value some candidate already
// returned) / 0.9 and get the same minimum.
__device__ inline double advancement_limit(V3 p0, V3 p1, V3 d0, V3 d1, double speed, V3 a, V3 b, V3 c, double target,
                                           double stop_at = 2.0)
{
    double alpha = 0.0;
    for (int it = 0; it < 80; ++it) {
        const double dist = closest_segment_triangle(add(p0, mul(d0, alpha)), add(p1, mul(d1, alpha)), a, b, c).distance;
        const double clear = dist - target;
        if (clear <= 0) return fmax(0.0, alpha * 0.9);
        const double advance = 0.8 * clear / speed;
        if (alpha + advance >= 1) return 1.0;
        if (advance < 1e-12) return alpha;
        alpha += advance;
        if (alpha >= stop_at) return alpha;
    }
    return alpha;
}

constexpr int kCcdLanes = 32;
constexpr int kCcdBatch = kBlock / kCcdLanes;
// The swept pass's gathers and the line-search step bound in one launch.
// The first kSweptGatherBlocks blocks take the elements the swept check
// queued, one wavefront each: the wavefront gathers the triangles
// overlapping the element's new box into LDS (an empty result keeps a box
// fifteen skins wider still when that is empty too; a list over capacity
// is gathered again for the thin box), stores the list, and bounds the
// element right away from it (an overflowed list by the wavefront's walk of
// the swept box, overflow_bound). The other blocks bound
// every listed element the check did not queue from its list: kBlock /
// kCcdLanes elements at a time, each group of kCcdLanes lanes box-tests one
// element's list against its swept segment box and appends the overlapping
// (element, triangle) pairs to the block's LDS queue, and every lane of the
// block takes pairs from the queue, so an element with many candidates
// spreads over the block. A queued element's gather thus overlaps the
// others' advancement loops instead of preceding them.
// Only the minimum step matters, so each advancement loop stops once it
// cannot go below the smallest step found so far: the lane's own, or the
// launch-wide one in `bound[slot]` (bits of a non-negative double, lowered
// with atomicMin and read without ordering; any value read is some
// candidate's step or 1, so the minimum is unchanged). Block 0 resets the
// other slot to 1, and the other pending counter to 0, for the next pass.
// The gathers run also when the Newton loop has stopped, so every list
// matches its box. 1,024 gathering blocks (4,096 wavefronts) measured
// against 256-2,048 on the groom: 256 and 512 leave queued elements waiting
// for a wavefront, more add idle ones.
constexpr unsigned kSweptGatherBlocks = 1024;
// Step bound of an element whose list overflowed, by its wavefront: a
// cooperative walk of the swept box, each lane advancing against the leaves
// it reaches (a loop stops once it cannot lower the lane's minimum, which
// leaves the minimum unchanged); an incomplete walk (ring overflow) is
// redone by lane 0 alone. Every lane returns the element's bound.
__device__ inline double overflow_bound(const ColliderView &col, V3 p0, V3 p1, V3 d0, V3 d1, double speed, V3 lo,
                                        V3 hi, double target, volatile int *queue)
{
    double limit = 1.0;
    const bool ok = wave_walk(col, lo, hi, target, queue, kGatherQueue, [&](uint32_t t) {
        limit = fmin(limit, advancement_limit(p0, p1, d0, d1, speed, triangle_vertex(col, t, 0),
                                              triangle_vertex(col, t, 1), triangle_vertex(col, t, 2), target,
                                              limit / 0.9));
    });
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) limit = fmin(limit, __shfl_xor(limit, o, kWave));
    if (!ok) {
        double full = 1.0;
        if ((threadIdx.x & (kWave - 1)) == 0)
            bvh_query(col.bvh, internal_count(col), col.leaf_box, lo, hi, target, [&](uint32_t t) {
                const uint32_t *t3 = col.triangles + 3 * size_t(t);
                full = fmin(full, advancement_limit(p0, p1, d0, d1, speed, vertex(col, t3[0]), vertex(col, t3[1]),
                                                    vertex(col, t3[2]), target));
            });
        limit = __shfl(full, 0, kWave);
    }
    return limit;
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_swept_ccd(const uint32_t *flags, ColliderView col, ContactCache cc,
                                                      StepParams p, const double *q, const double *dir,
                                                      double *block_alpha, unsigned long long *bound, int slot,
                                                      uint32_t *next_pending_count, const unsigned *stop)
{
    const bool stopped = stop && *stop;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        bound[slot ^ 1] = __double_as_longlong(1.0);
        *next_pending_count = 0;
    }
    double limit = 1.0;
    if (blockIdx.x < kSweptGatherBlocks) {
        __shared__ int queues[kTilesPerBlock][kGatherQueue];
        __shared__ uint32_t found_lists[kTilesPerBlock][kCacheCapacity];
        const int wave_in_block = threadIdx.x / kWave;
        volatile int *queue = queues[wave_in_block];
        uint32_t *found = found_lists[wave_in_block];
        uint32_t *hits = reinterpret_cast<uint32_t *>(const_cast<int *>(queue));  // after the walk
        const int lane = threadIdx.x & (kWave - 1);
        const unsigned long long below = __lanemask_lt();
        const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
        const uint32_t waves = kSweptGatherBlocks * kBlock / kWave;
        const uint32_t pending = *cc.pending_count;
        for (uint32_t j = wave; j < pending; j += waves) {
            const uint32_t L = cc.pending[j];
            double *box = cc.box + 6 * size_t(L);
            V3 lo = mk(box[0], box[1], box[2]), hi = mk(box[3], box[4], box[5]);
            uint32_t n = wave_gather(col, lo, hi, found, kCacheCapacity, queue, false);
            if (n > kCacheCapacity) {
                thin_box(cc.skin, lo, hi);
                n = wave_gather(col, lo, hi, found, kCacheCapacity, queue, false);
                if (lane == 0) {
                    store3(box, lo);
                    store3(box + 3, hi);
                }
            }
            if (n == 0) {
                const double far = 15.0 * cc.skin;
                const V3 flo = mk(lo.x - far, lo.y - far, lo.z - far), fhi = mk(hi.x + far, hi.y + far, hi.z + far);
                uint32_t none;
                if (wave_gather(col, flo, fhi, &none, 0, queue, true) == 0 && lane == 0) {
                    store3(box, flo);
                    store3(box + 3, fhi);
                }
            }
            __builtin_amdgcn_wave_barrier();
            const uint32_t kept = n <= kCacheCapacity ? n : 0u;
            uint32_t *tris = cc.tris + size_t(L) * kCacheCapacity;
            for (uint32_t k = lane; k < kept; k += kWave) tris[k] = found[k];
            if (lane == 0) {
                cc.count[L] = n <= kCacheCapacity ? n : kCacheOverflow;
                if (n > 0 && !(flags[L] & kNodeFixed) && !cc.listed[L]) {
                    cc.listed[L] = 1;
                    cc.list[atomicAdd(cc.list_count, 1u)] = L;
                }
            }
            // the element's step bound from its new list (the swept check
            // queues only elements with a free left node)
            if (stopped || n == 0) continue;
            const V3 p0 = load3(q + 6 * size_t(L)), p1 = load3(q + 6 * size_t(L + 1));
            const V3 d0 = load3(dir + 6 * size_t(L)), d1 = load3(dir + 6 * size_t(L + 1));
            const double speed = fmax(norm(d0), norm(d1));
            if (!(speed >= 1e-30)) continue;
            const double target = sweep_target_of(p, L);
            const V3 slo = vmin(vmin(p0, p1), vmin(add(p0, d0), add(p1, d1)));
            const V3 shi = vmax(vmax(p0, p1), vmax(add(p0, d0), add(p1, d1)));
            double own = 1.0;
            if (n > kCacheCapacity) {
                // an overflowed list: the wavefront walks the swept box
                limit = fmin(limit, overflow_bound(col, p0, p1, d0, d1, speed, slo, shi, target, queue));
                __builtin_amdgcn_wave_barrier();
                continue;
            }
            uint32_t taken = 0;
            for (uint32_t k0 = 0; k0 < n; k0 += kWave) {
                const uint32_t k = k0 + lane;
                uint32_t t = 0;
                bool hit = false;
                if (k < n) {
                    t = found[k];
                    const double *b = cc.tri_box + kTriRecord * size_t(t);
                    hit = overlap(slo, shi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), target);
                }
                const unsigned long long mask = __ballot(hit);
                if (hit) hits[taken + __popcll(mask & below)] = t;
                taken += __popcll(mask);
            }
            __builtin_amdgcn_wave_barrier();
            for (uint32_t i = lane; i < taken; i += kWave) {
                const uint32_t t = hits[i];
                const double seen = fmin(own, __longlong_as_double(__hip_atomic_load(
                                                  bound + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                const double r = advancement_limit(p0, p1, d0, d1, speed, record_vertex(cc.tri_box, t, 0),
                                                   record_vertex(cc.tri_box, t, 1), record_vertex(cc.tri_box, t, 2),
                                                   target, seen / 0.9);
                if (r < seen) atomicMin(bound + slot, (unsigned long long)__double_as_longlong(r));
                own = fmin(own, r);
            }
            limit = fmin(limit, own);
            __builtin_amdgcn_wave_barrier();
        }
    } else if (!stopped) {
        // every listed element the check did not queue, from its list
        constexpr int G = kCcdLanes;
        __shared__ uint32_t pair_tri[kCcdBatch * kCacheCapacity];
        __shared__ uint8_t pair_elem[kCcdBatch * kCacheCapacity];
        __shared__ double elem[kCcdBatch][14];  // p0, p1, d0, d1, speed, target clearance
        __shared__ uint32_t pair_count;
        const uint32_t cb = blockIdx.x - kSweptGatherBlocks, cblocks = gridDim.x - kSweptGatherBlocks;
        const int lane = threadIdx.x & (G - 1), my = threadIdx.x / G;
        const int shift = (threadIdx.x & (kWave - 1)) & ~(G - 1);
        const unsigned long long below = (1ull << lane) - 1;
        const uint32_t listed = *cc.listed_snapshot;
        for (uint32_t first = cb * kCcdBatch; first < listed; first += cblocks * kCcdBatch) {
            if (threadIdx.x == 0) pair_count = 0;
            __syncthreads();
            const uint32_t j = first + my;
            const uint32_t L = j < listed ? cc.list[j] : 0u;
            // everything of element L is requested together, before the
            // mark, count and flags decide whether it is bounded here (a
            // group past the end reads element 0, which exists when any is
            // listed, and discards it)
            const uint32_t mark = cc.pending_epoch[L], count = cc.count[L], fl = flags[L];
            const V3 q0 = load3(q + 6 * size_t(L)), q1 = load3(q + 6 * size_t(L + 1));
            const V3 e0 = load3(dir + 6 * size_t(L)), e1 = load3(dir + 6 * size_t(L + 1));
            const uint32_t *tris = cc.tris + size_t(L) * kCacheCapacity;
            const uint32_t first_tri = tris[lane];
            uint32_t c = 0;
            bool overflowed = false;
            if (j < listed && mark != cc.epoch && !(fl & kNodeFixed)) {
                c = count;
                overflowed = c == kCacheOverflow;
                if (overflowed) c = 0;
            }
            V3 p0 = mk(0, 0, 0), p1 = p0, d0 = p0, d1 = p0;
            double speed = 0.0;
            const double target = sweep_target_of(p, L);
            if (c > 0 || overflowed) {
                p0 = q0;
                p1 = q1;
                d0 = e0;
                d1 = e1;
                speed = fmax(norm(d0), norm(d1));
                if (speed < 1e-30) c = 0;
            }
            if (overflowed && lane == 0 && speed >= 1e-30) {
                // an overflowed list (rare): this lane walks the swept box
                const V3 lo = vmin(vmin(p0, p1), vmin(add(p0, d0), add(p1, d1)));
                const V3 hi = vmax(vmax(p0, p1), vmax(add(p0, d0), add(p1, d1)));
                bvh_query(col.bvh, internal_count(col), col.leaf_box, lo, hi, target, [&](uint32_t t) {
                    const uint32_t *t3 = col.triangles + 3 * size_t(t);
                    limit = fmin(limit, advancement_limit(p0, p1, d0, d1, speed, vertex(col, t3[0]),
                                                          vertex(col, t3[1]), vertex(col, t3[2]), target));
                });
            }
            if (c > 0 && lane == 0) {
                const double v[14] = {p0.x, p0.y, p0.z, p1.x, p1.y, p1.z, d0.x, d0.y, d0.z, d1.x, d1.y, d1.z, speed,
                                      target};
#pragma unroll
                for (int k = 0; k < 14; ++k) elem[my][k] = v[k];
            }
            const V3 lo = vmin(vmin(p0, p1), vmin(add(p0, d0), add(p1, d1)));
            const V3 hi = vmax(vmax(p0, p1), vmax(add(p0, d0), add(p1, d1)));
            for (uint32_t k0 = 0; k0 < c; k0 += G) {
                const uint32_t k = k0 + lane;
                uint32_t t = 0;
                bool hit = false;
                if (k < c) {
                    t = k0 == 0 ? first_tri : tris[k];
                    const double *b = cc.tri_box + kTriRecord * size_t(t);
                    hit = overlap(lo, hi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), target);
                }
                const unsigned long long mask = (__ballot(hit) >> shift) & ((G == 64) ? ~0ull : ((1ull << G) - 1));
                uint32_t base = 0;
                if (lane == 0 && mask) base = atomicAdd(&pair_count, uint32_t(__popcll(mask)));
                base = __shfl(base, shift, kWave);
                if (hit) {
                    const uint32_t at = base + __popcll(mask & below);
                    pair_tri[at] = t;
                    pair_elem[at] = uint8_t(my);
                }
            }
            __syncthreads();
            const uint32_t pairs = pair_count;
            for (uint32_t i = threadIdx.x; i < pairs; i += kBlock) {
                const uint32_t t = pair_tri[i];
                const double *e = elem[pair_elem[i]];
                const double seen = fmin(limit, __longlong_as_double(__hip_atomic_load(
                                                    bound + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
                const double r = advancement_limit(mk(e[0], e[1], e[2]), mk(e[3], e[4], e[5]), mk(e[6], e[7], e[8]),
                                                   mk(e[9], e[10], e[11]), e[12], record_vertex(cc.tri_box, t, 0),
                                                   record_vertex(cc.tri_box, t, 1), record_vertex(cc.tri_box, t, 2),
                                                   e[13], seen / 0.9);
                if (r < seen) atomicMin(bound + slot, (unsigned long long)__double_as_longlong(r));
                limit = fmin(limit, r);
            }
            __syncthreads();
        }
    }
    __shared__ double wave_alpha[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) limit = fmin(limit, __shfl_xor(limit, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) wave_alpha[threadIdx.x / kWave] = limit;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = wave_alpha[0];
        for (int w = 1; w < kBlock / kWave; ++w) a = fmin(a, wave_alpha[w]);
        block_alpha[blockIdx.x] = a;
    }
}

// Whether element L's query in this pass was answered by the list pass
// (after one, every query is).
__device__ __forceinline__ bool listed_query(const ContactCache &cc, const uint32_t *best_tri, uint32_t L)
{
    return best_tri != nullptr;
}

// ------------------------------------------------------------ evaluation
// LDS stride of one node's staged 36-entry block (odd in 8-byte words, so
// lanes writing consecutive nodes fall in different banks).
constexpr int kStageStride = 37;

// Writes a tile's staged blocks (nodes first .. first + count - 1, 36
// entries each) as one contiguous run, consecutive lanes on consecutive
// words. Every wavefront of the block calls it.
__device__ inline void store_tile_blocks(const double *st, double *out, long first, int count, int lane)
{
    __syncthreads();
    const int total = 36 * count;
    double *dst = out + 36 * size_t(first);
    for (int k = lane; k < total; k += kWave) {
        const int node = k / 36, e = k - 36 * node;
        dst[k] = st[node * kStageStride + e];
    }
    __syncthreads();
}

// Incremental potential of one substep (reference cuda_backend.cu:883-939):
// inertia and damping per node, Cosserat elasticity and log-barrier contact
// per element; with GRAD also the gradient and the Gauss-Newton blocks (the
// diagonal block D and the upper coupling block U of each node). With LS the
// state is q = alpha d + q_base, written back by the owners.
// Tiles of kEvalTile nodes: wavefront w covers nodes 62 w - 1 .. 62 w + 62,
// one per lane. Every lane forms its node's rotation (and with GRAD its
// right-Jacobian product) once; lanes 0..62 take the next node's from the
// lane above by a DPP shift and evaluate the element from their node to the
// next, lanes 1..62 own their nodes (lane 0's element is the previous tile's
// last, for its coupling terms; lane 63 only lends its node's rotation).
// The assembly is held to 256 registers so two wavefronts share a SIMD
// (a few values spill); at one wavefront per SIMD its loads are exposed.
// With LISTS the list pass before it has chosen every element's closest
// triangle, and the tree walk (and its stack) is not compiled in.
template <bool GRAD, bool LS, bool LISTS, bool MAT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(GRAD ? 2 : 1))) void k_evaluate(
    RodView m, ColliderView col, ContactCache cc, const uint32_t *best_tri, const double *best_pair, StepParams p,
    double *q,
    const double *q_old, const double *vel, const double *q_base, const double *dir, const double *alpha_at,
    double *grad, double *Dblk, double *Ublk, Partial *partials, const unsigned *stop)
{
    if (stop && *stop) return;
    const double alpha = LS ? *alpha_at : 0.0;
    const int lane = threadIdx.x & (kWave - 1);
    const long tile = long(blockIdx.x) * kTilesPerBlock + threadIdx.x / kWave;
    const long L = tile * kEvalTile - 1 + lane;
    const bool in_range = L >= 0 && L < long(m.node_count);
    const bool owner = lane >= 1 && lane <= kEvalTile && in_range;
    // Every input of the lane's node is requested here, before any is used
    // (a lane outside the rods reads node 0 and discards it), so the loads
    // are in flight together instead of one round trip per phase below.
    const size_t Lc = in_range ? size_t(L) : 0;
    const uint32_t fl_loaded = m.flags[Lc];
    double q_in[6], dir_in[6], qo_in[6], v_in[6], F_in[9], rest_in[8];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        q_in[k] = LS ? q_base[6 * Lc + k] : q[6 * Lc + k];
        dir_in[k] = LS ? dir[6 * Lc + k] : 0.0;
        qo_in[k] = q_old[6 * Lc + k];
        v_in[k] = vel[6 * Lc + k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) F_in[k] = m.frame[9 * Lc + k];
    // (the assembly, which holds more registers live here, requests the rest
    // strains below, before its node rotation)
    if constexpr (!GRAD) {
#pragma unroll
        for (int k = 0; k < 8; ++k) rest_in[k] = m.rest[8 * Lc + k];
    }
    const double mass_in = m.mass[Lc], rmass_in = m.rotational_mass[Lc];
    uint32_t count_in = 0, tri_in = kNoTriangle;
    if constexpr (LISTS) {
        count_in = cc.count[Lc];
        tri_in = best_tri[Lc];
    }
    const uint32_t fl = in_range ? fl_loaded : 0u;
    const bool elem = lane <= kEvalTile && (fl & kNodeHasElement) != 0;
    const bool fixed_i = (fl & kNodeFixed) != 0;
    // the next node's flags and state are the lane above's (an element's
    // lane is at most 62, its next node is lane + 1's)
    const uint32_t fl_above = uint32_t(__builtin_amdgcn_update_dpp(0, int(fl), 0x130, 0xf, 0xf, false));
    const bool fixed_j = elem && (fl_above & kNodeFixed) != 0;

    double qi[6] = {0, 0, 0, 0, 0, 0}, qj[6] = {0, 0, 0, 0, 0, 0};
    if (in_range) {
#pragma unroll
        for (int k = 0; k < 6; ++k) qi[k] = LS ? alpha * dir_in[k] + q_in[k] : q_in[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double above = from_lane_above(qi[k]);
        qj[k] = elem ? above : 0.0;
    }
    if (LS && owner) {
#pragma unroll
        for (int k = 0; k < 6; ++k) q[6 * L + k] = qi[k];
    }

    Partial acc = empty_partial();

    // Inertia around q_old + dt v + g dt^2 and damping (cuda_backend.cu:883-897).
    double gnode[6] = {0, 0, 0, 0, 0, 0}, dnode[6] = {0, 0, 0, 0, 0, 0};
    if (owner && !fixed_i) {
        const double kx = mass_in * p.inv_dt2, kr = rmass_in * p.inv_dt2;
        const double cx = mass_in * p.damping * p.inv_dt, cr = rmass_in * p.damping * p.inv_dt;
        double obj = 0.0, kin = 0.0;
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const double gr = d < 3 ? p.gravity[d] * p.dt * p.dt : 0.0;
            const double predicted = qo_in[d] + p.dt * v_in[d] + gr;
            const double delta = qi[d] - predicted;
            const double damp = qi[d] - qo_in[d];
            const double k = d < 3 ? kx : kr, c = d < 3 ? cx : cr;
            obj += 0.5 * k * delta * delta + 0.5 * c * damp * damp;
            if (d < 3) kin += 0.5 * mass_in * damp * damp * p.inv_dt2;
            gnode[d] = k * delta + c * damp;
            dnode[d] = k + c;
        }
        acc.sum[kObjective] += obj;
        acc.sum[kKinetic] += kin;
    }

    // Barrier contact against the closest triangle (cuda_backend.cu:920-938).
    double cg[6] = {0, 0, 0, 0, 0, 0}, dg[6] = {0, 0, 0, 0, 0, 0}, curvature = 0.0, contact_energy = 0.0;
    if (elem && p.has_collider && !fixed_i && (fl & kNodeContact)) {
        const V3 p0 = mk(qi[0], qi[1], qi[2]), p1 = mk(qj[0], qj[1], qj[2]);
        Closest pair;
        V3 listed_bp = mk(0, 0, 0);
        uint32_t tri = 0;
        bool found;
        if constexpr (LISTS) {
            // the list pass counted the candidates, chose the triangle and
            // left its closest pair
            tri = count_in ? tri_in : kNoTriangle;
            found = tri != kNoTriangle;
            if (found) {
                const double *w = best_pair + kPairWords * size_t(L);
                pair.distance = w[0];
                pair.t = w[1];
                listed_bp = mk(w[2], w[3], w[4]);
            }
        } else {
            unsigned long long count = 0;
            found = closest_triangle(col, p0, p1, contact_pad_of(p, L), pair, tri, count);
            // Lane 0 recomputes the previous tile's element for its coupling
            // terms only; statistics are counted by the element's owner lane.
            if (owner) acc.candidates += count;
        }
        if (found) {
            const double radius_offset = radius_offset_of(p, L);
            const double gap = pair.distance - radius_offset;
            if (owner) acc.min_gap = fmin(acc.min_gap, gap);
            if (gap <= p.minimum_gap) {
                if (owner) acc.infeasible = 1;
            } else if (gap < p.barrier) {
                if (owner) acc.active += 1;
                V3 bp = listed_bp;
                if constexpr (!LISTS) {
                    const V3 A = triangle_vertex(col, tri, 0), B = triangle_vertex(col, tri, 1),
                             C = triangle_vertex(col, tri, 2);
                    bp = add(add(mul(A, pair.bary.x), mul(B, pair.bary.y)), mul(C, pair.bary.z));
                }
                const double w0 = 1.0 - pair.t, w1 = pair.t;
                const V3 hp = add(mul(p0, w0), mul(p1, w1));
                const V3 diff = sub(hp, bp);
                const double dist = sqrt(dot(diff, diff) + 1e-30);
                const double dgap = dist - radius_offset;
                if (dgap < p.barrier) {
                    const double aa = p.barrier - dgap, lg = log(dgap * (1.0 / p.barrier));
                    contact_energy = -p.stiffness * aa * aa * lg;
                    if (!isfinite(contact_energy)) {
                        if (owner) acc.infeasible = 1;
                        contact_energy = 0.0;
                    } else if (GRAD) {
                        // b'(g) grad(g), grad(g) = (diff / dist) (w0, w1)
                        const double db = p.stiffness * (2.0 * aa * lg - aa * aa / dgap);
                        const double s = db / dist;
                        cg[0] = s * diff.x * w0; cg[1] = s * diff.y * w0; cg[2] = s * diff.z * w0;
                        cg[3] = s * diff.x * w1; cg[4] = s * diff.y * w1; cg[5] = s * diff.z * w1;
                        // b''(g) grad(g) grad(g)^T with the closest point
                        // held fixed
                        const V3 normal = mul(sub(add(mul(p0, 1.0 - pair.t), mul(p1, pair.t)), bp),
                                              1.0 / fmax(pair.distance, 1e-15));
                        const double a2 = p.barrier - gap;
                        curvature = p.stiffness * (-2.0 * log(gap / p.barrier) + 4.0 * a2 / gap + a2 * a2 / (gap * gap));
                        dg[0] = normal.x * (1.0 - pair.t); dg[1] = normal.y * (1.0 - pair.t); dg[2] = normal.z * (1.0 - pair.t);
                        dg[3] = normal.x * pair.t; dg[4] = normal.y * pair.t; dg[5] = normal.z * pair.t;
                    }
                }
            }
        }
    }

    // Cosserat elasticity (cuda_backend.cu:905-919). This lane's node
    // rotation, and the next node's from the lane above.
    if constexpr (GRAD) {
#pragma unroll
        for (int k = 0; k < 8; ++k) rest_in[k] = m.rest[8 * Lc + k];
    }
    NodeRotation ni = {}, nj;
    if (in_range) node_rotation<GRAD>(qi, F_in, ni);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        nj.R[k] = from_lane_above(ni.R[k]);
        if (GRAD) nj.M[k] = from_lane_above(ni.M[k]);
    }
    ElementJacobian J;
    double elastic = 0.0;
    double weight[6];
    weights_of<MAT>(p, elem ? L : 0, weight);
    if (elem) {
        element_strains_from<GRAD>(qi, qj, ni, nj, rest_in, J);
        const double L0 = rest_in[0];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const double c = (0.5 * L0) * weight[s];
            elastic = elastic + (c * J.strain[s]) * J.strain[s];
        }
    }
    if (owner) {
        acc.sum[kObjective] += elastic;
        acc.sum[kElastic] += elastic;
        acc.sum[kObjective] += contact_energy;
        acc.sum[kContact] += contact_energy;
        if (LS) {
            // |d|^2 of the step this trial takes (the host's increment test)
            double dd = 0.0;
#pragma unroll
            for (int d = 0; d < 6; ++d) dd += dir_in[d] * dir_in[d];
            acc.sum[kDirNorm2] += dd;
        }
    }

    if constexpr (GRAD) {
        const double L0 = elem ? rest_in[0] : 0.0;
        double Lw[6], cw[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            Lw[s] = L0 * weight[s];
            cw[s] = (0.5 * L0) * weight[s];
        }
        // Jacobian entry of strain s by coordinate a of (xi, ui, xj, uj).
        auto jac = [&](int s, int a) -> double {
            if (!elem) return 0.0;
            const int blk = a / 3, c = a % 3;
            if (s < 3) {
                if (blk == 0) return -J.Sx[3 * s + c];
                if (blk == 1) return J.Si[3 * s + c];
                if (blk == 2) return J.Sx[3 * s + c];
                return J.Sj[3 * s + c];
            }
            if (blk == 1) return J.Ki[3 * (s - 3) + c];
            if (blk == 3) return J.Kj[3 * (s - 3) + c];
            return 0.0;
        };
        auto structural_zero = [](int s, int a) { return s >= 3 && (a / 3 == 0 || a / 3 == 2); };
        // Gradient parts, rounded as the dual-number product rule rounds
        // sum_s (c J) e + (c e) J.
        double ge[12];
#pragma unroll
        for (int a = 0; a < 12; ++a) {
            double acc_a = 0.0;
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                if (structural_zero(s, a)) continue;
                const double j = jac(s, a);
                acc_a = acc_a + ((cw[s] * j) * J.strain[s] + (cw[s] * J.strain[s]) * j);
            }
            ge[a] = elem ? acc_a : 0.0;
        }
        // Left element's j-part and contact terms from the lane below.
        double g_out[6];
#pragma unroll
        for (int d = 0; d < 6; ++d) {
            const double from_left = from_lane_below(ge[6 + d]);
            const double ct_left = d < 3 ? from_lane_below(cg[3 + d]) : 0.0;
            g_out[d] = gnode[d] + from_left + ct_left + ge[d] + (d < 3 ? cg[d] : 0.0);
        }
        const double curv_left = from_lane_below(curvature);
        double dg_left[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) dg_left[k] = from_lane_below(dg[3 + k]);
        if (owner) {
            double gg = 0.0;
#pragma unroll
            for (int d = 0; d < 6; ++d) {
                const double gv = fixed_i ? 0.0 : g_out[d];
                grad[6 * L + d] = gv;
                gg += gv * gv;
            }
            acc.sum[kGradNorm2] += gg;
        }
        const bool write_u = owner && elem && !fixed_i && !fixed_j;
        // Each element term L0 w_s J_s[a] J_s[b] is added by a fused
        // multiply-add, as in the oracle. The blocks are staged per
        // wavefront in LDS and written out as the tile's contiguous 36-entry
        // records, whole lines at a time: first D
        // (inertia, both elements' j- and i-parts, both contact terms), then
        // U (this element's coupling and contact term).
        __shared__ double stage[kTilesPerBlock][kEvalTile * kStageStride];
        double *st = stage[threadIdx.x / kWave];
        const int slot = lane - 1;
        const long first_node = tile * kEvalTile;
        const int tile_nodes =
            first_node < long(m.node_count) ? int(min(long(kEvalTile), long(m.node_count) - first_node)) : 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                double hj = 0.0, hi = 0.0;
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    if (!structural_zero(s, 6 + a) && !structural_zero(s, 6 + b))
                        hj = fma(Lw[s] * jac(s, 6 + a), jac(s, 6 + b), hj);
                    if (!structural_zero(s, a) && !structural_zero(s, b)) hi = fma(Lw[s] * jac(s, a), jac(s, b), hi);
                }
                const double left = from_lane_below(hj);
                double v = (a == b ? dnode[a] : 0.0) + left;
                if (a < 3 && b < 3) v = v + (curv_left * dg_left[a]) * dg_left[b];
                v = v + hi;
                if (a < 3 && b < 3) v = v + (curvature * dg[a]) * dg[b];
                if (owner) st[slot * kStageStride + 6 * a + b] = fixed_i ? 0.0 : v;
            }
        }
        store_tile_blocks(st, Dblk, first_node, tile_nodes, lane);
#pragma unroll
        for (int a = 0; a < 6; ++a) {
#pragma unroll
            for (int b = 0; b < 6; ++b) {
                double hu = 0.0;
#pragma unroll
                for (int s = 0; s < 6; ++s)
                    if (!structural_zero(s, a) && !structural_zero(s, 6 + b)) hu = fma(Lw[s] * jac(s, a), jac(s, 6 + b), hu);
                if (a < 3 && b < 3) hu = hu + (curvature * dg[a]) * dg[3 + b];
                if (owner) st[slot * kStageStride + 6 * a + b] = write_u ? hu : 0.0;
            }
        }
        store_tile_blocks(st, Ublk, first_node, tile_nodes, lane);
    }
    block_reduce(acc, partials + blockIdx.x);
}

// ---------------------------------------------------------- linear solve
// Value of `v` in lane L of this lane's 16-lane row (DPP row_newbcast).
// (one 64-bit DPP move: every lane of the row has a valid source lane)
template <int L>
__device__ __forceinline__ double row_bcast(double v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x150 + L, 0xf, 0xf, false);
}
template <int L>
__device__ __forceinline__ float row_bcast(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + L, 0xf, 0xf, false));
}

// f(std::integral_constant<int, 0>), ..., f(std::integral_constant<int, N - 1>).
template <class F, int... Is>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Raw buffer access over one array with a wave-uniform resource. A lane
// whose byte offset is kNoAccess (plus at most 248) reads zero and writes
// nothing (the hardware's range check drops it), so predicated loads and
// stores need no branches and every pass of a loop issues the same memory
// operations, which keeps the compiler's wait counts exact across
// iterations.
constexpr uint32_t kNoAccess = 0xffffff00u;
constexpr uint64_t kMaxBufferBytes = kNoAccess;  // arrays read this way stay below this size

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void *p, uint64_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, int(uint32_t(bytes)), 0x00020000);
}
__device__ __forceinline__ double buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t offset)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, int(offset), 0, 0);
    return __hiloint2double(int(v[1]), int(v[0]));
}
__device__ __forceinline__ void buffer_store(__amdgpu_buffer_rsrc_t r, uint32_t offset, double x)
{
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    u2 v;
    v[0] = unsigned(__double2loint(x));
    v[1] = unsigned(__double2hiint(x));
    __builtin_amdgcn_raw_buffer_store_b64(v, r, int(offset), 0, 0);
}
__device__ __forceinline__ float buffer_load_f32(__amdgpu_buffer_rsrc_t r, uint32_t offset)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, int(offset), 0, 0));
}
__device__ __forceinline__ void buffer_store_f32(__amdgpu_buffer_rsrc_t r, uint32_t offset, float x)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, int(offset), 0, 0);
}
// Loads and stores of the solve's own factors (C', d') at their stored width.
__device__ __forceinline__ void factor_store(__amdgpu_buffer_rsrc_t r, uint32_t offset, double x)
{
    buffer_store(r, offset, x);
}
__device__ __forceinline__ void factor_store(__amdgpu_buffer_rsrc_t r, uint32_t offset, float x)
{
    buffer_store_f32(r, offset, x);
}
template <class T>
__device__ __forceinline__ T factor_load(__amdgpu_buffer_rsrc_t r, uint32_t offset)
{
    if constexpr (sizeof(T) == 8) return buffer_load(r, offset);
    else return buffer_load_f32(r, offset);
}


// Block Thomas per strand on the free nodes; fixed root nodes are
// eliminated and get a zero direction (cuda_backend.cu:866-881).
//
// One 16-lane DPP row per strand, four strands per wavefront (strands
// dispatched longest first, so a wavefront's strands have similar lengths
// and it runs the longest one's node count). In the forward sweep lane
// j < 6 holds column j of the block being inverted and of U, lanes 6..11
// the columns of its inverse, lane 12 the right-hand side. The pivot
// column of each Gauss-Jordan step, the previous node's U (for the Schur
// complement, kept by lanes 0..5) and the finished inverse (for C' = inv U
// and d' = inv rhs) all come by row broadcasts from the lanes that hold
// them, with no LDS round trip. Each entry is
// computed with the sequential algorithm's operations in its order:
// products sum k = 0..5 from zero by fused multiply-adds, as the
// reference's CUDA build contracts them, and the Gauss-Jordan inverse (diagonal
// shift max(1e-12 max|a_ii|, 1e-24), partial pivoting, pivot below 1e-30
// fails; cuda_backend.cu:855-864; each row update one fused multiply-add)
// applies each column step's row operations to all columns at once, so the
// result is bit-identical to a one-thread solve. A first pass without row exchanges stands whenever partial pivoting
// would exchange nothing in any strand of the wavefront (the usual case for
// these diagonally dominant blocks); otherwise the pivoting pass reruns.
// The back substitution runs row-wise on lanes 0..5; it also gathers the
// direction's statistics for the choice that follows (cuda_backend.cu:
// 623-630): g.d, each row lane summing its own products over the strand's
// nodes, and the largest free-node translation |d_xyz| (its square kept,
// one square root at the end), one Partial per block (a skipped launch
// writes a failed one, so nothing is chosen from it).
//
// Memory: a ring of kSolveAhead (forward) or kBackAhead (back substitution)
// nodes' inputs; a slot is refilled with the node that many steps later as
// soon as its step is done. Loads and stores are raw buffer operations
// predicated by offset, so no step waits for the stores of the one before.
//
// T is the arithmetic type: double for the FP64 step; float for the mixed
// step, where the FP64 blocks and gradient are rounded to FP32 as they are
// loaded, every Schur complement, inverse and factor is formed and stored in
// FP32 (C' and d' at half the bytes), and the direction is widened to FP64
// as it is stored.
constexpr int kSolveLanes = 16;
constexpr int kSolveGroups = kWave / kSolveLanes;
constexpr int kSolveAhead = 3;
constexpr int kBackAhead = 8;
constexpr uint64_t kMaxSolveNodes = kMaxBufferBytes / 288;  // 36 doubles per node block

template <class T>
struct ForwardSlot {
    T d[6];  // D column
    T g[6];  // gradient (rhs lane)
    T u[6];  // U column
};

template <class T>
__global__ __launch_bounds__(kWave) void k_block_solve(RodView m, const uint32_t *order, const double *Dblk,
                                                       const double *Ublk, const double *grad, double *Cp,
                                                       double *dp, double *dir, unsigned *failed,
                                                       unsigned *next_failed, Partial *partials, const unsigned *stop)
{
    constexpr uint32_t E = sizeof(T), EB = 36 * E, EV = 6 * E;  // factor entry, block and vector bytes
    // the next solve's flag starts at zero (the two alternate)
    if (blockIdx.x == 0 && threadIdx.x == 0 && next_failed) *next_failed = 0;
    if (stop && *stop) {
        Partial skipped = empty_partial();
        skipped.failed = 1;
        block_reduce(skipped, partials + blockIdx.x);
        return;
    }
    const int g = threadIdx.x / kSolveLanes, j = threadIdx.x % kSolveLanes;
    const bool acol = j < 6, icol = j >= 6 && j < 12, vlane = j == 12;
    const uint32_t gi = blockIdx.x * kSolveGroups + g;
    uint32_t begin = 0, end = 0;
    if (gi < m.strand_count) {
        const uint32_t s = order[gi];
        const uint32_t first = m.strand_first[s];
        begin = first + m.strand_fixed[s];
        end = first + m.strand_count_of[s];
        if (acol)
            for (uint32_t n = first; n < begin; ++n) dir[6 * size_t(n) + j] = 0.0;
        if (begin > end) begin = end;
    }
    const uint32_t len = end - begin;
    uint32_t steps = len;
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) steps = max(steps, uint32_t(__shfl_xor(int(steps), o, kWave)));
    steps = __builtin_amdgcn_readfirstlane(steps);  // equal in every lane: keep the loops scalar

    const uint64_t block_bytes = 288 * uint64_t(m.node_count), vec_bytes = 48 * uint64_t(m.node_count);
    const __amdgpu_buffer_rsrc_t rD = buffer_of(Dblk, block_bytes), rU = buffer_of(Ublk, block_bytes);
    const __amdgpu_buffer_rsrc_t rG = buffer_of(grad, vec_bytes), rC = buffer_of(Cp, block_bytes);
    const __amdgpu_buffer_rsrc_t rP = buffer_of(dp, vec_bytes), rX = buffer_of(dir, vec_bytes);

    // node t's inputs: D column j (acol), gradient (rhs lane), U column j
    // (acol, when a node follows)
    auto fetch = [&](ForwardSlot<T> &f, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = begin + t;
        const uint32_t od = (acol && t < len) ? 288 * n + 8 * j : kNoAccess;
        const uint32_t og = (vlane && t < len) ? 48 * n : kNoAccess;
        const uint32_t ou = (acol && t + 1 < len) ? 288 * n + 8 * j : kNoAccess;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            f.d[r] = T(buffer_load(rD, od + 48 * r));
            f.g[r] = T(buffer_load(rG, og + 8 * r));
            f.u[r] = T(buffer_load(rU, ou + 48 * r));
        }
    };

    bool ok = true;
    T p[6];  // previous C' column (d' on the rhs lane)
#pragma unroll
    for (int r = 0; r < 6; ++r) p[r] = T(0);
    ForwardSlot<T> ring[kSolveAhead];
#pragma unroll
    for (int a = 0; a < kSolveAhead; ++a) fetch(ring[a], a);
    T wprev[6];  // the previous node's U column (lanes 0..5); zero before the first
#pragma unroll
    for (int k = 0; k < 6; ++k) wprev[k] = T(0);

    auto forward = [&](ForwardSlot<T> &slot, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = begin + t;
        T col[6], w[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            col[r] = acol ? slot.d[r] : (vlane ? -slot.g[r] : ((icol && r == j - 6) ? T(1) : T(0)));
            w[r] = slot.u[r];
        }
        // col[r] -= sum_k U'[k][r] p[k]: D - U'^T C' on the block columns,
        // rhs - U'^T d' on the rhs lane, U'[k][r] from lane r. At the first
        // node U' and p are zero and this subtracts +0, which leaves every
        // value (signed zeros, infinities and NaNs included) unchanged.
        static_for<6>([&](auto R) {
            T sum = T(0);
#pragma unroll
            for (int k = 0; k < 6; ++k) sum = fma(row_bcast<R>(wprev[k]), p[k], sum);
            col[R] -= sum;
        });
#pragma unroll
        for (int k = 0; k < 6; ++k) wprev[k] = w[k];
        // The rhs lane keeps its column (the right-hand side) in rhs; the
        // Gauss-Jordan steps below also transform its col, which nothing
        // reads (lane 12 is never a pivot column).
        T rhs[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) rhs[r] = col[r];
        // Gauss-Jordan inverse on the columns of [a | inv].
        T scale = T(0);
        static_for<6>([&](auto I) { scale = fmax(scale, fabs(row_bcast<I>(col[I]))); });
        const T shift = fmax(T(1.0e-12) * scale, T(1.0e-24));
#pragma unroll
        for (int r = 0; r < 6; ++r)
            if (acol && r == j) col[r] += shift;
        // Fast pass without row exchanges. Where no column step finds an
        // entry below the diagonal larger in magnitude than the diagonal one,
        // partial pivoting would exchange nothing and this pass performs
        // exactly its operations; otherwise (some strand of the wavefront)
        // the pivoting pass reruns from the saved columns.
        T start[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) start[r] = col[r];
        bool exchange = false, singular = false;
        static_for<6>([&](auto C) {
            T cv[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) cv[i] = row_bcast<C>(col[i]);
            const T diag = fabs(cv[C]);
#pragma unroll
            for (int i = C + 1; i < 6; ++i) exchange |= fabs(cv[i]) > diag;
            singular |= diag < T(1.0e-30) || !isfinite(diag);
            const T invp = T(1) / cv[C];
            col[C] *= invp;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r != C) col[r] = fma(-cv[r], col[C], col[r]);
        });
        if (__builtin_amdgcn_ballot_w64(exchange)) {
            singular = false;
#pragma unroll
            for (int r = 0; r < 6; ++r) col[r] = start[r];
            static_for<6>([&](auto C) {
                T cv[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) cv[i] = row_bcast<C>(col[i]);
                int pivot = C;
                T best = fabs(cv[C]);
#pragma unroll
                for (int i = C + 1; i < 6; ++i) {
                    const T v = fabs(cv[i]);
                    if (v > best) { best = v; pivot = i; }
                }
                if (best < T(1.0e-30) || !isfinite(best)) {
                    singular = true;
                    pivot = C;
                }
#pragma unroll
                for (int i = C + 1; i < 6; ++i) {
                    const bool sw = pivot == i;
                    const T x = col[C], y = col[i];
                    col[C] = sw ? y : x;
                    col[i] = sw ? x : y;
                    const T cx = cv[C], cy = cv[i];
                    cv[C] = sw ? cy : cx;
                    cv[i] = sw ? cx : cy;
                }
                const T invp = T(1) / cv[C];
                col[C] *= invp;
#pragma unroll
                for (int r = 0; r < 6; ++r)
                    if (r != C) col[r] = fma(-cv[r], col[C], col[r]);
            });
        }
        if (singular && t < len) ok = false;
        // res[r] = sum_k inv[r][k] w[k]: C' = inv U on the block columns,
        // d' = inv rhs on the rhs lane; inverse column k is on lane 6 + k
#pragma unroll
        for (int r = 0; r < 6; ++r)
            if (vlane) w[r] = rhs[r];
        T res[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) res[r] = T(0);
        static_for<6>([&](auto K) {
#pragma unroll
            for (int r = 0; r < 6; ++r) res[r] = fma(row_bcast<6 + K>(col[r]), w[K], res[r]);
        });
        const bool carry = (acol && t + 1 < len) || vlane;  // C' columns while a node follows, and d'
#pragma unroll
        for (int r = 0; r < 6; ++r) p[r] = carry ? res[r] : T(0);
        const bool keep = t < len && ok;
        const uint32_t oc = (acol && keep) ? EB * n + E * j : kNoAccess;
        const uint32_t op = (vlane && keep) ? EV * n : kNoAccess;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            factor_store(rC, oc + EV * r, p[r]);
            factor_store(rP, op + E * r, p[r]);
        }
        // refilled once its values are dead, so it keeps its registers
        fetch(slot, t + kSolveAhead);
    };
    for (uint32_t t0 = 0; t0 < steps; t0 += kSolveAhead)
        static_for<kSolveAhead>([&](auto A) { forward(ring[A], t0 + A); });
    if (!ok && j == 0) atomicOr(failed, 1u);

    // Back substitution x_n = d'_n - C'_n x_{n+1}, lane r < 6 owning row r.
    const int r = acol ? j : 0;
    struct BackSlot {
        T c[6], d;
        double g;
    };
    auto fetch_back = [&](BackSlot &b, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = end - 1 - t;
        const bool live = acol && t < len;
        const uint32_t oc = live ? EB * n + EV * r : kNoAccess;
#pragma unroll
        for (int k = 0; k < 6; ++k) b.c[k] = factor_load<T>(rC, oc + E * k);
        b.d = factor_load<T>(rP, live ? EV * n + E * r : kNoAccess);
        b.g = buffer_load(rG, live ? 48 * n + 8 * r : kNoAccess);
    };
    BackSlot back[kBackAhead];
#pragma unroll
    for (int a = 0; a < kBackAhead; ++a) fetch_back(back[a], a);
    T next = T(0);
    double slope = 0.0, reach2 = 0.0;
    auto backward = [&](BackSlot &b, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = end - 1 - t;
        T crow[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) crow[k] = b.c[k];
        T value_t = b.d;
        if (t > 0) static_for<6>([&](auto K) { value_t = fma(-crow[K], row_bcast<K>(next), value_t); });
        next = value_t;
        const double value = double(value_t);
        buffer_store(rX, (t < len && ok && acol) ? 48 * n + 8 * j : kNoAccess, value);
        // this row's part of g.d; |d_xyz|^2 of node n on lane 0
        const double sq = value * value;
        const double t2 = (sq + row_bcast<1>(sq)) + row_bcast<2>(sq);
        if (t < len) {
            slope += b.g * value;
            reach2 = fmax(reach2, t2);
        }
        fetch_back(b, t + kBackAhead);
    };
    for (uint32_t t0 = 0; t0 < steps; t0 += kBackAhead)
        static_for<kBackAhead>([&](auto A) { backward(back[A], t0 + A); });
    double row_slope = 0.0;
    static_for<6>([&](auto K) { row_slope += row_bcast<K>(acol ? slope : 0.0); });
    Partial acc = empty_partial();
    if (j == 0 && gi < m.strand_count) {
        acc.sum[kSlope] = row_slope;
        acc.max_translation = sqrt(reach2);
        acc.failed = ok ? 0u : 1u;
    }
    block_reduce(acc, partials + blockIdx.x);
}

// The same solve from both ends (twisted block factorization; the oracle's
// block_solve_twisted restates it operation for operation). Two strands
// per wavefront, two 16-lane rows each: the first row runs block Thomas
// from the first free node up to the middle node m (free count - 1) / 2
// further on, the second from the tip down to m on the reversed chain,
// whose coupling blocks are the transposes (its U column j is row j of the
// original U of the node before). The first row then forms the middle
// node's Schur complement from both sides (the second row's last factor
// and coupling block come by lane shuffles), solves it, and both rows
// back-substitute outward from m. Lanes, registers, the Gauss-Jordan
// passes and the buffer accesses are those of k_block_solve; a row whose
// side is shorter than the wavefront's longest keeps its last factor
// (its later steps are masked). The serial chain is about half the
// strand's free length.
constexpr int kTwistGroups = kWave / (2 * kSolveLanes);

__global__ __launch_bounds__(kWave) void k_block_solve_twisted(RodView m, const uint32_t *order, const double *Dblk,
                                                               const double *Ublk, const double *grad, double *Cp,
                                                               double *dp, double *dir, unsigned *failed,
                                                               unsigned *next_failed, Partial *partials,
                                                               const unsigned *stop)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && next_failed) *next_failed = 0;
    if (stop && *stop) {
        Partial skipped = empty_partial();
        skipped.failed = 1;
        block_reduce(skipped, partials + blockIdx.x);
        return;
    }
    const int row = threadIdx.x / kSolveLanes, j = threadIdx.x % kSolveLanes;
    const int side = row & 1;  // 0: from the root, 1: from the tip
    const bool acol = j < 6, icol = j >= 6 && j < 12, vlane = j == 12;
    const uint32_t gi = blockIdx.x * kTwistGroups + (row >> 1);
    uint32_t begin = 0, end = 0;
    if (gi < m.strand_count) {
        const uint32_t s = order[gi];
        const uint32_t first = m.strand_first[s];
        begin = first + m.strand_fixed[s];
        end = first + m.strand_count_of[s];
        if (acol && side == 0)
            for (uint32_t n = first; n < begin; ++n) dir[6 * size_t(n) + j] = 0.0;
        if (begin > end) begin = end;
    }
    const uint32_t len = end - begin;
    const bool has_free = len > 0;
    const uint32_t na = has_free ? (len - 1) / 2 : 0, nb = has_free ? len - 1 - na : 0;
    const uint32_t mid = begin + na;
    const uint32_t lside = side ? nb : na;
    uint32_t steps = lside;
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) steps = max(steps, uint32_t(__shfl_xor(int(steps), o, kWave)));
    steps = __builtin_amdgcn_readfirstlane(steps);

    const uint64_t block_bytes = 288 * uint64_t(m.node_count), vec_bytes = 48 * uint64_t(m.node_count);
    const __amdgpu_buffer_rsrc_t rD = buffer_of(Dblk, block_bytes), rU = buffer_of(Ublk, block_bytes);
    const __amdgpu_buffer_rsrc_t rG = buffer_of(grad, vec_bytes), rC = buffer_of(Cp, block_bytes);
    const __amdgpu_buffer_rsrc_t rP = buffer_of(dp, vec_bytes), rX = buffer_of(dir, vec_bytes);
    // the root side's U column j (stride 48 between rows) or, on the tip
    // side, row j of the node before's U (stride 8 between columns)
    const uint32_t ustride = side ? 8 : 48;
    auto node_at = [&](uint32_t t) { return side ? end - 1 - t : begin + t; };
    auto fetch = [&](ForwardSlot<double> &f, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = node_at(t);
        const bool live = t < lside;
        const uint32_t od = (acol && live) ? 288 * n + 8 * j : kNoAccess;
        const uint32_t og = (vlane && live) ? 48 * n : kNoAccess;
        const uint32_t ou = (acol && live) ? (side ? 288 * (n - 1) + 48 * j : 288 * n + 8 * j) : kNoAccess;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            f.d[r] = buffer_load(rD, od + 48 * r);
            f.g[r] = buffer_load(rG, og + 8 * r);
            f.u[r] = buffer_load(rU, ou + ustride * r);
        }
    };

    bool ok = true;
    double p[6], wprev[6];  // last factor column (d' on the rhs lane) and coupling column
#pragma unroll
    for (int r = 0; r < 6; ++r) p[r] = wprev[r] = 0.0;

    // col[R] -= sum_k W[k][R] P[k], W[k][R] from lane R
    auto schur = [&](double col[6], const double wv[6], const double pv[6]) __attribute__((always_inline)) {
        static_for<6>([&](auto R) {
            double sum = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) sum = fma(row_bcast<R>(wv[k]), pv[k], sum);
            col[R] -= sum;
        });
    };
    // Gauss-Jordan inverse on the columns of [a | inv]; returns singular
    auto invert = [&](double col[6], bool live) __attribute__((always_inline)) {
        double scale = 0.0;
        static_for<6>([&](auto I) { scale = fmax(scale, fabs(row_bcast<I>(col[I]))); });
        const double shift = fmax(1.0e-12 * scale, 1.0e-24);
#pragma unroll
        for (int r = 0; r < 6; ++r)
            if (acol && r == j) col[r] += shift;
        double start[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) start[r] = col[r];
        bool exchange = false, singular = false;
        static_for<6>([&](auto C) {
            double cv[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) cv[i] = row_bcast<C>(col[i]);
            const double diag = fabs(cv[C]);
#pragma unroll
            for (int i = C + 1; i < 6; ++i) exchange |= live && fabs(cv[i]) > diag;
            singular |= diag < 1.0e-30 || !isfinite(diag);
            const double invp = 1.0 / cv[C];
            col[C] *= invp;
#pragma unroll
            for (int r = 0; r < 6; ++r)
                if (r != C) col[r] = fma(-cv[r], col[C], col[r]);
        });
        if (__builtin_amdgcn_ballot_w64(exchange)) {
            singular = false;
#pragma unroll
            for (int r = 0; r < 6; ++r) col[r] = start[r];
            static_for<6>([&](auto C) {
                double cv[6];
#pragma unroll
                for (int i = 0; i < 6; ++i) cv[i] = row_bcast<C>(col[i]);
                int pivot = C;
                double best = fabs(cv[C]);
#pragma unroll
                for (int i = C + 1; i < 6; ++i) {
                    const double v = fabs(cv[i]);
                    if (v > best) { best = v; pivot = i; }
                }
                if (best < 1.0e-30 || !isfinite(best)) {
                    singular = true;
                    pivot = C;
                }
#pragma unroll
                for (int i = C + 1; i < 6; ++i) {
                    const bool sw = pivot == i;
                    const double x = col[C], y = col[i];
                    col[C] = sw ? y : x;
                    col[i] = sw ? x : y;
                    const double cx = cv[C], cy = cv[i];
                    cv[C] = sw ? cy : cx;
                    cv[i] = sw ? cx : cy;
                }
                const double invp = 1.0 / cv[C];
                col[C] *= invp;
#pragma unroll
                for (int r = 0; r < 6; ++r)
                    if (r != C) col[r] = fma(-cv[r], col[C], col[r]);
            });
        }
        return singular;
    };
    // res[r] = sum_K inv[r][K] w[K], inverse column K on lane 6 + K
    auto apply_inverse = [&](const double col[6], const double w[6], double res[6]) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 6; ++r) res[r] = 0.0;
        static_for<6>([&](auto K) {
#pragma unroll
            for (int r = 0; r < 6; ++r) res[r] = fma(row_bcast<6 + K>(col[r]), w[K], res[r]);
        });
    };

    ForwardSlot<double> ring[kSolveAhead];
#pragma unroll
    for (int a = 0; a < kSolveAhead; ++a) fetch(ring[a], a);
    auto forward = [&](ForwardSlot<double> &slot, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = node_at(t);
        const bool live = t < lside;
        double col[6], w[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            col[r] = acol ? slot.d[r] : (vlane ? -slot.g[r] : ((icol && r == j - 6) ? 1.0 : 0.0));
            w[r] = slot.u[r];
        }
        schur(col, wprev, p);
#pragma unroll
        for (int k = 0; k < 6; ++k) wprev[k] = live ? w[k] : wprev[k];
        double rhs[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) rhs[r] = col[r];
        const bool singular = invert(col, live);
        if (singular && live) ok = false;
#pragma unroll
        for (int r = 0; r < 6; ++r)
            if (vlane) w[r] = rhs[r];
        double res[6];
        apply_inverse(col, w, res);
        // every side node has a follower (the middle node at last)
#pragma unroll
        for (int r = 0; r < 6; ++r) p[r] = live ? ((acol || vlane) ? res[r] : 0.0) : p[r];
        const bool keep = live && ok;
        const uint32_t oc = (acol && keep) ? 288 * n + 8 * j : kNoAccess;
        const uint32_t op = (vlane && keep) ? 48 * n : kNoAccess;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            buffer_store(rC, oc + 48 * r, p[r]);
            buffer_store(rP, op + 8 * r, p[r]);
        }
        fetch(slot, t + kSolveAhead);
    };
    for (uint32_t t0 = 0; t0 < steps; t0 += kSolveAhead)
        static_for<kSolveAhead>([&](auto A) { forward(ring[A], t0 + A); });

    // The middle node, on the root-side row: its Schur complement takes the
    // root side's term, then the tip side's (U_m times the tip side's last
    // factor: the tip row's coupling column j is row j of U_m).
    const int base = threadIdx.x & ~(2 * kSolveLanes - 1);  // the strand's first lane
    double pt[6], wt[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        pt[r] = __shfl(p[r], base + kSolveLanes + j, kWave);
        wt[r] = __shfl(wprev[r], base + kSolveLanes + j, kWave);
    }
    const bool mlive = side == 0 && has_free;
    double xm[6];
    {
        const uint32_t od = (acol && mlive) ? 288 * mid + 8 * j : kNoAccess;
        const uint32_t og = (vlane && mlive) ? 48 * mid : kNoAccess;
        double col[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const double dv = buffer_load(rD, od + 48 * r), gv = buffer_load(rG, og + 8 * r);
            col[r] = acol ? dv : (vlane ? -gv : ((icol && r == j - 6) ? 1.0 : 0.0));
        }
        schur(col, wprev, p);
        schur(col, wt, pt);
        double rhs[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) rhs[r] = col[r];
        const bool singular = invert(col, mlive);
        if (singular && mlive) ok = false;
        apply_inverse(col, rhs, xm);  // x_m on the rhs lane
    }
    // both rows of the strand learn the middle solution and the strand's
    // verdict
    const bool strand_ok = __shfl(int(ok), base, kWave) && __shfl(int(ok), base + kSolveLanes, kWave);
    const int r = acol ? j : 0;
    double next = 0.0;
    static_for<6>([&](auto Q) {
        const double v = __shfl(xm[Q], base + 12, kWave);
        if (r == Q) next = v;
    });
    double slope = 0.0, reach2 = 0.0;
    if (mlive && acol) {
        buffer_store(rX, strand_ok ? 48 * mid + 8 * j : kNoAccess, next);
        slope += buffer_load(rG, 48 * mid + 8 * r) * next;
    }
    {
        const double sq = next * next;
        const double t2 = (sq + row_bcast<1>(sq)) + row_bcast<2>(sq);
        if (mlive) reach2 = fmax(reach2, t2);
    }

    // Back substitution outward from m: x_n = d'_n - C'_n x_(n -+ 1), lane
    // r < 6 owning row r.
    struct BackSlot {
        double c[6], d, g;
    };
    auto back_node = [&](uint32_t t) { return side ? mid + 1 + t : mid - 1 - t; };
    auto fetch_back = [&](BackSlot &b, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = back_node(t);
        const bool live = acol && t < lside;
        const uint32_t oc = live ? 288 * n + 48 * r : kNoAccess;
#pragma unroll
        for (int k = 0; k < 6; ++k) b.c[k] = buffer_load(rC, oc + 8 * k);
        b.d = buffer_load(rP, live ? 48 * n + 8 * r : kNoAccess);
        b.g = buffer_load(rG, live ? 48 * n + 8 * r : kNoAccess);
    };
    BackSlot back[kBackAhead];
#pragma unroll
    for (int a = 0; a < kBackAhead; ++a) fetch_back(back[a], a);
    auto backward = [&](BackSlot &b, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = back_node(t);
        double crow[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) crow[k] = b.c[k];
        double value = b.d;
        static_for<6>([&](auto K) { value = fma(-crow[K], row_bcast<K>(next), value); });
        const bool live = t < lside;
        buffer_store(rX, (live && strand_ok && acol) ? 48 * n + 8 * j : kNoAccess, value);
        next = live ? value : next;
        const double sq = value * value;
        const double t2 = (sq + row_bcast<1>(sq)) + row_bcast<2>(sq);
        if (live) {
            slope += b.g * value;
            reach2 = fmax(reach2, t2);
        }
        fetch_back(b, t + kBackAhead);
    };
    for (uint32_t t0 = 0; t0 < steps; t0 += kBackAhead)
        static_for<kBackAhead>([&](auto A) { backward(back[A], t0 + A); });
    if (!strand_ok && j == 0 && side == 0 && has_free) atomicOr(failed, 1u);
    double row_slope = 0.0;
    static_for<6>([&](auto K) { row_slope += row_bcast<K>(acol ? slope : 0.0); });
    Partial acc = empty_partial();
    if (j == 0 && gi < m.strand_count) {
        acc.sum[kSlope] = row_slope;
        acc.max_translation = sqrt(reach2);
        acc.failed = strand_ok ? 0u : 1u;
    }
    block_reduce(acc, partials + blockIdx.x);
}

// Block Thomas with the factors taken straight out of one elimination
// (solve algorithm 2; the oracle's block_solve_fused restates it). Per free
// node n the 16-lane row holds the columns of [S | U_n | e]: S = D_n -
// U_{n-1}^T C'_{n-1}, e = g_n - U_{n-1}^T e_{n-1} with e = -d' (the rhs
// carried with the opposite sign, which rounds identically); Gauss-Jordan
// without pivoting reduces S to I and leaves C'_n = S^-1 U_n and e_n in the
// other columns, with no inverse formed and no product by it. The blocks
// are symmetric positive definite (Gauss-Newton Hessian, mass, the
// reference's diagonal shift), so their Schur complements need no row
// exchanges; a pivot below 1e-30 or a non-finite result fails the solve as
// a singular pivot fails the reference's (cuda_backend.cu:855-881).
//
// The S and U column halves alternate between lanes 0..5 and 6..11 from
// node to node: the U lanes of node n, whose columns become C'_n, are the
// S lanes of node n + 1, so the Schur complement reads its C' column in
// place and U_n (kept in the lane's ring slot) by row broadcast. Lane 12
// holds e; every lane reads its one column per node (6 loads of D, U or g
// out of one buffer holding all three). Each Schur and row-operation term
// is one v_fmac_f64_dpp: the product's first factor comes from the
// broadcasting lane inside the fused multiply-add (the compiler forms
// these only as a separate DPP move and FMA). Per node that is 36 DPP-FMAs
// for the Schur complement and 6 x (1 reciprocal, 5 DPP-FMAs) for the
// elimination, against the reference form's inverse, its products and
// its moves: about a third of the instructions.
constexpr int kFusedAhead = 6;  // ring slots; even, as the lane halves alternate (4: 2.5 us slower, 8: 1 us)
// Nodes per strand whose factors stay in LDS: 28 is the largest window that
// keeps four of these one-wavefront workgroups on a CU (38.9 KB each);
// measured on the 6,757-strand groom against 12-40, the solve takes 88 us
// (20: 93 us, 32: 104 us)
constexpr int kFusedWindow = 28;
constexpr int kFusedRecord = 42;  // a node's C' (36, row-major) and e (6)

// v_fmac_f64_dpp ACC, X, -Y: ACC += X(lane L of the row) * -Y
#define KHS_FMAC_NEG(ACC, X, Y, L) "v_fmac_f64_dpp " ACC ", " X ", -" Y " row_newbcast:" L " row_mask:0xf bank_mask:0xf\n\t"
// col[R] += U[k][R] * -p[k] for R = 0..5; U[k][R] is w[k] of lane B + R
#define KHS_SCHUR_K(W, P, L0, L1, L2, L3, L4, L5)                                                                  \
    KHS_FMAC_NEG("%0", W, P, L0) KHS_FMAC_NEG("%1", W, P, L1) KHS_FMAC_NEG("%2", W, P, L2)                         \
        KHS_FMAC_NEG("%3", W, P, L3) KHS_FMAC_NEG("%4", W, P, L4) KHS_FMAC_NEG("%5", W, P, L5)
// k outermost: consecutive DPP-FMAs never read what the one before wrote;
// the leading s_nop covers a VALU write of an input just before the block
#define KHS_SCHUR(L0, L1, L2, L3, L4, L5)                                                                          \
    "s_nop 1\n\t" KHS_SCHUR_K("%6", "%12", L0, L1, L2, L3, L4, L5) KHS_SCHUR_K("%7", "%13", L0, L1, L2, L3, L4, L5) \
        KHS_SCHUR_K("%8", "%14", L0, L1, L2, L3, L4, L5) KHS_SCHUR_K("%9", "%15", L0, L1, L2, L3, L4, L5)           \
            KHS_SCHUR_K("%10", "%16", L0, L1, L2, L3, L4, L5) KHS_SCHUR_K("%11", "%17", L0, L1, L2, L3, L4, L5)
// pivot row C scaled by %6 (1 / pivot), then row r -= S[r][C] * row C for
// the other rows, S[r][C] from the pivot column's lane L
#define KHS_GJ(C, L, A1, A2, A3, A4, A5)                                                                           \
    "v_mul_f64 %" #C ", %" #C ", %6\n\ts_nop 1\n\t" KHS_FMAC_NEG("%" #A1, "%" #A1, "%" #C, #L)                    \
        KHS_FMAC_NEG("%" #A2, "%" #A2, "%" #C, #L) KHS_FMAC_NEG("%" #A3, "%" #A3, "%" #C, #L)                      \
            KHS_FMAC_NEG("%" #A4, "%" #A4, "%" #C, #L) KHS_FMAC_NEG("%" #A5, "%" #A5, "%" #C, #L)
#define KHS_GJ_ASM(C, L, A1, A2, A3, A4, A5)                                                                       \
    asm(KHS_GJ(C, L, A1, A2, A3, A4, A5)                                                                         \
        : "+v"(col[0]), "+v"(col[1]), "+v"(col[2]), "+v"(col[3]), "+v"(col[4]), "+v"(col[5])                    \
        : "v"(inv))

template <int B>  // lanes B..B+5 hold the S columns and the previous node's U columns
__device__ __forceinline__ void fused_schur(double (&col)[6], const double (&w)[6], const double (&p)[6])
{
#define KHS_SCHUR_OPERANDS                                                                                         \
    : "+v"(col[0]), "+v"(col[1]), "+v"(col[2]), "+v"(col[3]), "+v"(col[4]), "+v"(col[5])                        \
    : "v"(w[0]), "v"(w[1]), "v"(w[2]), "v"(w[3]), "v"(w[4]), "v"(w[5]), "v"(p[0]), "v"(p[1]), "v"(p[2]),         \
      "v"(p[3]), "v"(p[4]), "v"(p[5])
    if constexpr (B == 0) asm(KHS_SCHUR("0", "1", "2", "3", "4", "5") KHS_SCHUR_OPERANDS);
    else asm(KHS_SCHUR("6", "7", "8", "9", "10", "11") KHS_SCHUR_OPERANDS);
#undef KHS_SCHUR_OPERANDS
}

template <int B, int C>
__device__ __forceinline__ void fused_gj_step(double (&col)[6], double inv)
{
    static_assert(B == 0 || B == 6, "S columns on lanes 0..5 or 6..11");
    if constexpr (B == 0) {
        if constexpr (C == 0) KHS_GJ_ASM(0, 0, 1, 2, 3, 4, 5);
        if constexpr (C == 1) KHS_GJ_ASM(1, 1, 0, 2, 3, 4, 5);
        if constexpr (C == 2) KHS_GJ_ASM(2, 2, 0, 1, 3, 4, 5);
        if constexpr (C == 3) KHS_GJ_ASM(3, 3, 0, 1, 2, 4, 5);
        if constexpr (C == 4) KHS_GJ_ASM(4, 4, 0, 1, 2, 3, 5);
        if constexpr (C == 5) KHS_GJ_ASM(5, 5, 0, 1, 2, 3, 4);
    } else {
        if constexpr (C == 0) KHS_GJ_ASM(0, 6, 1, 2, 3, 4, 5);
        if constexpr (C == 1) KHS_GJ_ASM(1, 7, 0, 2, 3, 4, 5);
        if constexpr (C == 2) KHS_GJ_ASM(2, 8, 0, 1, 3, 4, 5);
        if constexpr (C == 3) KHS_GJ_ASM(3, 9, 0, 1, 2, 4, 5);
        if constexpr (C == 4) KHS_GJ_ASM(4, 10, 0, 1, 2, 3, 5);
        if constexpr (C == 5) KHS_GJ_ASM(5, 11, 0, 1, 2, 3, 4);
    }
}
#undef KHS_GJ_ASM
#undef KHS_GJ
#undef KHS_SCHUR
#undef KHS_SCHUR_K
#undef KHS_FMAC_NEG

// s_waitcnt vmcnt(0). The compiler's wait at a loop header covers every
// path into it: when the loads of a prefetch ring's first slot are issued
// last before the loop, it drains all memory operations at the top of every
// iteration. Waiting for them once before the loop leaves the header the
// back edge's wait alone (slots issued several steps earlier).
__device__ __forceinline__ void wait_memory_all() { __builtin_amdgcn_s_waitcnt(0x0f70); }

// 1 / x from the hardware reciprocal and two Newton steps (pivots are
// normal numbers here: the shift keeps them above 1e-24 unless the solve
// fails anyway)
__device__ __forceinline__ double reciprocal(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// blocks: D (36 N doubles), U (36 N) and g (6 N) in one allocation, in that
// order, below kMaxBufferBytes in all.
__global__ __launch_bounds__(kWave) void k_block_solve_fused(RodView m, const uint32_t *order, const double *blocks,
                                                             double *Cp, double *dp, double *dir, unsigned *failed,
                                                             unsigned *next_failed, Partial *partials,
                                                             const unsigned *stop)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && next_failed) *next_failed = 0;
    if (stop && *stop) {
        Partial skipped = empty_partial();
        skipped.failed = 1;
        block_reduce(skipped, partials + blockIdx.x);
        return;
    }
    const int g = threadIdx.x / kSolveLanes, j = threadIdx.x % kSolveLanes;
    const bool acol = j < 6, rlane = j == 12;
    const int half = j < 6 ? 0 : (j < 12 ? 1 : 2), ci = j < 12 ? j % 6 : 0;
    const uint32_t gi = blockIdx.x * kSolveGroups + g;
    uint32_t begin = 0, end = 0;
    if (gi < m.strand_count) {
        const uint32_t s = order[gi];
        const uint32_t first = m.strand_first[s];
        begin = first + m.strand_fixed[s];
        end = first + m.strand_count_of[s];
        if (acol)
            for (uint32_t n = first; n < begin; ++n) dir[6 * size_t(n) + j] = 0.0;
        if (begin > end) begin = end;
    }
    const uint32_t len = end - begin;
    uint32_t steps = len;
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) steps = max(steps, uint32_t(__shfl_xor(int(steps), o, kWave)));
    steps = __builtin_amdgcn_readfirstlane(steps);

    const uint32_t N = m.node_count;
    const uint32_t u_part = 288 * N, g_part = 576 * N;
    const __amdgpu_buffer_rsrc_t rB = buffer_of(blocks, 624 * uint64_t(N));
    const uint64_t block_bytes = 288 * uint64_t(N), vec_bytes = 48 * uint64_t(N);
    const __amdgpu_buffer_rsrc_t rC = buffer_of(Cp, block_bytes), rP = buffer_of(dp, vec_bytes);
    const __amdgpu_buffer_rsrc_t rG = buffer_of(blocks + 72 * size_t(N), vec_bytes), rX = buffer_of(dir, vec_bytes);

    // this lane's column of node t in parity P (S lanes: D; U lanes: U, while
    // a node follows; lane 12: g); row r at base + r * stride
    const uint32_t stride = rlane ? 8 : 48, per_node = rlane ? 48 : 288;
    const uint32_t lane_at[2] = {rlane ? g_part : 8 * ci + (half == 1 ? u_part : 0),
                                 rlane ? g_part : 8 * ci + (half == 0 ? u_part : 0)};
    auto fetch = [&](double (&f)[6], uint32_t t, auto P) __attribute__((always_inline)) {
        // bitwise and by selects, so that no branch splits the wavefront
        const uint32_t u_lane = half == 1 - int(P), now = (half == int(P)) | rlane;
        const uint32_t want = (now & uint32_t(t < len)) | (u_lane & uint32_t(t + 1 < len));
        const uint32_t at = __umul24(begin + t, per_node) + lane_at[P];
        const uint32_t base = want ? at : kNoAccess;
#pragma unroll
        for (int r = 0; r < 6; ++r) f[r] = buffer_load(rB, base + stride * r);
    };

    // The factors of each strand's last kFusedWindow nodes stay in LDS for
    // the back substitution (slot W is where lanes with nothing to keep
    // write); only earlier nodes' factors go through memory.
    __shared__ double window[kSolveGroups][kFusedWindow + 1][kFusedRecord];
    const bool ulane0 = half == 1, ulane1 = half == 0;  // U lanes at parity 0 and 1
    const int rec_at[2] = {ulane0 ? ci : (rlane ? 36 : 41), ulane1 ? ci : (rlane ? 36 : 41)};
    const int rec_stride[2] = {ulane0 ? 6 : (rlane ? 1 : 0), ulane1 ? 6 : (rlane ? 1 : 0)};

    bool ok = true;
    double col[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) col[r] = 0.0;
    // slot t % 4 holds node t's column; the slot of node t - 1 keeps U_{t-1}
    // (on the lanes that are node t's S lanes) until node t's Schur
    // complement, then takes node t + 3
    double ring[kFusedAhead][6];
#pragma unroll
    for (int r = 0; r < 6; ++r) ring[kFusedAhead - 1][r] = 0.0;
    static_for<kFusedAhead - 1>([&](auto A) { fetch(ring[A], A, std::integral_constant<int, A & 1>{}); });
    // g of each strand's last kFusedWindow nodes for the back substitution
    // (lane r < 6 its row), held from here so that its latency is long past
    // when the back substitution's window steps, which run from LDS, use it
    double g_window[kFusedWindow];
    static_for<kFusedWindow>([&](auto T) {
        const uint32_t n = end - 1 - T;
        g_window[T] = buffer_load(rG, (acol && T < len) ? 48 * n + 8 * j : kNoAccess);
    });
    wait_memory_all();

    auto forward = [&](uint32_t t, auto A) __attribute__((always_inline)) {
        constexpr int P = A & 1, B = P ? 6 : 0, prev = (A + kFusedAhead - 1) % kFusedAhead;
        const uint32_t n = begin + t;
        const bool live = t < len;
        // C'_{n-1} (S lanes) and e_{n-1} (lane 12) stay; the rest is zero
        const double keep_p = (half == P || rlane) ? 1.0 : 0.0;
        double p[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            p[r] = col[r] * keep_p;
            col[r] = ring[A][r];
        }
        fused_schur<B>(col, ring[prev], p);
        fetch(ring[prev], t + kFusedAhead - 1, std::integral_constant<int, (A + kFusedAhead - 1) & 1>{});
        double scale = 0.0;
        static_for<6>([&](auto I) { scale = fmax(scale, fabs(row_bcast<B + I>(col[I]))); });
        const double shift = fmax(1.0e-12 * scale, 1.0e-24);
        double smallest = 0.0;
        static_for<6>([&](auto C) {
            const double pivot = row_bcast<B + C>(col[C]) + shift;
            smallest = C == 0 ? pivot : fmin(smallest, pivot);
            fused_gj_step<B, C>(col, reciprocal(pivot));
        });
        const double e0 = row_bcast<12>(col[0]);
        const bool bad = !(smallest >= 1.0e-30) || !isfinite(e0);
        if (live && bad) ok = false;
        const bool keep = live && ok;
        const bool on_chip = t + kFusedWindow >= len;  // one of the strand's last kFusedWindow nodes
        const bool keeper = (half == 1 - P) || rlane;   // C' columns and e
        const uint32_t oc = (keep && !on_chip && half == 1 - P) ? 288 * n + 8 * ci : kNoAccess;
        const uint32_t op = (keep && !on_chip && rlane) ? 48 * n : kNoAccess;
        double *rec = window[g][(keep && on_chip && keeper) ? t + kFusedWindow - len : kFusedWindow];
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            buffer_store(rC, oc + 48 * r, col[r]);
            buffer_store(rP, op + 8 * r, col[r]);
            rec[rec_at[P] + rec_stride[P] * r] = col[r];
        }
    };
    for (uint32_t t0 = 0; t0 < steps; t0 += kFusedAhead)
        static_for<kFusedAhead>([&](auto A) { forward(t0 + A, A); });
    if (!ok && j == 0) atomicOr(failed, 1u);
    __syncthreads();  // the window's factors, written by other lanes, before they are read

    // Back substitution x_n = d'_n - C'_n x_{n+1} with d' = -e, lane r < 6
    // owning row r (as in k_block_solve).
    const int r = acol ? j : 0;
    struct BackSlot {
        double c[6], e, g;
    };
    // Two loops, so that no register is loaded from memory on one path and
    // from LDS on the other (the compiler then drains memory before every
    // LDS read): the first kFusedWindow steps (each strand's last nodes) take
    // C' and e from LDS and prefetch g; the rest run from memory through a
    // ring whose first slots load while the window steps run.
    auto fetch_back = [&](BackSlot &b, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = end - 1 - t;
        const bool bl = acol && t < len;
        const uint32_t oc = bl ? 288 * n + 48 * r : kNoAccess;
#pragma unroll
        for (int k = 0; k < 6; ++k) b.c[k] = buffer_load(rC, oc + 8 * k);
        b.e = buffer_load(rP, bl ? 48 * n + 8 * r : kNoAccess);
        b.g = buffer_load(rG, bl ? 48 * n + 8 * r : kNoAccess);
    };
    const bool beyond = steps > uint32_t(kFusedWindow);  // some strand has nodes outside the window
    BackSlot back[kBackAhead];
#pragma unroll
    for (int a = 0; a < kBackAhead; ++a) fetch_back(back[a], kFusedWindow + a);  // no access unless beyond
    double next = 0.0, slope = 0.0, reach2 = 0.0;
    auto backward = [&](const double (&c)[6], double e, double gv, uint32_t t) __attribute__((always_inline)) {
        const uint32_t n = end - 1 - t;
        double value = -e;
        if (t > 0) static_for<6>([&](auto K) { value = fma(-c[K], row_bcast<K>(next), value); });
        next = value;
        buffer_store(rX, (t < len && ok && acol) ? 48 * n + 8 * j : kNoAccess, value);
        const double sq = value * value;
        const double t2 = (sq + row_bcast<1>(sq)) + row_bcast<2>(sq);
        if (t < len) {
            slope += gv * value;
            reach2 = fmax(reach2, t2);
        }
    };
    static_for<kFusedWindow>([&](auto T) {
        if (T < steps) {
            const double *rec = window[g][kFusedWindow - 1 - T];
            double c[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) c[k] = rec[6 * r + k];
            backward(c, rec[36 + r], g_window[T], T);
        }
    });
    if (beyond) {
        wait_memory_all();  // the ring's first slots, loaded during the window steps
        for (uint32_t t0 = kFusedWindow; t0 < steps; t0 += kBackAhead)
            static_for<kBackAhead>([&](auto A) {
                backward(back[A].c, back[A].e, back[A].g, t0 + A);
                fetch_back(back[A], t0 + A + kBackAhead);
            });
    }
    double row_slope = 0.0;
    static_for<6>([&](auto K) { row_slope += row_bcast<K>(acol ? slope : 0.0); });
    Partial acc = empty_partial();
    if (j == 0 && gi < m.strand_count) {
        acc.sum[kSlope] = row_slope;
        acc.max_translation = sqrt(reach2);
        acc.failed = ok ? 0u : 1u;
    }
    block_reduce(acc, partials + blockIdx.x);
}

// Slope g.d and the largest translation of a free node (host then limits
// the direction by the trust radius, cuda_backend.cu:623-630).
__global__ __launch_bounds__(kBlock) void k_direction_stats(RodView m, const double *grad, const double *dir,
                                                            const unsigned *failed, const DirectionChoice *skip_if_used,
                                                            Partial *partials)
{
    if (skip_if_used && skip_if_used->used) return;
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    Partial acc = empty_partial();
    if (n < m.node_count) {
        double gd = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) gd += grad[6 * size_t(n) + k] * dir[6 * size_t(n) + k];
        acc.sum[kSlope] = gd;
        if (!(m.flags[n] & kNodeFixed)) acc.max_translation = norm(load3(dir + 6 * size_t(n)));
    }
    if (n == 0 && failed) acc.failed = *failed;
    block_reduce(acc, partials + blockIdx.x);
}

// Steepest descent preconditioned by the Jacobi diagonal (cuda_backend.cu:632-638),
// unless the choice kept the block direction.
__global__ void k_steepest(RodView m, const DirectionChoice *choice, const double *grad, const double *precond,
                           double *dir)
{
    if (choice->used) return;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(m.node_count) * 6) return;
    dir[i] = (m.flags[i / 6] & kNodeFixed) ? 0.0 : (-grad[i]) * precond[i];
}

// The host decisions of the direction step (cuda_backend.cu:623-648) made
// on the device from the reduced statistics: keep the block direction if
// it was solved and g.d < 0; scale a direction whose largest free-node
// translation exceeds the trust radius down to it.
__global__ void k_choose_direction(const Partial *stats, double trust, DirectionChoice *choice)
{
    if (threadIdx.x || blockIdx.x) return;
    const double slope = stats->sum[kSlope];
    const int used = !stats->failed && slope < 0 && isfinite(slope);
    choice->used = used;
    choice->scale = (used && stats->max_translation > trust) ? trust / stats->max_translation : 1.0;
}

// k_choose_direction and the block direction's scaling in one pass: every
// thread derives the choice from the reduced statistics, one records it.
// The choice is written to the device copy and to its host mirror.
__global__ void k_apply_direction_choice(const Partial *stats, double trust, DirectionChoice *choice,
                                         DirectionChoice *mirror, double *x, size_t count)
{
    const double slope = stats->sum[kSlope];
    const int used = !stats->failed && slope < 0 && isfinite(slope);
    const double scale = (used && stats->max_translation > trust) ? trust / stats->max_translation : 1.0;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i == 0) {
        choice->used = used;
        choice->scale = scale;
        mirror->used = used;
        mirror->scale = scale;
    }
    if (i < count && used && scale != 1.0) x[i] *= scale;
}

__global__ void k_force_steepest(DirectionChoice *choice, DirectionChoice *mirror)
{
    if (threadIdx.x || blockIdx.x) return;
    choice->used = mirror->used = 0;
    choice->scale = mirror->scale = 1.0;
}

// Trust-radius factor of the steepest-descent direction.
__global__ void k_choose_steepest_scale(const Partial *stats, double trust, DirectionChoice *choice,
                                        DirectionChoice *mirror)
{
    if (threadIdx.x || blockIdx.x || choice->used) return;
    choice->scale = mirror->scale = stats->max_translation > trust ? trust / stats->max_translation : 1.0;
}

// Applies the trust-radius factor to the direction of the given kind only
// (used = 1 block, 0 steepest), so each direction is scaled once.
__global__ void k_scale_direction(const DirectionChoice *choice, int used, double *x, size_t count)
{
    const double s = choice->scale;
    if (choice->used != used || s == 1.0) return;
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) x[i] *= s;
}

// First line-search step: the CCD bound times 0.9 when it binds
// (cuda_backend.cu:650-688).
__global__ void k_choose_alpha(const Partial *ccd, int has_collider, double *alpha)
{
    if (threadIdx.x || blockIdx.x) return;
    *alpha = step_from_ccd(ccd, has_collider);
}

__global__ void k_set_double(double *x, double v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *x = v;
}

__global__ void k_fill_u32(unsigned *x, size_t count, unsigned v)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) x[i] = v;
}

__global__ void k_fill_u64(unsigned long long *x, size_t count, unsigned long long v)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) x[i] = v;
}

// Jacobi diagonal: inertia plus per-element stretch and bend/twist
// stiffness bounds, node-gathered in element order (cuda_backend.cu:815-841).
__global__ void k_preconditioner(RodView m, double inv_dt, double damping, double translation_k, double rotation_a,
                                 double rotation_b, const double *material, double poisson, double shear_correction,
                                 double *precond)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.node_count) return;
    const double inv2 = inv_dt * inv_dt;
    const double pt = m.mass[n] * (inv2 + damping * inv_dt);
    const double pr = m.rotational_mass[n] * (inv2 + damping * inv_dt);
    // the material's bounds, or those of element L's strand
    auto bounds = [&](size_t L, double &tk, double &ra, double &rb) {
        if (!material) {
            tk = translation_k;
            ra = rotation_a;
            rb = rotation_b;
            return;
        }
        const double E = material[2 * L], rad = material[2 * L + 1];
        const double area = kPi * rad * rad, inertia = kPi * rad * rad * rad * rad / 4.0;
        const double G = E / (2.0 * (1.0 + poisson));
        tk = fmax(E * area, shear_correction * G * area);
        ra = shear_correction * G * area;
        rb = fmax(E * inertia, 2.0 * G * inertia);
    };
    double t = pt, r = pr;
    if (n > 0 && (m.flags[n - 1] & kNodeHasElement)) {
        const double L0 = m.rest[8 * size_t(n - 1)];
        double tk, ra, rb;
        bounds(n - 1, tk, ra, rb);
        t += tk / L0;
        r += ra * L0 + rb / L0;
    }
    if (m.flags[n] & kNodeHasElement) {
        const double L0 = m.rest[8 * size_t(n)];
        double tk, ra, rb;
        bounds(n, tk, ra, rb);
        t += tk / L0;
        r += ra * L0 + rb / L0;
    }
    const bool fixed = m.flags[n] & kNodeFixed;
    for (int d = 0; d < 3; ++d) {
        precond[6 * size_t(n) + d] = fixed ? 0.0 : 1.0 / fmax(t, 1.0e-30);
        precond[6 * size_t(n) + 3 + d] = fixed ? 0.0 : 1.0 / fmax(r, 1.0e-30);
    }
}

// ------------------------------------------------------------- substeps
// q_old = q for every node; fixed nodes move to the interpolated root
// targets and take the matching velocity (cuda_backend.cu:794-801).
__global__ void k_begin_substep(RodView m, double *q, double *q_old, double *vel, const double *root_a,
                                const double *rot_a, const double *root_b, const double *rot_b, double f, double dt)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.node_count) return;
    double cur[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        cur[k] = q[6 * size_t(n) + k];
        q_old[6 * size_t(n) + k] = cur[k];
    }
    if (!(m.flags[n] & kNodeFixed)) return;
    double val[6];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        val[k] = root_a[3 * size_t(n) + k] * (1.0 - f) + root_b[3 * size_t(n) + k] * f;
        val[3 + k] = rot_a[3 * size_t(n) + k] * (1.0 - f) + rot_b[3 * size_t(n) + k] * f;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        q[6 * size_t(n) + k] = val[k];
        vel[6 * size_t(n) + k] = (val[k] - cur[k]) / dt;
    }
}

// v = (q - q_old) / dt on free nodes (cuda_backend.cu:803-808).
__global__ void k_update_velocity(RodView m, const double *q, const double *q_old, double *vel, double inv_dt)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.node_count || (m.flags[n] & kNodeFixed)) return;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        vel[6 * size_t(n) + k] = (q[6 * size_t(n) + k] - q_old[6 * size_t(n) + k]) * inv_dt;
}

// previous = current; current = a (1 - f) + b f (cuda_backend.cu:770-774).
__global__ void k_collider_lerp(uint32_t count, double *current, double *previous, const double *a, const double *b,
                                double f)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(count) * 3) return;
    previous[i] = current[i];
    current[i] = a[i] * (1.0 - f) + b[i] * f;
}

__global__ void k_copy(double *dst, const double *src, size_t count)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i];
}

// Slope g.d and the step bound of every contact element by its own tree
// walk, one lane each: the path without candidate lists (with lists the
// swept check takes the slope and k_swept_ccd the bound).
__global__ __launch_bounds__(kBlock) void k_ccd_slope(RodView m, ColliderView col, StepParams p, const double *q,
                                                      const double *grad, const double *dir, Partial *partials)
{
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    Partial acc = empty_partial();
    if (L < m.node_count) {
        double gd = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) gd += grad[6 * size_t(L) + k] * dir[6 * size_t(L) + k];
        acc.sum[kSlope] = gd;
        const uint32_t fl = m.flags[L];
        if (p.has_collider && (fl & kNodeHasElement) && (fl & kNodeContact) && !(fl & kNodeFixed)) {
            const V3 p0 = load3(q + 6 * size_t(L)), p1 = load3(q + 6 * size_t(L + 1));
            const V3 d0 = load3(dir + 6 * size_t(L)), d1 = load3(dir + 6 * size_t(L + 1));
            const double speed = fmax(norm(d0), norm(d1));
            const double target = sweep_target_of(p, L);
            if (speed >= 1e-30) {
                const V3 lo = vmin(vmin(p0, p1), vmin(add(p0, d0), add(p1, d1)));
                const V3 hi = vmax(vmax(p0, p1), vmax(add(p0, d0), add(p1, d1)));
                double limit = 1.0;
                bvh_query(col.bvh, internal_count(col), col.leaf_box, lo, hi, target, [&](uint32_t t) {
                    const uint32_t *t3 = col.triangles + 3 * size_t(t);
                    limit = fmin(limit, advancement_limit(p0, p1, d0, d1, speed, vertex(col, t3[0]),
                                                          vertex(col, t3[1]), vertex(col, t3[2]), target));
                });
                acc.alpha = limit;
            }
        }
    }
    block_reduce(acc, partials + blockIdx.x);
}

// ---------------------------------------------------- collider sweep test
// Conservative advancement of each candidate triangle from its previous to
// its current position against the stationary segment
// (cuda_backend.cu:975-987); the lowest failing (element, triangle) wins.
// Whether triangle t, moving from its previous to its current position,
// cannot be advanced to the end of the substep against the stationary
// segment while keeping the target clearance (conservative advancement).
__device__ inline bool sweep_fails(const ColliderView &col, V3 p0, V3 p1, uint32_t t, double target)
{
    const uint32_t *t3 = col.triangles + 3 * size_t(t);
    const V3 oa = vertex_prev(col, t3[0]), ob = vertex_prev(col, t3[1]), oc = vertex_prev(col, t3[2]);
    const V3 da = sub(vertex(col, t3[0]), oa), db = sub(vertex(col, t3[1]), ob), dc = sub(vertex(col, t3[2]), oc);
    const double speed = fmax(norm(da), fmax(norm(db), norm(dc)));
    if (speed < 1e-30) return false;
    double alpha = 0.0;
    for (int it = 0; it < 80; ++it) {
        const Closest pair =
            closest_segment_triangle(p0, p1, add(oa, mul(da, alpha)), add(ob, mul(db, alpha)), add(oc, mul(dc, alpha)));
        const double clearance = pair.distance - target;
        if (clearance <= 0) return true;
        const double advance = 0.8 * clearance / speed;
        if (alpha + advance >= 1) return false;
        if (advance < 1e-12) return true;
        alpha += advance;
    }
    return true;
}

// Collider sweep by walking the swept tree, one lane per element; with
// only_overflow (the list counts) just the elements whose list overflowed.
__global__ void k_sweep(RodView m, ColliderView col, StepParams p, const double *q, const uint32_t *only_overflow,
                        unsigned long long *first_failure)
{
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= m.node_count) return;
    const uint32_t fl = m.flags[L];
    if (!(fl & kNodeHasElement) || !(fl & kNodeContact) || (fl & kNodeFixed)) return;
    if (only_overflow && only_overflow[L] != kCacheOverflow) return;
    const V3 p0 = load3(q + 6 * size_t(L)), p1 = load3(q + 6 * size_t(L + 1));
    const uint32_t element = L - m.strand_of[L];
    const double target = sweep_target_of(p, L);
    unsigned long long worst = ~0ull;
    bvh_query(col.bvh, internal_count(col), col.leaf_box, vmin(p0, p1), vmax(p0, p1), target, [&](uint32_t t) {
        if (sweep_fails(col, p0, p1, t, target)) {
            const unsigned long long key = (static_cast<unsigned long long>(element) << 32) | t;
            worst = key < worst ? key : worst;
        }
    });
    if (worst != ~0ull) atomicMin(first_failure, worst);
}

// Collider sweep from the candidate lists, one wavefront per listed element
// with a free left node, its lanes striding the list. The lists were
// gathered against the swept triangle boxes (`swept_box`), which the sweep
// queries too; the lowest failing (element, triangle) is kept.
__global__ __launch_bounds__(kBlock) void k_list_sweep(const uint32_t *flags, const uint32_t *strand_of,
                                                       ColliderView col, ContactCache cc, const double *swept_box,
                                                       StepParams p, const double *q,
                                                       unsigned long long *first_failure)
{
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t waves = gridDim.x * blockDim.x / kWave;
    const uint32_t listed = *cc.list_count;
    for (uint32_t j = wave; j < listed; j += waves) {
        const uint32_t L = cc.list[j];
        const uint32_t c = cc.count[L];
        if (c == 0 || c == kCacheOverflow || (flags[L] & kNodeFixed)) continue;
        const V3 p0 = load3(q + 6 * size_t(L)), p1 = load3(q + 6 * size_t(L + 1));
        const V3 lo = vmin(p0, p1), hi = vmax(p0, p1);
        const uint32_t element = L - strand_of[L];
        const double target = sweep_target_of(p, L);
        const uint32_t *tris = cc.tris + size_t(L) * kCacheCapacity;
        unsigned long long worst = ~0ull;
        for (uint32_t k = lane; k < c; k += kWave) {
            const uint32_t t = tris[k];
            const double *b = swept_box + 6 * size_t(t);
            if (!overlap(lo, hi, mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5]), target)) continue;
            if (sweep_fails(col, p0, p1, t, target)) {
                const unsigned long long key = (static_cast<unsigned long long>(element) << 32) | t;
                worst = key < worst ? key : worst;
            }
        }
#pragma unroll
        for (int o = kWave / 2; o >= 1; o >>= 1) {
            const unsigned long long other = __shfl_xor(worst, o, kWave);
            worst = other < worst ? other : worst;
        }
        if (lane == 0 && worst != ~0ull) atomicMin(first_failure, worst);
    }
}

// Recomputes the advancement of the reported pair to fill the diagnostics
// (cuda_backend.cu:966-973).
__global__ void k_sweep_record(RodView m, ColliderView col, StepParams p, const double *q,
                               const unsigned long long *first_failure, SweepRecord *out)
{
    if (threadIdx.x || blockIdx.x) return;
    const unsigned long long key = *first_failure;
    SweepRecord r{};
    if (key == ~0ull) {
        *out = r;
        return;
    }
    const uint32_t element = uint32_t(key >> 32), t = uint32_t(key & 0xffffffffu);
    // element e of strand s has left node e + s: find s with strand_first.
    uint32_t lo = 0, hi = m.strand_count;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (m.strand_first[mid] - mid <= element) lo = mid; else hi = mid;
    }
    const uint32_t L = element + lo;
    const double target = sweep_target_of(p, L);
    const V3 p0 = load3(q + 6 * size_t(L)), p1 = load3(q + 6 * size_t(L + 1));
    const uint32_t *t3 = col.triangles + 3 * size_t(t);
    const V3 oa = vertex_prev(col, t3[0]), ob = vertex_prev(col, t3[1]), oc = vertex_prev(col, t3[2]);
    const V3 da = sub(vertex(col, t3[0]), oa), db = sub(vertex(col, t3[1]), ob), dc = sub(vertex(col, t3[2]), oc);
    const double speed = fmax(norm(da), fmax(norm(db), norm(dc)));
    double alpha = 0.0;
    Closest pair{};
    V3 a = oa, b = ob, c = oc;
    for (int it = 0; it < 80; ++it) {
        a = add(oa, mul(da, alpha));
        b = add(ob, mul(db, alpha));
        c = add(oc, mul(dc, alpha));
        pair = closest_segment_triangle(p0, p1, a, b, c);
        const double clearance = pair.distance - target;
        if (clearance <= 0) break;
        const double advance = 0.8 * clearance / speed;
        if (alpha + advance >= 1 || advance < 1e-12) break;
        alpha += advance;
    }
    r.claimed = 1;
    r.strand = m.strand_of[L];
    r.element = element;
    r.triangle = t;
    r.distance = pair.distance;
    r.required = target;
    r.clearance = pair.distance - target;
    r.displacement = speed;
    store3(r.hair_start, p0);
    store3(r.hair_end, p1);
    store3(r.collider_point, add(add(mul(a, pair.bary.x), mul(b, pair.bary.y)), mul(c, pair.bary.z)));
    *out = r;
}

// -------------------------------------------------------------- friction
// Post-substep smoothed Coulomb impulse against the closest triangle
// (cuda_backend.cu:941-964); the velocity change of each node is gathered
// from its two elements in element order.
__global__ __launch_bounds__(kBlock) void k_friction(RodView m, ColliderView col, ContactCache cc,
                                                     const uint32_t *best_tri, StepParams p, const double *q,
                                                     const double *vel, double *dv)
{
    const int lane = threadIdx.x & (kWave - 1);
    const long tile = long(blockIdx.x) * kTilesPerBlock + threadIdx.x / kWave;
    const long L = tile * kTile - 1 + lane;
    const bool in_range = L >= 0 && L < long(m.node_count);
    const bool owner = lane >= 1 && in_range;
    const uint32_t fl = in_range ? m.flags[L] : 0u;
    const bool elem = (fl & kNodeHasElement) != 0;
    double di[3] = {0, 0, 0}, dj[3] = {0, 0, 0};
    if (elem && (fl & kNodeContact)) {
        const bool fixed_i = fl & kNodeFixed, fixed_j = m.flags[L + 1] & kNodeFixed;
        if (!(fixed_i && fixed_j)) {
            const V3 p0 = load3(q + 6 * L), p1 = load3(q + 6 * (L + 1));
            Closest pair;
            uint32_t tri;
            bool found;
            if (listed_query(cc, best_tri, uint32_t(L))) {
                tri = cc.count[L] ? best_tri[L] : kNoTriangle;
                found = tri != kNoTriangle;
                if (found)
                    pair = closest_segment_triangle(p0, p1, triangle_vertex(col, tri, 0), triangle_vertex(col, tri, 1),
                                                    triangle_vertex(col, tri, 2));
            } else {
                unsigned long long count = 0;
                found = closest_triangle(col, p0, p1, contact_pad_of(p, L), pair, tri, count);
            }
            if (found) {
                const double gap = pair.distance - radius_offset_of(p, L);
                if (gap > 0.0 && gap < p.barrier) {
                    const uint32_t *t3 = col.triangles + 3 * size_t(tri);
                    const double w0 = 1.0 - pair.t, w1 = pair.t;
                    const V3 hair = add(mul(p0, w0), mul(p1, w1));
                    const V3 body = add(add(mul(triangle_vertex(col, tri, 0), pair.bary.x),
                                            mul(triangle_vertex(col, tri, 1), pair.bary.y)),
                                        mul(triangle_vertex(col, tri, 2), pair.bary.z));
                    const V3 old_body =
                        add(add(mul(vertex_prev(col, t3[0]), pair.bary.x), mul(vertex_prev(col, t3[1]), pair.bary.y)),
                            mul(vertex_prev(col, t3[2]), pair.bary.z));
                    const V3 normal = mul(sub(hair, body), 1.0 / fmax(pair.distance, 1.0e-15));
                    const V3 v0 = load3(vel + 6 * L), v1 = load3(vel + 6 * (L + 1));
                    const V3 rel = sub(add(mul(v0, w0), mul(v1, w1)), mul(sub(body, old_body), 1.0 / p.dt));
                    const V3 tangent = sub(rel, mul(normal, dot(normal, rel)));
                    const double tspeed = norm(tangent);
                    double inv_mass = 0.0;
                    if (!fixed_i) inv_mass += w0 * w0 / fmax(m.mass[L], 1.0e-30);
                    if (!fixed_j) inv_mass += w1 * w1 / fmax(m.mass[L + 1], 1.0e-30);
                    if (tspeed >= 1.0e-15 && inv_mass > 0.0) {
                        const double aa = p.barrier - gap, lg = log(gap / p.barrier);
                        const double fn = p.stiffness * (aa * aa / gap - 2.0 * aa * lg);
                        const double mag = fmin(tspeed / inv_mass, p.friction * fn * p.dt);
                        const double smooth = p.smoothing / p.dt;
                        const V3 imp = mul(tangent, -mag / fmax(sqrt(tspeed * tspeed + smooth * smooth), 1.0e-30));
                        if (!fixed_i) {
                            const double sc = w0 / fmax(m.mass[L], 1.0e-30);
                            di[0] = imp.x * sc; di[1] = imp.y * sc; di[2] = imp.z * sc;
                        }
                        if (!fixed_j) {
                            const double sc = w1 / fmax(m.mass[L + 1], 1.0e-30);
                            dj[0] = imp.x * sc; dj[1] = imp.y * sc; dj[2] = imp.z * sc;
                        }
                    }
                }
            }
        }
    }
    double total[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) total[k] = (0.0 + from_lane_below(dj[k])) + di[k];
    if (owner) {
#pragma unroll
        for (int k = 0; k < 3; ++k) dv[3 * L + k] = total[k];
    }
}

__global__ void k_apply_dv(uint32_t n, const double *dv, double *vel)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) vel[6 * size_t(i) + k] = 1.0 * dv[3 * size_t(i) + k] + vel[6 * size_t(i) + k];
}

// --------------------------------------------------- build-time checks
// Ray-parity point-in-mesh test along a slightly skewed +x ray for free,
// visible nodes (reference solver.cpp:608-645); candidates come from the
// BVH with a box enclosing the ray inside the collider bounds.
__global__ void k_point_inside(RodView m, ColliderView col, const double *pos, double x_max, unsigned *first_inside)
{
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.node_count) return;
    if (m.flags[n] & (kNodeFixed | kNodeVirtual)) return;
    const V3 pt = load3(pos + 3 * size_t(n));
    const V3 dirv = mk(1.0, 0.0000012345, 0.0000023456);
    const double reach = fmax(x_max - pt.x, 0.0) + 1.0e-6;
    const V3 lo = pt, hi = add(pt, mul(dirv, reach));
    unsigned hits = 0;
    bvh_query(col.bvh, internal_count(col), col.leaf_box, vmin(lo, hi), vmax(lo, hi), 1.0e-9, [&](uint32_t t) {
        const uint32_t *t3 = col.triangles + 3 * size_t(t);
        const V3 a = vertex(col, t3[0]), b = vertex(col, t3[1]), c = vertex(col, t3[2]);
        const V3 e1 = sub(b, a), e2 = sub(c, a);
        const V3 h = mk(dirv.y * e2.z - dirv.z * e2.y, dirv.z * e2.x - dirv.x * e2.z, dirv.x * e2.y - dirv.y * e2.x);
        const double det = dot(e1, h);
        if (fabs(det) < 1.0e-14) return;
        const double inv = 1.0 / det;
        const V3 s = sub(pt, a);
        const double u = inv * dot(s, h);
        if (u < 0.0 || u > 1.0) return;
        const V3 qv = mk(s.y * e1.z - s.z * e1.y, s.z * e1.x - s.x * e1.z, s.x * e1.y - s.y * e1.x);
        const double v = inv * dot(dirv, qv);
        if (v < 0.0 || u + v > 1.0) return;
        if (inv * dot(e2, qv) > 1.0e-12) ++hits;
    });
    if (hits & 1u) atomicMin(first_inside, n);
}

// Minimum gap of free contact elements with the intersection-aware
// distance (reference solver.cpp:659-681).
__global__ __launch_bounds__(kBlock) void k_initial_gap(RodView m, ColliderView col, StepParams p, const double *pos,
                                                        Partial *partials)
{
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    Partial acc = empty_partial();
    if (L < m.node_count) {
        const uint32_t fl = m.flags[L];
        if ((fl & kNodeHasElement) && (fl & kNodeContact) && !(fl & kNodeFixed)) {
            const V3 p0 = load3(pos + 3 * size_t(L)), p1 = load3(pos + 3 * size_t(L + 1));
            const double radius_offset = radius_offset_of(p, L);
            bvh_query(col.bvh, internal_count(col), col.leaf_box, vmin(p0, p1), vmax(p0, p1), contact_pad_of(p, L),
                      [&](uint32_t t) {
                const uint32_t *t3 = col.triangles + 3 * size_t(t);
                const Closest c = closest_segment_triangle(p0, p1, vertex(col, t3[0]), vertex(col, t3[1]),
                                                           vertex(col, t3[2]), true);
                acc.min_gap = fmin(acc.min_gap, c.distance - radius_offset);
                ++acc.candidates;
            });
        }
    }
    block_reduce(acc, partials + blockIdx.x);
}

// Every (element, triangle) box overlap of contact-enabled free elements.
__global__ void k_broadphase_pairs(RodView m, ColliderView col, StepParams p, const double *q, uint32_t *pairs,
                                   unsigned long long capacity, unsigned long long *count)
{
    const uint32_t L = blockIdx.x * blockDim.x + threadIdx.x;
    if (L >= m.node_count) return;
    const uint32_t fl = m.flags[L];
    if (!(fl & kNodeHasElement) || !(fl & kNodeContact) || (fl & kNodeFixed)) return;
    const V3 p0 = load3(q + 6 * size_t(L)), p1 = load3(q + 6 * size_t(L + 1));
    const uint32_t element = L - m.strand_of[L];
    bvh_query(col.bvh, internal_count(col), col.leaf_box, vmin(p0, p1), vmax(p0, p1), contact_pad_of(p, L),
              [&](uint32_t t) {
        const unsigned long long slot = atomicAdd(count, 1ull);
        if (slot < capacity) {
            pairs[2 * slot] = element;
            pairs[2 * slot + 1] = t;
        }
    });
}

__global__ void k_gather_positions(const double *q, const uint32_t *map, uint32_t count, double *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t n = map ? map[i] : i;
#pragma unroll
    for (int k = 0; k < 3; ++k) out[3 * size_t(i) + k] = q[6 * size_t(n) + k];
}

}  // namespace dev
}  // namespace khs
