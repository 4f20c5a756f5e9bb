// CPU FP64 oracle of the reference hair step, exported through the same C
// ABI (kami_hair_solver.h) and test hooks (kami_hair_solver_ext.h) as the
// HIP library, so one test script drives both.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's CPU baseline as the checker; never linked by the product.
//
// It restates, function by function:
//   model build        reference src/solver.cpp:227-447, 608-681
//   incremental step   reference src/cuda_backend.cu:604-723 (host control)
//                      and :794-1000 (per-node / per-element formulas)
// with fixed-order reductions (strand by strand, node then element order)
// instead of the reference's floating-point atomics. Broad-phase candidates
// are the exact set of (element, triangle) pairs whose boxes overlap with
// the padding; among candidates the closest triangle wins and equal
// distances go to the lower triangle index.
#include "../include/kami_hair_solver.h"
#include "../include/kami_hair_solver_ext.h"
#include "oracle_geometry.hpp"
#include "oracle_math.hpp"

#include <atomic>
#include <cerrno>
#include <chrono>
#include <cstring>
#include <limits>
#include <new>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include <unistd.h>

namespace oracle {
namespace {

constexpr double kPi = 3.1415926535897932384626433832795;
constexpr uint64_t kCheckpointMagic = 0x3154504348534bULL;
const double kInf = std::numeric_limits<double>::infinity();

struct Node {
    V3 x, rotation;
    M3 frame;
    double mass = 0.0, rot_mass = 0.0;
    bool fixed = false, virt = false;
    uint32_t strand = 0;
    uint32_t bind_left = 0, bind_right = 0;
    double bind_t = 0.0;
};

struct Element {
    uint32_t i = 0, j = 0, strand = 0;
    bool contact = true;
    double rest_length = 0.0;
    V3 rest_shear, rest_curvature;
};

struct Strand {
    uint32_t first = 0, count = 0, fixed_count = 0;
};

struct Stats {
    double objective = 0, elastic = 0, contact = 0, friction = 0, kinetic = 0;
    double min_gap = kInf;
    uint64_t candidates = 0, active = 0;
    bool feasible = true;
    void add(const Stats &o)
    {
        objective += o.objective;
        elastic += o.elastic;
        contact += o.contact;
        friction += o.friction;
        kinetic += o.kinetic;
        min_gap = std::fmin(min_gap, o.min_gap);
        candidates += o.candidates;
        active += o.active;
        feasible = feasible && o.feasible;
    }
};

struct SweepFailure {
    bool claimed = false;
    uint32_t strand = 0, element = 0, triangle = 0;
    double distance = 0, required = 0, clearance = 0, displacement = 0;
    V3 hair_start, hair_end, collider_point;
};

template <class F>
void parallel_for(uint32_t threads, uint32_t n, F &&f)
{
    if (threads <= 1 || n < 2) {
        for (uint32_t i = 0; i < n; ++i) f(i);
        return;
    }
    const uint32_t T = std::min(threads, n);
    std::vector<std::thread> pool;
    pool.reserve(T);
    for (uint32_t t = 0; t < T; ++t)
        pool.emplace_back([&, t] {
            const uint32_t b = uint32_t(uint64_t(n) * t / T), e = uint32_t(uint64_t(n) * (t + 1) / T);
            for (uint32_t i = b; i < e; ++i) f(i);
        });
    for (auto &th : pool) th.join();
}

// Rod strains (shear/stretch, then curvature/twist) of one element as
// functions of its 12 coordinates (cuda_backend.cu:234-245).
template <int N>
void rod_strains(const Dual<N> *q, const M3 &fi, const M3 &fj, const Element &e, Dual<N> *strain)
{
    typedef Dual<N> D;
    TV3<D> xi{q[0], q[1], q[2]}, ui{q[3], q[4], q[5]}, xj{q[6], q[7], q[8]}, uj{q[9], q[10], q[11]};
    TM3<D> ri0, rj0;
    for (int k = 0; k < 9; ++k) {
        ri0.m[k] = D(fi.m[k]);
        rj0.m[k] = D(fj.m[k]);
    }
    const TM3<D> ri = tmmul(exp_so3(ui), ri0), rj = tmmul(exp_so3(uj), rj0);
    const TV3<D> tangent = tscale(tsub(xj, xi), D(1.0 / e.rest_length));
    const TV3<D> shear = tscale(tmvec(tmadd(ttranspose(ri), ttranspose(rj)), tangent), D(0.5));
    const TV3<D> curvature = tscale(cayley(tmmul(ttranspose(ri), rj)), D(1.0 / e.rest_length));
    strain[0] = shear.x - D(e.rest_shear.x);
    strain[1] = shear.y - D(e.rest_shear.y);
    strain[2] = shear.z - D(e.rest_shear.z);
    strain[3] = curvature.x - D(e.rest_curvature.x);
    strain[4] = curvature.y - D(e.rest_curvature.y);
    strain[5] = curvature.z - D(e.rest_curvature.z);
}

// Gauss-Jordan inverse of a 6x6 block with partial pivoting after adding
// max(1e-12 * max|a_ii|, 1e-24) to the diagonal; fails on a pivot below
// 1e-30 or a non-finite pivot (cuda_backend.cu:855-864). Each row update
// a - f b is one fused multiply-add, as the reference's CUDA build contracts
// it (nvcc's default); so are the sums of products in the block solve.
// The same Gauss-Jordan inverse in single precision, for the
// mixed-precision step (the blocks and factors rounded to FP32).
bool inverse6f(const float *src, float *inv)
{
    float a[36];
    for (int i = 0; i < 36; ++i) { a[i] = src[i]; inv[i] = 0.0f; }
    for (int i = 0; i < 6; ++i) inv[6 * i + i] = 1.0f;
    float scale = 0.0f;
    for (int i = 0; i < 6; ++i) scale = std::fmax(scale, std::fabs(a[6 * i + i]));
    for (int i = 0; i < 6; ++i) a[6 * i + i] += std::fmax(1.0e-12f * scale, 1.0e-24f);
    for (int col = 0; col < 6; ++col) {
        int pivot = col;
        float best = std::fabs(a[6 * col + col]);
        for (int row = col + 1; row < 6; ++row)
            if (std::fabs(a[6 * row + col]) > best) { best = std::fabs(a[6 * row + col]); pivot = row; }
        if (best < 1.0e-30f || !std::isfinite(best)) return false;
        if (pivot != col)
            for (int k = 0; k < 6; ++k) {
                std::swap(a[6 * col + k], a[6 * pivot + k]);
                std::swap(inv[6 * col + k], inv[6 * pivot + k]);
            }
        const float invp = 1.0f / a[6 * col + col];
        for (int k = 0; k < 6; ++k) { a[6 * col + k] *= invp; inv[6 * col + k] *= invp; }
        for (int row = 0; row < 6; ++row)
            if (row != col) {
                const float f = a[6 * row + col];
                for (int k = 0; k < 6; ++k) {
                    a[6 * row + k] = std::fma(-f, a[6 * col + k], a[6 * row + k]);
                    inv[6 * row + k] = std::fma(-f, inv[6 * col + k], inv[6 * row + k]);
                }
            }
    }
    return true;
}

bool inverse6(const double *src, double *inv)
{
    double a[36];
    for (int i = 0; i < 36; ++i) { a[i] = src[i]; inv[i] = 0.0; }
    for (int i = 0; i < 6; ++i) inv[6 * i + i] = 1.0;
    double scale = 0.0;
    for (int i = 0; i < 6; ++i) scale = std::fmax(scale, std::fabs(a[6 * i + i]));
    for (int i = 0; i < 6; ++i) a[6 * i + i] += std::fmax(1.0e-12 * scale, 1.0e-24);
    for (int col = 0; col < 6; ++col) {
        int pivot = col;
        double best = std::fabs(a[6 * col + col]);
        for (int row = col + 1; row < 6; ++row)
            if (std::fabs(a[6 * row + col]) > best) { best = std::fabs(a[6 * row + col]); pivot = row; }
        if (best < 1.0e-30 || !std::isfinite(best)) return false;
        if (pivot != col)
            for (int k = 0; k < 6; ++k) {
                std::swap(a[6 * col + k], a[6 * pivot + k]);
                std::swap(inv[6 * col + k], inv[6 * pivot + k]);
            }
        const double invp = 1.0 / a[6 * col + col];
        for (int k = 0; k < 6; ++k) { a[6 * col + k] *= invp; inv[6 * col + k] *= invp; }
        for (int row = 0; row < 6; ++row)
            if (row != col) {
                const double f = a[6 * row + col];
                for (int k = 0; k < 6; ++k) {
                    a[6 * row + k] = std::fma(-f, a[6 * col + k], a[6 * row + k]);
                    inv[6 * row + k] = std::fma(-f, inv[6 * col + k], inv[6 * row + k]);
                }
            }
    }
    return true;
}

uint64_t edge_key(uint32_t a, uint32_t b)
{
    if (a > b) std::swap(a, b);
    return (uint64_t(a) << 32u) | uint64_t(b);
}

V3 from_c(const KhsVec3 &v) { return {v.x, v.y, v.z}; }
KhsVec3 to_c(V3 v) { return {v.x, v.y, v.z}; }

class Solver {
public:
    explicit Solver(const KhsSolverDesc &desc) : desc_(desc)
    {
        build_stats_.struct_size = sizeof(KhsBuildStats);
        step_stats_.struct_size = sizeof(KhsStepStats);
        step_stats_.phase = KHS_PHASE_IDLE;
        clear_failure();
    }

    std::string error;
    uint32_t threads = 1;
    uint32_t convergence_mode = 0;
    uint32_t direction_mode = 0;
    uint32_t precision_mode = 0;
    uint32_t solve_algorithm = 0;  // 0 block Thomas from the root, 1 twisted (from both ends), 2 fused

    // ---------------------------------------------------------------- input
    KhsResult set_hair(const KhsVec3 *points, uint32_t point_count, const uint32_t *offsets,
                       uint32_t strand_count, const KhsHairMaterial &m)
    {
        // solver.cpp:136-177
        if (!points || !offsets || point_count < 2 || strand_count == 0 || offsets[0] != 0 ||
            offsets[strand_count] != point_count) {
            error = "Invalid hair points or strand offsets.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        for (uint32_t s = 0; s < strand_count; ++s)
            if (offsets[s + 1] <= offsets[s] || offsets[s + 1] > point_count) {
                error = "Hair strand offsets are not increasing.";
                return KHS_ERROR_INVALID_HAIR;
            }
        if (!(m.density > 0.0 && m.radius > 0.0 && m.young_modulus > 0.0 && m.poisson_ratio > -1.0 &&
              m.poisson_ratio < 0.5 && m.shear_correction > 0.0 && m.contact_stiffness > 0.0 &&
              m.barrier_distance > 0.0 && m.friction >= 0.0 && m.friction_smoothing > 0.0)) {
            error = "Hair material is out of range.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        std::vector<V3> pts(point_count);
        for (uint32_t i = 0; i < point_count; ++i) {
            pts[i] = from_c(points[i]);
            if (!finite3(pts[i])) {
                error = "Hair point " + std::to_string(i) + " is not finite.";
                return KHS_ERROR_INVALID_HAIR;
            }
        }
        clear_built();
        input_points_ = pts;
        input_offsets_.assign(offsets, offsets + strand_count + 1);
        strand_E_.clear();
        strand_r_.clear();
        material_ = m;
        has_hair_ = true;
        error.clear();
        return KHS_OK;
    }

    KhsResult set_collider(const KhsVec3 *vertices, uint32_t vertex_count, const KhsTriangle *triangles,
                           uint32_t triangle_count)
    {
        // solver.cpp:179-225
        if (triangle_count == 0) {
            has_collider_ = false;
            collider_pending_.clear();
            collider_input_.clear();
            tris_.clear();
            excluded_ = 0;
            built_ = false;
            return KHS_OK;
        }
        if (!vertices || !triangles || vertex_count < 3) {
            error = "Invalid collider vertices or triangles.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        collider_input_.resize(vertex_count);
        for (uint32_t i = 0; i < vertex_count; ++i) collider_input_[i] = from_c(vertices[i]);
        tris_.clear();
        excluded_ = 0;
        for (uint32_t k = 0; k < triangle_count; ++k) {
            const KhsTriangle &t = triangles[k];
            if (t.i0 >= vertex_count || t.i1 >= vertex_count || t.i2 >= vertex_count || t.i0 == t.i1 ||
                t.i1 == t.i2 || t.i2 == t.i0 || !finite3(collider_input_[t.i0]) ||
                !finite3(collider_input_[t.i1]) || !finite3(collider_input_[t.i2])) {
                ++excluded_;
                continue;
            }
            const V3 c = cross(collider_input_[t.i1] - collider_input_[t.i0],
                               collider_input_[t.i2] - collider_input_[t.i0]);
            if (norm2(c) <= 1.0e-24) {
                ++excluded_;
                continue;
            }
            tris_.push_back({t.i0, t.i1, t.i2});
        }
        if (tris_.empty()) {
            error = "The collider has no valid triangle.";
            return KHS_ERROR_INVALID_COLLIDER;
        }
        has_collider_ = true;
        built_ = false;
        error.clear();
        return KHS_OK;
    }

    // ---------------------------------------------------------------- build
    KhsResult build()
    {
        // solver.cpp:449-556
        if (!has_hair_) {
            error = "Set hair curves first.";
            return KHS_ERROR_INVALID_STATE;
        }
        step_stats_.phase = KHS_PHASE_PREPARING;
        if (!make_internal_mesh()) {
            step_stats_.phase = KHS_PHASE_FAILED;
            return KHS_ERROR_INVALID_HAIR;
        }
        init_frames_and_elements();
        init_masses();
        free_dofs_ = 0;
        for (const Node &n : nodes_) if (!n.fixed) free_dofs_ += 6;
        std::vector<V3> tp, tr;
        if (!root_targets(input_points_, tp, tr)) {
            error = "Cannot build the initial root orientation.";
            step_stats_.phase = KHS_PHASE_FAILED;
            return KHS_ERROR_INVALID_HAIR;
        }
        if (has_collider_) {
            validate_topology();
            collider_cur_ = collider_input_;
            bvh_.build(collider_cur_, tris_);
            collider_x_max_ = -kInf;
            for (const V3 &v : collider_cur_) collider_x_max_ = std::fmax(collider_x_max_, v.x);
            if (!initial_feasibility()) {
                step_stats_.phase = KHS_PHASE_FAILED;
                return KHS_ERROR_INITIAL_INTERSECTION;
            }
        }
        init_state();
        built_ = true;
        build_stats_ = {};
        build_stats_.struct_size = sizeof(KhsBuildStats);
        build_stats_.strand_count = uint32_t(strands_.size());
        build_stats_.original_point_count = uint32_t(input_points_.size());
        build_stats_.internal_node_count = uint32_t(nodes_.size());
        build_stats_.element_count = uint32_t(elements_.size());
        uint32_t fixed = 0;
        for (const Node &n : nodes_) fixed += n.fixed ? 1 : 0;
        build_stats_.fixed_node_count = fixed;
        build_stats_.collider_vertex_count = uint32_t(has_collider_ ? collider_input_.size() : 0);
        build_stats_.collider_triangle_count = uint32_t(tris_.size());
        build_stats_.excluded_collider_triangle_count = excluded_;
        build_stats_.collider_boundary_edge_count = boundary_edges_;
        build_stats_.collider_nonmanifold_edge_count = nonmanifold_edges_;
        build_stats_.collider_inconsistent_edge_count = inconsistent_edges_;
        build_stats_.collider_inverted_closed_component_count = inverted_;
        build_stats_.merged_zero_length_segment_count = merged_;
        build_stats_.virtual_extension_strand_count = ext_strands_;
        build_stats_.virtual_extension_node_count = ext_nodes_;
        build_stats_.virtual_extension_rest_length = ext_length_;
        build_stats_.degree_of_freedom_count = free_dofs_;
        build_stats_.estimated_bytes = resident_bytes();
        build_stats_.initial_minimum_gap = has_collider_ ? current_minimum_gap(nullptr) : kInf;
        step_stats_.phase = KHS_PHASE_IDLE;
        error.clear();
        return KHS_OK;
    }

    KhsResult update_runtime(const KhsSolverDesc &d, const KhsHairMaterial &m)
    {
        // solver.cpp:558-606, cuda_backend.cu:409-430
        if (!built_) {
            error = "Build the solver before updating runtime parameters.";
            return KHS_ERROR_INVALID_STATE;
        }
        if (d.maximum_element_length != desc_.maximum_element_length ||
            d.minimum_dynamic_length != desc_.minimum_dynamic_length || d.fixed_root_nodes != desc_.fixed_root_nodes) {
            error = "maximum_element_length, minimum_dynamic_length and fixed_root_nodes cannot change.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        desc_ = d;
        material_ = m;
        init_masses();
        trust_radius_ = std::fmax(0.25 * max_rest_length_,
                                  8.0 * (max_radius() + material_.collider_offset + material_.barrier_distance));
        error.clear();
        return KHS_OK;
    }

    KhsResult update_collider(const KhsVec3 *vertices, uint32_t count)
    {
        // solver.cpp:683-708
        if (!built_ || !has_collider_) {
            error = "No built collider.";
            return KHS_ERROR_INVALID_STATE;
        }
        if (!vertices || count != collider_input_.size()) {
            error = "Collider vertex count changed.";
            return KHS_ERROR_INVALID_COLLIDER;
        }
        std::vector<V3> v(count);
        for (uint32_t i = 0; i < count; ++i) {
            v[i] = from_c(vertices[i]);
            if (!finite3(v[i])) {
                error = "Collider vertex is not finite.";
                return KHS_ERROR_INVALID_COLLIDER;
            }
        }
        collider_pending_ = v;
        return KHS_OK;
    }

    KhsResult update_roots(const KhsVec3 *points, uint32_t count)
    {
        // solver.cpp:710-742
        if (!built_ || !points || count != input_points_.size()) {
            error = "Root target count does not match the hair curves.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        std::vector<V3> p(count);
        for (uint32_t i = 0; i < count; ++i) {
            p[i] = from_c(points[i]);
            if (!finite3(p[i])) {
                error = "Root target is not finite.";
                return KHS_ERROR_INVALID_HAIR;
            }
        }
        std::vector<V3> tp, tr;
        if (!root_targets(p, tp, tr)) {
            error = "Cannot build root orientation from the targets.";
            return KHS_ERROR_INVALID_HAIR;
        }
        root_pending_ = tp;
        rot_pending_ = tr;
        return KHS_OK;
    }

    // ------------------------------------------------------------ animation
    KhsResult allocate_animation(uint32_t frames)
    {
        if (!built_) return KHS_ERROR_INVALID_STATE;
        if (frames < 1) {
            error = "Invalid animation frame count.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        anim_frames_ = frames;
        std::fill(ring_frame_.begin(), ring_frame_.end(), ~0u);
        anim_root_.assign(size_t(frames) * nodes_.size(), V3{});
        anim_rot_.assign(size_t(frames) * nodes_.size(), V3{});
        anim_collider_.assign(has_collider_ ? size_t(frames) * collider_input_.size() : 0, V3{});
        root_set_.assign(frames, false);
        collider_set_.assign(frames, !has_collider_);
        return KHS_OK;
    }

    KhsResult set_root_frame(uint32_t frame, const KhsVec3 *points, uint32_t count)
    {
        if (!built_ || !points || count != input_points_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        std::vector<V3> p(count);
        for (uint32_t i = 0; i < count; ++i) {
            p[i] = from_c(points[i]);
            if (!finite3(p[i])) return KHS_ERROR_INVALID_HAIR;
        }
        std::vector<V3> tp, tr;
        if (!root_targets(p, tp, tr)) return KHS_ERROR_INVALID_HAIR;
        if (frame >= anim_frames_) {
            error = "Invalid root animation frame.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        std::copy(tp.begin(), tp.end(), anim_root_.begin() + size_t(frame) * nodes_.size());
        std::copy(tr.begin(), tr.end(), anim_rot_.begin() + size_t(frame) * nodes_.size());
        root_set_[frame] = true;
        return KHS_OK;
    }

    KhsResult set_collider_frame(uint32_t frame, const KhsVec3 *vertices, uint32_t count)
    {
        if (!built_) return KHS_ERROR_INVALID_STATE;
        const size_t vc = has_collider_ ? collider_input_.size() : 0;
        if (frame >= anim_frames_ || count != vc || (count > 0 && !vertices)) {
            error = "Invalid collider animation frame.";
            return KHS_ERROR_INVALID_COLLIDER;
        }
        for (uint32_t i = 0; i < count; ++i) anim_collider_[size_t(frame) * vc + i] = from_c(vertices[i]);
        collider_set_[frame] = true;
        return KHS_OK;
    }

    KhsResult finalize_animation()
    {
        if (!built_) return KHS_ERROR_INVALID_STATE;
        if (anim_frames_ < 1 || std::find(root_set_.begin(), root_set_.end(), false) != root_set_.end() ||
            std::find(collider_set_.begin(), collider_set_.end(), false) != collider_set_.end()) {
            error = "Animation has missing frames.";
            return KHS_ERROR_INVALID_STATE;
        }
        anim_ready_ = true;
        anim_current_ = 0;
        return KHS_OK;
    }

    // ---------------------------------------------------------------- steps
    KhsResult step(double frame_dt)
    {
        if (!built_) {
            error = "Build the solver first.";
            return KHS_ERROR_INVALID_STATE;
        }
        if (!(frame_dt > 0.0) || !std::isfinite(frame_dt)) {
            error = "The time step must be positive and finite.";
            return KHS_ERROR_INVALID_ARGUMENT;
        }
        const size_t vc = collider_input_.size();
        return run_frame(frame_dt, root_cur_.data(), rot_cur_.data(), root_pending_.data(), rot_pending_.data(),
                         has_collider_ ? collider_frame_start_.data() : nullptr,
                         has_collider_ ? collider_pending_.data() : nullptr, 0, 0, vc);
    }

    KhsResult step_animation_frame(uint32_t frame, double frame_dt)
    {
        // cuda_backend.cu:488-498
        if (!built_) return KHS_ERROR_INVALID_STATE;
        if (!anim_ready_ || frame >= anim_frames_) {
            error = "Invalid animation frame index.";
            return fail(KHS_ERROR_INVALID_ARGUMENT, error);
        }
        if (frame == 0) {
            anim_current_ = 0;
            keep_in_ring(0);
            return KHS_OK;
        }
        if (frame != anim_current_ + 1) return fail(KHS_ERROR_INVALID_STATE, "Animation frames must be stepped in order.");
        const size_t n = nodes_.size(), vc = has_collider_ ? collider_input_.size() : 0;
        const uint32_t a = frame - 1, b = frame;
        KhsResult r = run_frame(frame_dt, &anim_root_[a * n], &anim_rot_[a * n], &anim_root_[b * n], &anim_rot_[b * n],
                                vc ? &anim_collider_[a * vc] : nullptr, vc ? &anim_collider_[b * vc] : nullptr, frame,
                                anim_frames_, vc);
        if (r == KHS_OK) {
            anim_current_ = frame;
            keep_in_ring(frame);
        }
        return r;
    }

    void keep_in_ring(uint32_t frame)
    {
        if (ring_frame_.empty()) return;
        const size_t slot = frame % ring_frame_.size();
        ring_q_[slot] = q_;
        ring_v_[slot] = v_;
        ring_frame_[slot] = frame;
    }

    // One .khc frame record: 3 doubles per original point, native order.
    KhsResult write_cache_frame(int fd)
    {
        if (!built_) return KHS_ERROR_INVALID_STATE;
        if (fd < 0) return KHS_ERROR_INVALID_ARGUMENT;
        std::vector<KhsVec3> out(mapping_.size());
        copy_original(out.data(), uint32_t(out.size()));
        const char *p = reinterpret_cast<const char *>(out.data());
        for (size_t left = out.size() * sizeof(KhsVec3); left;) {
            const ssize_t w = ::write(fd, p, left);
            if (w < 0 && errno == EINTR) continue;
            if (w <= 0) {
                error = std::string("Writing the cache frame failed: ") + std::strerror(w < 0 ? errno : EIO);
                return KHS_ERROR_INTERNAL;
            }
            p += w;
            left -= size_t(w);
        }
        return KHS_OK;
    }

    // Rollback ring (host memory here): the state after each stepped frame
    // in slot frame % size.
    KhsResult set_rollback_ring(uint32_t frames)
    {
        if (!built_) return KHS_ERROR_INVALID_STATE;
        ring_frame_.assign(frames, ~0u);
        ring_q_.assign(frames, {});
        ring_v_.assign(frames, {});
        return KHS_OK;
    }

    KhsResult rollback(uint32_t frame)
    {
        if (!built_ || !anim_ready_) return KHS_ERROR_INVALID_STATE;
        if (ring_frame_.empty() || ring_frame_[frame % ring_frame_.size()] != frame)
            return fail(KHS_ERROR_INVALID_ARGUMENT, "The rollback ring does not hold that frame.");
        const size_t slot = frame % ring_frame_.size();
        q_ = ring_q_[slot];
        v_ = ring_v_[slot];
        restored(frame);
        return KHS_OK;
    }

    uint64_t checkpoint_size() const
    {
        if (!built_ || !anim_ready_) return 0;
        return 24 + uint64_t(2 * q_.size() * sizeof(double));
    }

    KhsResult save_checkpoint(void *data, uint64_t capacity)
    {
        // cuda_backend.cu:500-517; header {magic, abi, frame, dof}
        if (!built_) return KHS_ERROR_INVALID_STATE;
        const uint64_t required = checkpoint_size();
        if (!data || !required || capacity < required) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid checkpoint output buffer.");
        unsigned char *out = static_cast<unsigned char *>(data);
        const uint64_t magic = kCheckpointMagic, dof = q_.size();
        const uint32_t abi = KHS_ABI_VERSION, frame = anim_current_;
        std::memcpy(out, &magic, 8);
        std::memcpy(out + 8, &abi, 4);
        std::memcpy(out + 12, &frame, 4);
        std::memcpy(out + 16, &dof, 8);
        std::memcpy(out + 24, q_.data(), q_.size() * 8);
        std::memcpy(out + 24 + q_.size() * 8, v_.data(), v_.size() * 8);
        error.clear();
        return KHS_OK;
    }

    KhsResult restore_checkpoint(const void *data, uint64_t size)
    {
        // cuda_backend.cu:519-537
        if (!built_) return KHS_ERROR_INVALID_STATE;
        const uint64_t required = checkpoint_size();
        if (!data || !required || size != required) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid checkpoint size.");
        const unsigned char *in = static_cast<const unsigned char *>(data);
        uint64_t magic = 0, dof = 0;
        uint32_t abi = 0, frame = 0;
        std::memcpy(&magic, in, 8);
        std::memcpy(&abi, in + 8, 4);
        std::memcpy(&frame, in + 12, 4);
        std::memcpy(&dof, in + 16, 8);
        if (magic != kCheckpointMagic || abi != KHS_ABI_VERSION || dof != q_.size() || frame >= anim_frames_)
            return fail(KHS_ERROR_INVALID_ARGUMENT, "Checkpoint does not match this solver.");
        std::memcpy(q_.data(), in + 24, q_.size() * 8);
        std::memcpy(v_.data(), in + 24 + q_.size() * 8, v_.size() * 8);
        restored(frame);
        return KHS_OK;
    }

    // Bookkeeping of a restored state (checkpoint or rollback slot).
    void restored(uint32_t frame)
    {
        q_old_ = q_;
        snap_q_ = q_;
        snap_v_ = v_;
        anim_current_ = frame;
        cancel_.store(false);
        prog_frame_.store(frame);
        prog_substep_.store(0);
        prog_substep_count_.store(0);
        prog_iteration_.store(0);
        prog_phase_.store(KHS_PHASE_FINISHED);
        step_stats_ = {};
        step_stats_.struct_size = sizeof(step_stats_);
        step_stats_.phase = KHS_PHASE_FINISHED;
        clear_failure();
        error.clear();
    }

    // --------------------------------------------------------------- output
    uint32_t original_count() const { return uint32_t(input_points_.size()); }
    uint32_t node_count() const { return uint32_t(nodes_.size()); }

    KhsResult copy_original(KhsVec3 *out, uint32_t capacity) const
    {
        if (!built_ || !out || capacity < mapping_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        for (size_t i = 0; i < mapping_.size(); ++i) {
            const size_t n = mapping_[i];
            out[i] = {q_[6 * n], q_[6 * n + 1], q_[6 * n + 2]};
        }
        return KHS_OK;
    }

    KhsResult copy_internal(KhsVec3 *out, uint32_t capacity) const
    {
        if (!built_ || !out || capacity < nodes_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        for (size_t n = 0; n < nodes_.size(); ++n) out[n] = {q_[6 * n], q_[6 * n + 1], q_[6 * n + 2]};
        return KHS_OK;
    }

    KhsResult copy_mapping(uint32_t *out, uint32_t capacity) const
    {
        if (!built_ || !out || capacity < mapping_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        std::copy(mapping_.begin(), mapping_.end(), out);
        return KHS_OK;
    }

    KhsResult get_build_stats(KhsBuildStats *s) const
    {
        if (!built_ || !s || s->struct_size < sizeof(KhsBuildStats)) return KHS_ERROR_INVALID_ARGUMENT;
        *s = build_stats_;
        return KHS_OK;
    }

    KhsResult get_step_stats(KhsStepStats *s) const
    {
        if (!s || s->struct_size < sizeof(KhsStepStats)) return KHS_ERROR_INVALID_ARGUMENT;
        *s = step_stats_;
        return KHS_OK;
    }

    KhsResult get_gpu_stats(KhsGpuStats *s) const
    {
        if (!built_ || !s || s->struct_size < sizeof(KhsGpuStats)) return KHS_ERROR_INVALID_ARGUMENT;
        KhsGpuStats g{};
        g.struct_size = sizeof(g);
        g.animation_frame_count = anim_frames_;
        g.resident_bytes = resident_bytes();
        g.last_frame_milliseconds = last_frame_ms_;
        *s = g;
        return KHS_OK;
    }

    KhsResult get_progress(KhsProgress *p) const
    {
        if (!built_ || !p || p->struct_size < sizeof(KhsProgress)) return KHS_ERROR_INVALID_ARGUMENT;
        KhsProgress r{};
        r.struct_size = sizeof(r);
        r.phase = prog_phase_.load();
        r.frame_index = prog_frame_.load();
        r.frame_count = prog_frame_count_.load();
        r.substep = prog_substep_.load();
        r.substep_count = prog_substep_count_.load();
        r.nonlinear_iteration = prog_iteration_.load();
        r.nonlinear_iteration_limit = desc_.newton_iterations;
        r.cancelled = cancel_.load() ? 1u : 0u;
        r.frame_elapsed_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                                                std::chrono::steady_clock::time_point(
                                                                    std::chrono::steady_clock::duration(frame_clock_.load())))
                                      .count();
        *p = r;
        return KHS_OK;
    }

    KhsResult get_failure(KhsFailureDiagnostics *d) const
    {
        if (!built_ || !d || d->struct_size < sizeof(KhsFailureDiagnostics)) return KHS_ERROR_INVALID_ARGUMENT;
        *d = failure_;
        return KHS_OK;
    }

    KhsResult request_cancel()
    {
        if (!built_) return KHS_ERROR_INVALID_STATE;
        cancel_.store(true);
        return KHS_OK;
    }

    // ----------------------------------------------------------- test hooks
    uint64_t dof_count() const { return built_ ? q_.size() : 0; }

    KhsResult get_state(double *q, double *v, uint64_t count) const
    {
        if (!built_ || count != q_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        if (q) std::copy(q_.begin(), q_.end(), q);
        if (v) std::copy(v_.begin(), v_.end(), v);
        return KHS_OK;
    }

    KhsResult set_state(const double *q, const double *q_old, const double *v, uint64_t count)
    {
        if (!built_ || count != q_.size() || !q || !v) return KHS_ERROR_INVALID_ARGUMENT;
        std::copy(q, q + count, q_.begin());
        std::copy(v, v + count, v_.begin());
        if (q_old) std::copy(q_old, q_old + count, q_old_.begin());
        else q_old_ = q_;
        return KHS_OK;
    }

    KhsResult copy_node_model(double *frames9, double *mass, double *rot_mass, uint32_t *flags, uint32_t capacity) const
    {
        if (!built_ || capacity < nodes_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        for (size_t n = 0; n < nodes_.size(); ++n) {
            if (frames9) for (int k = 0; k < 9; ++k) frames9[9 * n + k] = nodes_[n].frame.m[k];
            if (mass) mass[n] = nodes_[n].mass;
            if (rot_mass) rot_mass[n] = nodes_[n].rot_mass;
            if (flags) flags[n] = (nodes_[n].fixed ? 1u : 0u) | (nodes_[n].virt ? 2u : 0u);
        }
        return KHS_OK;
    }

    KhsResult copy_element_model(uint32_t *nodes2, uint32_t *strand, uint32_t *contact, double *rest7,
                                 uint32_t capacity) const
    {
        if (!built_ || capacity < elements_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        for (size_t k = 0; k < elements_.size(); ++k) {
            const Element &e = elements_[k];
            if (nodes2) { nodes2[2 * k] = e.i; nodes2[2 * k + 1] = e.j; }
            if (strand) strand[k] = e.strand;
            if (contact) contact[k] = e.contact ? 1u : 0u;
            if (rest7) {
                rest7[7 * k] = e.rest_length;
                rest7[7 * k + 1] = e.rest_shear.x;
                rest7[7 * k + 2] = e.rest_shear.y;
                rest7[7 * k + 3] = e.rest_shear.z;
                rest7[7 * k + 4] = e.rest_curvature.x;
                rest7[7 * k + 5] = e.rest_curvature.y;
                rest7[7 * k + 6] = e.rest_curvature.z;
            }
        }
        return KHS_OK;
    }

    KhsResult test_evaluate(double dt, double *gradient, double *direction, KhsxEvaluation *out)
    {
        if (!built_ || !(dt > 0.0) || !out || out->struct_size < sizeof(KhsxEvaluation)) return KHS_ERROR_INVALID_ARGUMENT;
        if (has_collider_) bvh_.refit(collider_cur_, nullptr);
        const Stats s = evaluate(true, dt);
        KhsxEvaluation r{};
        r.struct_size = sizeof(r);
        r.feasible = s.feasible ? 1u : 0u;
        r.candidate_pairs = s.candidates;
        r.active_contacts = s.active;
        r.objective = s.objective;
        r.elastic = s.elastic;
        r.contact = s.contact;
        r.kinetic = s.kinetic;
        r.minimum_gap = s.min_gap;
        if (gradient) std::copy(g_.begin(), g_.end(), gradient);
        if (direction && direction_mode == 1) {
            // the forced fallback: -P g with the Jacobi preconditioner, no
            // trust-radius scaling
            preconditioner(dt);
            for (size_t i = 0; i < d_.size(); ++i) d_[i] = (-g_[i]) * P_[i];
            for (size_t n = 0; n < nodes_.size(); ++n)
                if (nodes_[n].fixed) for (int d = 0; d < 6; ++d) d_[6 * n + d] = 0.0;
            std::copy(d_.begin(), d_.end(), direction);
        } else if (direction) {
            std::fill(d_.begin(), d_.end(), 0.0);
            const bool failed = block_solve();
            r.solve_failed = failed ? 1u : 0u;
            std::copy(d_.begin(), d_.end(), direction);
        }
        *out = r;
        return KHS_OK;
    }

    KhsResult set_collider_state(const double *current, const double *previous, uint32_t count)
    {
        if (!built_ || !has_collider_ || !current || count != collider_cur_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        for (uint32_t i = 0; i < count; ++i) {
            collider_cur_[i] = {current[3 * i], current[3 * i + 1], current[3 * i + 2]};
            const double *p = previous ? previous : current;
            collider_old_[i] = {p[3 * i], p[3 * i + 1], p[3 * i + 2]};
        }
        return KHS_OK;
    }

    // The friction stage alone, with the tree on the current boxes as the
    // step leaves it after the sweep (cuda_backend.cu:786-792).
    KhsResult friction_stage(double dt)
    {
        if (!built_ || !(dt > 0.0)) return KHS_ERROR_INVALID_ARGUMENT;
        if (has_collider_) bvh_.refit(collider_cur_, nullptr);
        apply_friction(dt);
        return KHS_OK;
    }

    KhsResult step_bound(const double *direction, uint64_t count, double *alpha)
    {
        if (!built_ || !direction || !alpha || count != q_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        std::copy(direction, direction + count, d_.begin());
        double a = 1.0;
        if (has_collider_) {
            bvh_.refit(collider_cur_, nullptr);
            a = ccd_limit();
        }
        *alpha = a;
        return KHS_OK;
    }

    KhsResult sweep_check(uint32_t *feasible, KhsFailureDiagnostics *diag)
    {
        if (!built_ || !feasible) return KHS_ERROR_INVALID_ARGUMENT;
        if (diag && diag->struct_size < sizeof(KhsFailureDiagnostics)) return KHS_ERROR_INVALID_ARGUMENT;
        SweepFailure sw;
        bool ok = true;
        if (has_collider_) {
            bvh_.refit(collider_cur_, &collider_old_);
            ok = sweep_feasible(sw);
        }
        *feasible = ok ? 1u : 0u;
        if (diag) {
            KhsFailureDiagnostics f{};
            f.struct_size = sizeof(f);
            f.strand_index = f.element_index = f.collider_triangle_index = 0xffffffffu;
            if (!ok) {
                f.kind = KHS_FAILURE_MOVING_COLLIDER_SWEEP;
                f.strand_index = sw.strand;
                f.element_index = sw.element;
                f.collider_triangle_index = sw.triangle;
                f.distance = sw.distance;
                f.required_distance = sw.required;
                f.clearance = sw.clearance;
                f.collider_substep_displacement = sw.displacement;
                f.hair_start = to_c(sw.hair_start);
                f.hair_end = to_c(sw.hair_end);
                f.collider_point = to_c(sw.collider_point);
            }
            *diag = f;
        }
        return KHS_OK;
    }

    double radius_of(uint32_t strand) const { return strand_r_.empty() ? material_.radius : strand_r_[strand]; }
    double young_of(uint32_t strand) const { return strand_E_.empty() ? material_.young_modulus : strand_E_[strand]; }
    double max_radius() const
    {
        double r = material_.radius;
        if (!strand_r_.empty()) {
            r = strand_r_[0];
            for (double x : strand_r_) r = std::fmax(r, x);
        }
        return r;
    }

    // Per-strand Young's modulus and radius; both NULL clears them. Applies
    // at once to a built solver (masses and trust radius) and to later
    // builds; new hair curves clear them.
    KhsResult set_strand_materials(const double *E, const double *r, uint32_t count)
    {
        if (!E && !r) {
            strand_E_.clear();
            strand_r_.clear();
        } else {
            if (!E || !r || !has_hair_ || count != input_offsets_.size() - 1) {
                error = "Per-strand materials need both arrays, one value per hair strand.";
                return KHS_ERROR_INVALID_ARGUMENT;
            }
            for (uint32_t k = 0; k < count; ++k)
                if (!(E[k] > 0.0 && std::isfinite(E[k]) && r[k] > 0.0 && std::isfinite(r[k]))) {
                    error = "Strand " + std::to_string(k) + " has a material value out of range.";
                    return KHS_ERROR_INVALID_ARGUMENT;
                }
            strand_E_.assign(E, E + count);
            strand_r_.assign(r, r + count);
        }
        if (built_) {
            init_masses();
            trust_radius_ = std::fmax(0.25 * max_rest_length_,
                                      8.0 * (max_radius() + material_.collider_offset + material_.barrier_distance));
        }
        return KHS_OK;
    }

    KhsResult direction_counts(uint64_t *block, uint64_t *steepest) const
    {
        if (!block || !steepest) return KHS_ERROR_INVALID_ARGUMENT;
        *block = block_directions_;
        *steepest = steepest_directions_;
        return KHS_OK;
    }

    KhsResult copy_blocks(double *diagonal, double *upper, uint32_t capacity) const
    {
        if (!built_ || capacity < nodes_.size()) return KHS_ERROR_INVALID_ARGUMENT;
        if (diagonal) std::copy(D_.begin(), D_.end(), diagonal);
        if (upper) std::copy(U_.begin(), U_.end(), upper);
        return KHS_OK;
    }

    KhsResult broadphase_pairs(uint32_t *pairs, uint64_t capacity, uint64_t *count)
    {
        if (!built_ || !count) return KHS_ERROR_INVALID_ARGUMENT;
        uint64_t total = 0;
        if (has_collider_) {
            bvh_.refit(collider_cur_, nullptr);
            std::vector<uint32_t> cand;
            for (uint32_t k = 0; k < elements_.size(); ++k) {
                const Element &e = elements_[k];
                if (nodes_[e.i].fixed || !e.contact) continue;
                const double pad = radius_of(e.strand) + material_.collider_offset + material_.barrier_distance;
                const V3 p0 = pos(e.i), p1 = pos(e.j);
                bvh_.query(vmin(p0, p1), vmax(p0, p1), pad, cand);
                for (uint32_t t : cand) {
                    if (pairs && total < capacity) { pairs[2 * total] = k; pairs[2 * total + 1] = t; }
                    ++total;
                }
            }
        }
        *count = total;
        return (pairs && total > capacity) ? KHS_ERROR_INVALID_ARGUMENT : KHS_OK;
    }

    KhsStepStats &stats_ref() { return step_stats_; }
    void set_phase_failed() { step_stats_.phase = KHS_PHASE_FAILED; }

private:
    // ----------------------------------------------------- model building
    void clear_built()
    {
        built_ = false;
        anim_ready_ = false;
        nodes_.clear();
        elements_.clear();
        strands_.clear();
        mapping_.clear();
    }

    bool make_internal_mesh()
    {
        // solver.cpp:227-318
        nodes_.clear();
        elements_.clear();
        const uint32_t strand_count = uint32_t(input_offsets_.size() - 1);
        std::vector<std::vector<uint32_t>> strand_nodes(strand_count);
        mapping_.assign(input_points_.size(), 0);
        const double hmax = desc_.maximum_element_length, zero = 1.0e-12;
        merged_ = ext_strands_ = ext_nodes_ = 0;
        ext_length_ = 0.0;
        for (uint32_t s = 0; s < strand_count; ++s) {
            const uint32_t begin = input_offsets_[s], end = input_offsets_[s + 1];
            Node first;
            first.x = input_points_[begin];
            first.strand = s;
            first.bind_left = first.bind_right = begin;
            first.bind_t = 0.0;
            nodes_.push_back(first);
            strand_nodes[s].push_back(uint32_t(nodes_.size() - 1));
            mapping_[begin] = uint32_t(nodes_.size() - 1);
            V3 previous = input_points_[begin];
            uint32_t unique = 0, term_left = begin, term_right = begin;
            double rest = 0.0, term_len = 0.0;
            V3 term_dir;
            for (uint32_t p = begin + 1; p < end; ++p) {
                const V3 target = input_points_[p];
                const double length = norm(target - previous);
                if (length <= zero) {
                    mapping_[p] = strand_nodes[s].back();
                    ++merged_;
                    previous = target;
                    continue;
                }
                const uint32_t divisions = std::max(1u, uint32_t(std::ceil(length / hmax)));
                for (uint32_t d = 1; d <= divisions; ++d) {
                    const double t = double(d) / divisions;
                    Node node;
                    node.x = (1.0 - t) * previous + t * target;
                    node.strand = s;
                    node.bind_left = p - 1;
                    node.bind_right = p;
                    node.bind_t = t;
                    nodes_.push_back(node);
                    strand_nodes[s].push_back(uint32_t(nodes_.size() - 1));
                }
                mapping_[p] = strand_nodes[s].back();
                previous = target;
                rest += length;
                term_left = p - 1;
                term_right = p;
                term_len = length;
                term_dir = div_each(target - input_points_[p - 1], length);
                ++unique;
            }
            if (unique == 0 || strand_nodes[s].size() < 2) {
                error = "Strand " + std::to_string(s) + " has all points at one position.";
                return false;
            }
            if (desc_.minimum_dynamic_length > rest + zero) {
                const double ext = desc_.minimum_dynamic_length - rest;
                const uint32_t divisions = std::max(1u, uint32_t(std::ceil(ext / hmax)));
                const V3 tip = input_points_[end - 1];
                for (uint32_t d = 1; d <= divisions; ++d) {
                    const double distance = ext * double(d) / divisions;
                    Node node;
                    node.x = tip + distance * term_dir;
                    node.strand = s;
                    node.virt = true;
                    node.bind_left = term_left;
                    node.bind_right = term_right;
                    node.bind_t = 1.0 + distance / term_len;
                    nodes_.push_back(node);
                    strand_nodes[s].push_back(uint32_t(nodes_.size() - 1));
                }
                ++ext_strands_;
                ext_nodes_ += divisions;
                ext_length_ += ext;
            }
            const uint32_t fixed = std::min<uint32_t>(desc_.fixed_root_nodes, uint32_t(strand_nodes[s].size()));
            for (uint32_t k = 0; k < fixed; ++k) nodes_[strand_nodes[s][k]].fixed = true;
        }
        strands_.resize(strand_count);
        for (uint32_t s = 0; s < strand_count; ++s) {
            uint32_t fixed = 0;
            while (fixed < strand_nodes[s].size() && nodes_[strand_nodes[s][fixed]].fixed) ++fixed;
            strands_[s] = {strand_nodes[s].front(), uint32_t(strand_nodes[s].size()), fixed};
        }
        return true;
    }

    void init_frames_and_elements()
    {
        // solver.cpp:320-356
        elements_.clear();
        for (uint32_t s = 0; s < strands_.size(); ++s) {
            const uint32_t first = strands_[s].first, count = strands_[s].count;
            std::vector<V3> tangents(count);
            for (uint32_t k = 0; k < count; ++k) {
                V3 t;
                if (k == 0) t = nodes_[first + 1].x - nodes_[first].x;
                else if (k + 1 == count) t = nodes_[first + k].x - nodes_[first + k - 1].x;
                else t = nodes_[first + k + 1].x - nodes_[first + k - 1].x;
                tangents[k] = normalized(t);
            }
            M3 frame = frame_from_tangent(tangents[0], v3(1, 0, 0));
            nodes_[first].frame = frame;
            for (uint32_t k = 1; k < count; ++k) {
                const V3 d1 = parallel_transport(frame.col(0), tangents[k - 1], tangents[k]);
                frame = frame_from_tangent(tangents[k], d1);
                nodes_[first + k].frame = frame;
            }
            for (uint32_t k = 0; k + 1 < count; ++k) {
                Element e;
                e.i = first + k;
                e.j = first + k + 1;
                e.strand = s;
                e.contact = !nodes_[e.j].virt;
                e.rest_length = norm(nodes_[e.j].x - nodes_[e.i].x);
                const M3 &ri = nodes_[e.i].frame, &rj = nodes_[e.j].frame;
                const V3 tangent = div_each(nodes_[e.j].x - nodes_[e.i].x, e.rest_length);
                const M3 rit = transpose(ri), rjt = transpose(rj);
                M3 sum;
                for (int k2 = 0; k2 < 9; ++k2) sum.m[k2] = rit.m[k2] + rjt.m[k2];
                e.rest_shear = 0.5 * matvec(sum, tangent);
                e.rest_curvature = div_each(cayley_rotation_vector(matmul(rit, rj)), e.rest_length);
                elements_.push_back(e);
            }
        }
        max_rest_length_ = 0.0;
        for (const Element &e : elements_) max_rest_length_ = std::fmax(max_rest_length_, e.rest_length);
    }

    void init_masses()
    {
        // solver.cpp:358-373
        const double pi = 3.14159265358979323846;
        for (Node &n : nodes_) { n.mass = 0.0; n.rot_mass = 0.0; }
        for (const Element &e : elements_) {
            const double r = radius_of(e.strand);
            const double area = pi * r * r;
            const double inertia = pi * std::pow(r, 4.0) / 4.0;
            const double polar = 2.0 * inertia;
            const double mass = material_.density * area * e.rest_length;
            const double rot = material_.density * (2.0 * inertia + polar) / 3.0 * e.rest_length;
            nodes_[e.i].mass += 0.5 * mass;
            nodes_[e.j].mass += 0.5 * mass;
            nodes_[e.i].rot_mass += 0.5 * rot;
            nodes_[e.j].rot_mass += 0.5 * rot;
        }
    }

    bool root_targets(const std::vector<V3> &points, std::vector<V3> &positions, std::vector<V3> &rotations) const
    {
        // solver.cpp:416-447
        if (points.size() != input_points_.size()) return false;
        positions.resize(nodes_.size());
        rotations.resize(nodes_.size());
        for (size_t n = 0; n < nodes_.size(); ++n) {
            const Node &node = nodes_[n];
            positions[n] = (1.0 - node.bind_t) * points[node.bind_left] + node.bind_t * points[node.bind_right];
            rotations[n] = node.rotation;
        }
        for (const Strand &s : strands_) {
            std::vector<V3> tangents(s.count);
            for (uint32_t k = 0; k < s.count; ++k) {
                V3 t;
                if (k == 0) t = positions[s.first + 1] - positions[s.first];
                else if (k + 1 == s.count) t = positions[s.first + k] - positions[s.first + k - 1];
                else t = positions[s.first + k + 1] - positions[s.first + k - 1];
                if (norm2(t) < 1.0e-24) return false;
                tangents[k] = normalized(t);
            }
            M3 frame = frame_from_tangent(tangents[0], nodes_[s.first].frame.col(0));
            rotations[s.first] = rotation_vector(matmul(frame, transpose(nodes_[s.first].frame)));
            for (uint32_t k = 1; k < s.count; ++k) {
                const V3 d1 = parallel_transport(frame.col(0), tangents[k - 1], tangents[k]);
                frame = frame_from_tangent(tangents[k], d1);
                rotations[s.first + k] = rotation_vector(matmul(frame, transpose(nodes_[s.first + k].frame)));
            }
        }
        return true;
    }

    void validate_topology()
    {
        // solver.cpp:381-414
        boundary_edges_ = nonmanifold_edges_ = inconsistent_edges_ = inverted_ = 0;
        struct Use { uint32_t count = 0; int dir = 0; };
        std::unordered_map<uint64_t, Use> uses;
        uses.reserve(tris_.size() * 3);
        for (const Tri &t : tris_) {
            const uint32_t e[3][2] = {{t.i0, t.i1}, {t.i1, t.i2}, {t.i2, t.i0}};
            for (auto &ed : e) {
                Use &u = uses[edge_key(ed[0], ed[1])];
                ++u.count;
                u.dir += ed[0] < ed[1] ? 1 : -1;
            }
        }
        for (const auto &kv : uses) {
            if (kv.second.count == 1) ++boundary_edges_;
            else if (kv.second.count > 2) ++nonmanifold_edges_;
            else if (kv.second.dir != 0) ++inconsistent_edges_;
        }
        closed_ = boundary_edges_ == 0 && nonmanifold_edges_ == 0 && inconsistent_edges_ == 0;
        if (closed_) {
            double six_volume = 0.0;
            for (const Tri &t : tris_)
                six_volume += dot(collider_input_[t.i0], cross(collider_input_[t.i1], collider_input_[t.i2]));
            if (six_volume < 0.0) inverted_ = 1;
        }
    }

    bool point_inside(V3 p) const
    {
        // solver.cpp:608-632: parity of ray crossings along a skewed +x ray.
        // Only triangles whose box meets the box of the ray inside the
        // collider's x range can be hit; the count equals the brute-force one.
        uint32_t hits = 0;
        const V3 dir{1.0, 0.0000012345, 0.0000023456};
        const double reach = std::fmax(collider_x_max_ - p.x, 0.0) + 1.0e-6;
        const V3 end = p + dir * reach;
        std::vector<uint32_t> cand;
        bvh_.query(vmin(p, end), vmax(p, end), 1.0e-9, cand);
        for (uint32_t ti : cand) {
            const Tri &t = tris_[ti];
            const V3 a = collider_cur_[t.i0], b = collider_cur_[t.i1], c = collider_cur_[t.i2];
            const V3 e1 = b - a, e2 = c - a, h = cross(dir, e2);
            const double det = dot(e1, h);
            if (std::fabs(det) < 1.0e-14) continue;
            const double inv = 1.0 / det;
            const V3 s = p - a;
            const double u = inv * dot(s, h);
            if (u < 0.0 || u > 1.0) continue;
            const V3 q = cross(s, e1);
            const double v = inv * dot(dir, q);
            if (v < 0.0 || u + v > 1.0) continue;
            if (inv * dot(e2, q) > 1.0e-12) ++hits;
        }
        return (hits & 1u) != 0;
    }

  public:
    // Brute-force ray parity over every triangle (checks the culled test).
    int point_inside_both(const double *pt) const
    {
        const V3 p{pt[0], pt[1], pt[2]};
        uint32_t hits = 0;
        const V3 dir{1.0, 0.0000012345, 0.0000023456};
        for (const Tri &t : tris_) {
            const V3 a = collider_cur_[t.i0], b = collider_cur_[t.i1], c = collider_cur_[t.i2];
            const V3 e1 = b - a, e2 = c - a, h = cross(dir, e2);
            const double det = dot(e1, h);
            if (std::fabs(det) < 1.0e-14) continue;
            const double inv = 1.0 / det;
            const V3 s = p - a;
            const double u = inv * dot(s, h);
            if (u < 0.0 || u > 1.0) continue;
            const V3 q = cross(s, e1);
            const double v = inv * dot(dir, q);
            if (v < 0.0 || u + v > 1.0) continue;
            if (inv * dot(e2, q) > 1.0e-12) ++hits;
        }
        return int(hits & 1u) | (point_inside(p) ? 2 : 0);
    }

  private:
    bool initial_feasibility()
    {
        // solver.cpp:634-657
        if (closed_) {
            for (uint32_t n = 0; n < nodes_.size(); ++n) {
                if (nodes_[n].fixed || nodes_[n].virt) continue;
                if (point_inside(nodes_[n].x)) {
                    error = "Initial intersection: internal node " + std::to_string(n) + " is inside the closed collider.";
                    return false;
                }
            }
        }
        uint64_t candidates = 0;
        const double gap = current_minimum_gap(&candidates);
        if (gap <= desc_.minimum_gap) {
            std::ostringstream ss;
            ss << "Initial intersection or negative gap: minimum gap " << gap << " m, " << candidates << " candidates.";
            error = ss.str();
            return false;
        }
        return true;
    }

    double current_minimum_gap(uint64_t *count_out) const
    {
        // solver.cpp:659-681 (host segment/triangle distance with the
        // intersection test)
        double minimum = kInf;
        uint64_t count = 0;
        std::vector<uint32_t> cand;
        for (const Element &e : elements_) {
            if (nodes_[e.i].fixed || !e.contact) continue;
            const double pad = radius_of(e.strand) + material_.collider_offset + material_.barrier_distance;
            bvh_.query(vmin(nodes_[e.i].x, nodes_[e.j].x), vmax(nodes_[e.i].x, nodes_[e.j].x), pad, cand);
            for (uint32_t t : cand) {
                const Tri &tr = tris_[t];
                const Closest c = closest_segment_triangle(nodes_[e.i].x, nodes_[e.j].x, collider_cur_[tr.i0],
                                                           collider_cur_[tr.i1], collider_cur_[tr.i2], true);
                minimum = std::min(minimum, c.distance - radius_of(e.strand) - material_.collider_offset);
                ++count;
            }
        }
        if (count_out) *count_out = count;
        return minimum;
    }

    void init_state()
    {
        // cuda_backend.cu:341-407: q from node positions and (zero)
        // rotations; root targets start at the node positions.
        const size_t n = nodes_.size();
        q_.assign(6 * n, 0.0);
        v_.assign(6 * n, 0.0);
        root_cur_.resize(n);
        rot_cur_.resize(n);
        for (size_t i = 0; i < n; ++i) {
            q_[6 * i] = nodes_[i].x.x;
            q_[6 * i + 1] = nodes_[i].x.y;
            q_[6 * i + 2] = nodes_[i].x.z;
            q_[6 * i + 3] = nodes_[i].rotation.x;
            q_[6 * i + 4] = nodes_[i].rotation.y;
            q_[6 * i + 5] = nodes_[i].rotation.z;
            root_cur_[i] = nodes_[i].x;
            rot_cur_[i] = nodes_[i].rotation;
        }
        root_pending_ = root_cur_;
        rot_pending_ = rot_cur_;
        q_old_ = q_;
        snap_q_ = q_;
        snap_v_ = v_;
        g_.assign(6 * n, 0.0);
        d_.assign(6 * n, 0.0);
        P_.assign(6 * n, 0.0);
        q_base_.assign(6 * n, 0.0);
        dv_.assign(6 * n, 0.0);
        D_.assign(36 * n, 0.0);
        U_.assign(36 * n, 0.0);
        C_.assign(36 * n, 0.0);
        dp_.assign(6 * n, 0.0);
        collider_cur_ = collider_input_;
        collider_old_ = collider_input_;
        collider_frame_start_ = collider_input_;
        collider_pending_ = collider_input_;
        trust_radius_ = std::fmax(0.25 * max_rest_length_,
                                  8.0 * (max_radius() + material_.collider_offset + material_.barrier_distance));
        anim_ready_ = false;
        anim_frames_ = 0;
        anim_current_ = 0;
        ring_frame_.clear();
        ring_q_.clear();
        ring_v_.clear();
        clear_failure();
    }

    uint64_t resident_bytes() const
    {
        const uint64_t n = nodes_.size();
        return n * (6 * 8 * 10 + 36 * 8 * 3 + 9 * 8 + 2 * 8) + elements_.size() * 64 +
               uint64_t(collider_input_.size()) * 24 * 4 + tris_.size() * 12 +
               (anim_root_.size() + anim_rot_.size() + anim_collider_.size()) * 24;
    }

    // ------------------------------------------------------------ stepping
    V3 pos(uint32_t n) const { return {q_[6 * n], q_[6 * n + 1], q_[6 * n + 2]}; }

    // Energy of node n: implicit-Euler inertia around q_old + dt v + g dt^2
    // plus mass-proportional damping, with gradient and diagonal block
    // (cuda_backend.cu:883-897).
    void node_term(uint32_t n, double dt, bool grad, Stats &st)
    {
        if (grad) for (int k = 0; k < 36; ++k) D_[36 * n + k] = 0.0;
        const Node &node = nodes_[n];
        if (node.fixed) return;
        const double inv = 1.0 / dt, inv2 = inv * inv;
        const double kx = node.mass * inv2, kr = node.rot_mass * inv2;
        const double cx = node.mass * material_.mass_damping * inv, cr = node.rot_mass * material_.mass_damping * inv;
        double objective = 0, kinetic = 0;
        const double gravity[3] = {desc_.gravity.x, desc_.gravity.y, desc_.gravity.z};
        for (int d = 0; d < 6; ++d) {
            const double gr = d < 3 ? gravity[d] * dt * dt : 0.0;
            const double predicted = q_old_[6 * n + d] + dt * v_[6 * n + d] + gr;
            const double delta = q_[6 * n + d] - predicted;
            const double damp = q_[6 * n + d] - q_old_[6 * n + d];
            const double k = d < 3 ? kx : kr, c = d < 3 ? cx : cr;
            objective += 0.5 * k * delta * delta + 0.5 * c * damp * damp;
            if (d < 3) kinetic += 0.5 * node.mass * damp * damp * inv2;
            if (grad) {
                g_[6 * n + d] = k * delta + c * damp;
                D_[36 * n + 6 * d + d] = k + c;
            }
        }
        st.objective += objective;
        st.kinetic += kinetic;
    }

    // Closest candidate triangle of a segment under the padding.
    bool closest_triangle(V3 p0, V3 p1, double pad, std::vector<uint32_t> &cand, Closest &best, uint32_t &best_t,
                          uint64_t *count) const
    {
        bvh_.query(vmin(p0, p1), vmax(p0, p1), pad, cand);
        best = Closest{};
        best_t = 0xffffffffu;
        for (uint32_t t : cand) {
            if (count) ++*count;
            const Tri &tr = tris_[t];
            const Closest c = closest_segment_triangle(p0, p1, collider_cur_[tr.i0], collider_cur_[tr.i1],
                                                       collider_cur_[tr.i2], false);
            if (c.distance < best.distance) {
                best = c;
                best_t = t;
            }
        }
        return best_t != 0xffffffffu;
    }

    // Elastic energy of an element (Dual<12> strains, Gauss-Newton blocks
    // L0 sum_s w_s J_s^T J_s, each term added by a fused multiply-add as
    // the reference's CUDA build contracts it) and its log-barrier contact
    // against the closest collider triangle (cuda_backend.cu:899-939).
    void element_term(uint32_t k, bool grad, std::vector<uint32_t> &cand, Stats &st)
    {
        const Element &e = elements_[k];
        const Node &ni = nodes_[e.i], &nj = nodes_[e.j];
        Dual<12> dq[12];
        for (int d = 0; d < 6; ++d) {
            dq[d] = Dual<12>::var(q_[6 * e.i + d], d);
            dq[6 + d] = Dual<12>::var(q_[6 * e.j + d], 6 + d);
        }
        Dual<12> strain[6];
        rod_strains(dq, ni.frame, nj.frame, e, strain);
        const double r = radius_of(e.strand), E = young_of(e.strand);
        const double area = kPi * r * r, inertia = kPi * r * r * r * r / 4.0;
        const double G = E / (2.0 * (1.0 + material_.poisson_ratio));
        const double w[6] = {material_.shear_correction * G * area, material_.shear_correction * G * area,
                             E * area, E * inertia, E * inertia, 2.0 * G * inertia};
        Dual<12> re(0.0);
        for (int s = 0; s < 6; ++s) re = re + Dual<12>(0.5 * e.rest_length * w[s]) * strain[s] * strain[s];
        st.objective += re.v;
        st.elastic += re.v;
        if (grad) {
            if (!ni.fixed) for (int d = 0; d < 6; ++d) g_[6 * e.i + d] += re.d[d];
            if (!nj.fixed) for (int d = 0; d < 6; ++d) g_[6 * e.j + d] += re.d[6 + d];
            for (int a = 0; a < 12; ++a)
                for (int b = 0; b < 12; ++b) {
                    double h = 0.0;
                    for (int s = 0; s < 6; ++s) h = std::fma(e.rest_length * w[s] * strain[s].d[a], strain[s].d[b], h);
                    if (a < 6 && b < 6) { if (!ni.fixed) D_[36 * e.i + 6 * a + b] += h; }
                    else if (a >= 6 && b >= 6) { if (!nj.fixed) D_[36 * e.j + 6 * (a - 6) + (b - 6)] += h; }
                    else if (a < 6 && b >= 6 && !ni.fixed && !nj.fixed) U_[36 * e.i + 6 * a + (b - 6)] += h;
                }
        }
        if (!has_collider_ || ni.fixed || !e.contact) return;
        const V3 p0 = pos(e.i), p1 = pos(e.j);
        const double pad = r + material_.collider_offset + material_.barrier_distance;
        Closest pair;
        uint32_t ti = 0;
        if (!closest_triangle(p0, p1, pad, cand, pair, ti, &st.candidates)) return;
        const Tri &t = tris_[ti];
        const double gap = pair.distance - r - material_.collider_offset;
        st.min_gap = std::fmin(st.min_gap, gap);
        if (gap <= desc_.minimum_gap) { st.feasible = false; return; }
        if (gap >= material_.barrier_distance) return;
        ++st.active;
        // Barrier b(g) = -kappa (dhat - g)^2 ln(g / dhat) of the distance
        // between the hair point at parameter t and the fixed barycentric
        // point (cuda_backend.cu:255-265).
        Dual<6> cq[6];
        for (int d = 0; d < 3; ++d) {
            cq[d] = Dual<6>::var(q_[6 * e.i + d], d);
            cq[3 + d] = Dual<6>::var(q_[6 * e.j + d], 3 + d);
        }
        const V3 A = collider_cur_[t.i0], B = collider_cur_[t.i1], C = collider_cur_[t.i2];
        typedef Dual<6> D6;
        const D6 tt(pair.t);
        const TV3<D6> x0{cq[0], cq[1], cq[2]}, x1{cq[3], cq[4], cq[5]};
        const TV3<D6> p = tadd(tscale(x0, D6(1.0) - tt), tscale(x1, tt));
        const V3 bp = A * pair.bary.x + B * pair.bary.y + C * pair.bary.z;
        const TV3<D6> diff = tsub(p, TV3<D6>{D6(bp.x), D6(bp.y), D6(bp.z)});
        const D6 dist = dsqrt(tdot(diff, diff) + D6(1e-30));
        const D6 dgap = dist - D6(r + material_.collider_offset);
        D6 ce(0.0);
        if (dgap.v >= material_.barrier_distance) {
            ce = cq[0] * D6(0.0);
        } else {
            const D6 aa = D6(material_.barrier_distance) - dgap;
            const D6 lg = dlog(dgap / D6(material_.barrier_distance));
            ce = -D6(material_.contact_stiffness) * aa * aa * lg;
        }
        if (!std::isfinite(ce.v)) { st.feasible = false; return; }
        st.objective += ce.v;
        st.contact += ce.v;
        if (grad) {
            for (int d = 0; d < 3; ++d) g_[6 * e.i + d] += ce.d[d];
            for (int d = 0; d < 3; ++d) g_[6 * e.j + d] += ce.d[3 + d];
            // Rank-one curvature b''(g) grad(g) grad(g)^T with the closest
            // point held fixed.
            const V3 normal = div_recip(p0 * (1.0 - pair.t) + p1 * pair.t - bp, std::fmax(pair.distance, 1e-15));
            const double aa = material_.barrier_distance - gap;
            const double curv = material_.contact_stiffness *
                                (-2.0 * std::log(gap / material_.barrier_distance) + 4.0 * aa / gap + aa * aa / (gap * gap));
            const double dg[6] = {normal.x * (1.0 - pair.t), normal.y * (1.0 - pair.t), normal.z * (1.0 - pair.t),
                                  normal.x * pair.t, normal.y * pair.t, normal.z * pair.t};
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) {
                    const double h = curv * dg[a] * dg[b];
                    if (a < 3 && b < 3) D_[36 * e.i + 6 * a + b] += h;
                    else if (a >= 3 && b >= 3) D_[36 * e.j + 6 * (a - 3) + (b - 3)] += h;
                    else if (a < 3 && b >= 3) U_[36 * e.i + 6 * a + (b - 3)] += h;
                }
        }
    }

    // One evaluation of the incremental potential; with `grad` also the
    // gradient and the block-tridiagonal Gauss-Newton matrix
    // (cuda_backend.cu:604-613).
    Stats evaluate(bool grad, double dt)
    {
        std::vector<Stats> partial(strands_.size());
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            Stats acc;
            std::vector<uint32_t> cand;
            if (grad) {
                for (uint32_t n = st.first; n < st.first + st.count; ++n) {
                    for (int d = 0; d < 6; ++d) g_[6 * n + d] = 0.0;
                    for (int k = 0; k < 36; ++k) U_[36 * n + k] = 0.0;
                }
            }
            for (uint32_t n = st.first; n < st.first + st.count; ++n) node_term(n, dt, grad, acc);
            const uint32_t e0 = st.first - s;  // elements of strand s start here
            for (uint32_t k = e0; k + 1 < e0 + st.count; ++k) element_term(k, grad, cand, acc);
            partial[s] = acc;
        });
        Stats total;
        for (const Stats &p : partial) total.add(p);
        return total;
    }

    // Per-strand block Thomas on the free nodes; the root Dirichlet nodes
    // are eliminated (cuda_backend.cu:866-881). Returns true on failure.
    // Mixed precision: the same block Thomas with the Gauss-Newton blocks,
    // the gradient and every factor rounded to FP32 (products summed by
    // single-precision fused multiply-adds); the direction is widened to
    // FP64. The Newton loop around it keeps its FP64 gradient, energy,
    // barrier, step bound and convergence test, so each iteration corrects
    // the FP32 direction's error from an FP64 residual.
    bool block_solve_fp32()
    {
        std::vector<uint8_t> failed(strands_.size(), 0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const uint32_t begin = st.first + st.fixed_count, end = st.first + st.count;
            if (begin >= end) return;
            std::vector<float> C(36 * size_t(end - begin)), dp(6 * size_t(end - begin));
            for (uint32_t n = begin; n < end; ++n) {
                const size_t k = n - begin;
                float b[36], inv[36], rhs[6];
                for (int i = 0; i < 36; ++i) b[i] = float(D_[36 * n + i]);
                for (int i = 0; i < 6; ++i) rhs[i] = -float(g_[6 * n + i]);
                if (n > begin) {
                    const float *pc = &C[36 * (k - 1)], *pd = &dp[6 * (k - 1)];
                    for (int r = 0; r < 6; ++r) {
                        for (int c = 0; c < 6; ++c) {
                            float sum = 0;
                            for (int j = 0; j < 6; ++j) sum = std::fma(float(U_[36 * (n - 1) + 6 * j + r]), pc[6 * j + c], sum);
                            b[6 * r + c] -= sum;
                        }
                        float sum = 0;
                        for (int j = 0; j < 6; ++j) sum = std::fma(float(U_[36 * (n - 1) + 6 * j + r]), pd[j], sum);
                        rhs[r] -= sum;
                    }
                }
                if (!inverse6f(b, inv)) { failed[s] = 1; return; }
                float *cc = &C[36 * k], *cd = &dp[6 * k];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) {
                        float sum = 0;
                        if (n + 1 < end)
                            for (int j = 0; j < 6; ++j) sum = std::fma(inv[6 * r + j], float(U_[36 * n + 6 * j + c]), sum);
                        cc[6 * r + c] = sum;
                    }
                for (int r = 0; r < 6; ++r) {
                    float sum = 0;
                    for (int j = 0; j < 6; ++j) sum = std::fma(inv[6 * r + j], rhs[j], sum);
                    cd[r] = sum;
                }
            }
            std::vector<float> x(6 * size_t(end - begin));
            for (uint32_t n = end; n-- > begin;) {
                const size_t k = n - begin;
                for (int r = 0; r < 6; ++r) {
                    float value = dp[6 * k + r];
                    if (n + 1 < end)
                        for (int j = 0; j < 6; ++j) value = std::fma(-C[36 * k + 6 * r + j], x[6 * (k + 1) + j], value);
                    x[6 * k + r] = value;
                    d_[6 * n + r] = double(value);
                }
            }
        });
        for (uint8_t f : failed) if (f) return true;
        return false;
    }

    // The same block-tridiagonal system solved from both ends ("twisted"
    // block factorization): block Thomas eliminates the free nodes from the
    // root up to the middle node m and, on the reversed chain (coupling
    // blocks transposed), from the tip down to m; the middle node's Schur
    // complement takes both sides' terms (the root side's first), its
    // solution starts both back substitutions. Mathematically the same
    // direction as block_solve's; the serial chain is half as long. Each
    // entry follows the sequential operations of the GPU kernel
    // (k_block_solve_twisted): sums of products by fused multiply-adds from
    // zero, the same inverse6.
    bool block_solve_twisted()
    {
        std::vector<uint8_t> failed(strands_.size(), 0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const uint32_t begin = st.first + st.fixed_count, end = st.first + st.count;
            if (begin >= end) return;
            const uint32_t len = end - begin, na = (len - 1) / 2, m = begin + na;
            double b[36], inv[36], rhs[6];
            auto load = [&](uint32_t n) {
                for (int i = 0; i < 36; ++i) b[i] = D_[36 * n + i];
                for (int i = 0; i < 6; ++i) rhs[i] = -g_[6 * n + i];
            };
            // b -= L C and rhs -= L d with L(r, j) = lower(r, j)
            auto schur = [&](auto lower, const double *pc, const double *pd) {
                for (int r = 0; r < 6; ++r) {
                    for (int c = 0; c < 6; ++c) {
                        double sum = 0;
                        for (int j = 0; j < 6; ++j) sum = std::fma(lower(r, j), pc[6 * j + c], sum);
                        b[6 * r + c] -= sum;
                    }
                    double sum = 0;
                    for (int j = 0; j < 6; ++j) sum = std::fma(lower(r, j), pd[j], sum);
                    rhs[r] -= sum;
                }
            };
            // C = inv W with W(j, c) = w(j, c); d = inv rhs
            auto factor = [&](uint32_t n, auto w) {
                double *cc = &C_[36 * n], *cd = &dp_[6 * n];
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c) {
                        double sum = 0;
                        for (int j = 0; j < 6; ++j) sum = std::fma(inv[6 * r + j], w(j, c), sum);
                        cc[6 * r + c] = sum;
                    }
                for (int r = 0; r < 6; ++r) {
                    double sum = 0;
                    for (int j = 0; j < 6; ++j) sum = std::fma(inv[6 * r + j], rhs[j], sum);
                    cd[r] = sum;
                }
            };
            // root side: nodes begin .. m - 1, coupling U_{n-1}^T C'_{n-1}
            for (uint32_t n = begin; n < m; ++n) {
                load(n);
                if (n > begin) {
                    const double *U = &U_[36 * (n - 1)];
                    schur([&](int r, int j) { return U[6 * j + r]; }, &C_[36 * (n - 1)], &dp_[6 * (n - 1)]);
                }
                if (!inverse6(b, inv)) { failed[s] = 1; return; }
                const double *Un = &U_[36 * n];
                factor(n, [&](int j, int c) { return Un[6 * j + c]; });
            }
            // tip side: nodes end - 1 .. m + 1, coupling U_n C''_{n+1}; its
            // factor is inv U_{n-1}^T
            for (uint32_t n = end; n-- > m + 1;) {
                load(n);
                if (n + 1 < end) {
                    const double *U = &U_[36 * n];
                    schur([&](int r, int j) { return U[6 * r + j]; }, &C_[36 * (n + 1)], &dp_[6 * (n + 1)]);
                }
                if (!inverse6(b, inv)) { failed[s] = 1; return; }
                const double *Up = &U_[36 * (n - 1)];
                factor(n, [&](int j, int c) { return Up[6 * c + j]; });
            }
            // middle node m: both sides' Schur terms, the root side's first
            load(m);
            if (m > begin) {
                const double *U = &U_[36 * (m - 1)];
                schur([&](int r, int j) { return U[6 * j + r]; }, &C_[36 * (m - 1)], &dp_[6 * (m - 1)]);
            }
            if (m + 1 < end) {
                const double *U = &U_[36 * m];
                schur([&](int r, int j) { return U[6 * r + j]; }, &C_[36 * (m + 1)], &dp_[6 * (m + 1)]);
            }
            if (!inverse6(b, inv)) { failed[s] = 1; return; }
            for (int r = 0; r < 6; ++r) {
                double sum = 0;
                for (int j = 0; j < 6; ++j) sum = std::fma(inv[6 * r + j], rhs[j], sum);
                d_[6 * m + r] = sum;
            }
            // back substitution outward from m
            for (uint32_t n = m; n-- > begin;)
                for (int r = 0; r < 6; ++r) {
                    double value = dp_[6 * n + r];
                    for (int k = 0; k < 6; ++k) value = std::fma(-C_[36 * n + 6 * r + k], d_[6 * (n + 1) + k], value);
                    d_[6 * n + r] = value;
                }
            for (uint32_t n = m + 1; n < end; ++n)
                for (int r = 0; r < 6; ++r) {
                    double value = dp_[6 * n + r];
                    for (int k = 0; k < 6; ++k) value = std::fma(-C_[36 * n + 6 * r + k], d_[6 * (n - 1) + k], value);
                    d_[6 * n + r] = value;
                }
        });
        for (uint8_t f : failed) if (f) return true;
        return false;
    }

    // Block Thomas whose factors come straight out of one elimination
    // (k_block_solve_fused): per node, S = D_n - U_{n-1}^T C'_{n-1} and
    // e = g_n - U_{n-1}^T e_{n-1} (e = -d'), then Gauss-Jordan without
    // pivoting on the columns of [S | U_n | e] turns them into
    // [I | C'_n | e_n]; no inverse is formed. The diagonal shift is the
    // reference's, max(1e-12 max|S_ii|, 1e-24), added to each pivot as it is
    // taken. The blocks are symmetric positive definite (Gauss-Newton
    // Hessian, mass and shift), whose Schur complements need no pivoting; a
    // pivot below 1e-30 (or negative) or a non-finite result fails the solve
    // as a singular pivot fails the reference's. The device forms 1 / pivot
    // from the hardware reciprocal and two Newton steps (within an ulp of
    // the division here); everything else is the device's operations in its
    // order: Schur terms accumulated into S by fused multiply-adds over
    // k = 0..5, row operations S[r][c] = fma(S[r][C], -S[C][c], S[r][c]).
    bool block_solve_fused()
    {
        std::vector<uint8_t> failed(strands_.size(), 0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const uint32_t begin = st.first + st.fixed_count, end = st.first + st.count;
            if (begin >= end) return;
            double e_prev[6] = {0, 0, 0, 0, 0, 0};
            for (uint32_t n = begin; n < end; ++n) {
                double S[36], W[36], e[6];
                for (int i = 0; i < 36; ++i) {
                    S[i] = D_[36 * n + i];
                    W[i] = n + 1 < end ? U_[36 * n + i] : 0.0;
                }
                for (int i = 0; i < 6; ++i) e[i] = g_[6 * n + i];
                if (n > begin) {
                    const double *U = &U_[36 * (n - 1)], *Cprev = &C_[36 * (n - 1)];
                    for (int k = 0; k < 6; ++k)
                        for (int R = 0; R < 6; ++R) {
                            const double u = U[6 * k + R];
                            for (int c = 0; c < 6; ++c) S[6 * R + c] = std::fma(u, -Cprev[6 * k + c], S[6 * R + c]);
                            e[R] = std::fma(u, -e_prev[k], e[R]);
                        }
                }
                double scale = 0.0;
                for (int i = 0; i < 6; ++i) scale = std::fmax(scale, std::fabs(S[6 * i + i]));
                const double shift = std::fmax(1.0e-12 * scale, 1.0e-24);
                double smallest = 0.0;
                for (int C = 0; C < 6; ++C) {
                    const double pivot = S[6 * C + C] + shift;
                    smallest = C == 0 ? pivot : std::fmin(smallest, pivot);
                    const double inv = 1.0 / pivot;
                    for (int c = 0; c < 6; ++c) {
                        S[6 * C + c] *= inv;
                        W[6 * C + c] *= inv;
                    }
                    e[C] *= inv;
                    for (int r = 0; r < 6; ++r) {
                        if (r == C) continue;
                        const double f = S[6 * r + C];
                        for (int c = 0; c < 6; ++c) {
                            S[6 * r + c] = std::fma(f, -S[6 * C + c], S[6 * r + c]);
                            W[6 * r + c] = std::fma(f, -W[6 * C + c], W[6 * r + c]);
                        }
                        e[r] = std::fma(f, -e[C], e[r]);
                    }
                }
                if (!(smallest >= 1.0e-30) || !std::isfinite(e[0])) { failed[s] = 1; return; }
                for (int i = 0; i < 36; ++i) C_[36 * n + i] = W[i];
                for (int i = 0; i < 6; ++i) {
                    dp_[6 * n + i] = -e[i];
                    e_prev[i] = e[i];
                }
            }
            for (uint32_t n = end; n-- > begin;)
                for (int r = 0; r < 6; ++r) {
                    double value = dp_[6 * n + r];
                    if (n + 1 < end)
                        for (int k = 0; k < 6; ++k) value = std::fma(-C_[36 * n + 6 * r + k], d_[6 * (n + 1) + k], value);
                    d_[6 * n + r] = value;
                }
        });
        for (uint8_t f : failed) if (f) return true;
        return false;
    }

    bool block_solve()
    {
        if (precision_mode == 1) return block_solve_fp32();
        if (solve_algorithm == 1) return block_solve_twisted();
        if (solve_algorithm == 2) return block_solve_fused();
        std::vector<uint8_t> failed(strands_.size(), 0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const uint32_t begin = st.first + st.fixed_count, end = st.first + st.count;
            if (begin >= end) return;
            for (uint32_t n = begin; n < end; ++n) {
                double b[36], inv[36], rhs[6];
                for (int i = 0; i < 36; ++i) b[i] = D_[36 * n + i];
                for (int i = 0; i < 6; ++i) rhs[i] = -g_[6 * n + i];
                if (n > begin) {
                    const double *lower = &U_[36 * (n - 1)], *pc = &C_[36 * (n - 1)], *pd = &dp_[6 * (n - 1)];
                    for (int r = 0; r < 6; ++r) {
                        for (int c = 0; c < 6; ++c) {
                            double sum = 0;
                            for (int k = 0; k < 6; ++k) sum = std::fma(lower[6 * k + r], pc[6 * k + c], sum);
                            b[6 * r + c] -= sum;
                        }
                        double sum = 0;
                        for (int k = 0; k < 6; ++k) sum = std::fma(lower[6 * k + r], pd[k], sum);
                        rhs[r] -= sum;
                    }
                }
                if (!inverse6(b, inv)) { failed[s] = 1; return; }
                double *cc = &C_[36 * n], *cd = &dp_[6 * n];
                if (n + 1 < end) {
                    const double *up = &U_[36 * n];
                    for (int r = 0; r < 6; ++r)
                        for (int c = 0; c < 6; ++c) {
                            double sum = 0;
                            for (int k = 0; k < 6; ++k) sum = std::fma(inv[6 * r + k], up[6 * k + c], sum);
                            cc[6 * r + c] = sum;
                        }
                } else {
                    for (int i = 0; i < 36; ++i) cc[i] = 0.0;
                }
                for (int r = 0; r < 6; ++r) {
                    double sum = 0;
                    for (int k = 0; k < 6; ++k) sum = std::fma(inv[6 * r + k], rhs[k], sum);
                    cd[r] = sum;
                }
            }
            for (uint32_t n = end; n-- > begin;)
                for (int r = 0; r < 6; ++r) {
                    double value = dp_[6 * n + r];
                    if (n + 1 < end)
                        for (int k = 0; k < 6; ++k) value = std::fma(-C_[36 * n + 6 * r + k], d_[6 * (n + 1) + k], value);
                    d_[6 * n + r] = value;
                }
        });
        for (uint8_t f : failed) if (f) return true;
        return false;
    }

    // Fixed-order sums over strands of per-strand partial sums.
    double dot_vec(const std::vector<double> &a, const std::vector<double> &b) const
    {
        std::vector<double> part(strands_.size(), 0.0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            double acc = 0.0;
            const Strand &st = strands_[s];
            for (uint32_t i = 6 * st.first; i < 6 * (st.first + st.count); ++i) acc += a[i] * b[i];
            part[s] = acc;
        });
        double total = 0.0;
        for (double p : part) total += p;
        return total;
    }
    double nrm2(const std::vector<double> &a) const { return std::sqrt(dot_vec(a, a)); }

    void preconditioner(double dt)
    {
        // cuda_backend.cu:815-841
        const double inv = 1.0 / dt, inv2 = inv * inv;
        // per element: stretch and bend/twist bounds of its strand's material
        auto bounds = [&](uint32_t strand, double &tr_el, double &rot_a, double &rot_b) {
            const double r = radius_of(strand), E = young_of(strand);
            const double area = kPi * r * r, inertia = kPi * r * r * r * r / 4.0;
            const double shear = E / (2.0 * (1.0 + material_.poisson_ratio));
            tr_el = std::fmax(E * area, material_.shear_correction * shear * area);
            rot_a = material_.shear_correction * shear * area;
            rot_b = std::fmax(E * inertia, 2.0 * shear * inertia);
        };
        for (size_t n = 0; n < nodes_.size(); ++n) {
            const double p = nodes_[n].mass * (inv2 + material_.mass_damping * inv);
            const double rr = nodes_[n].rot_mass * (inv2 + material_.mass_damping * inv);
            for (int d = 0; d < 3; ++d) { P_[6 * n + d] = p; P_[6 * n + 3 + d] = rr; }
        }
        for (const Element &e : elements_) {
            double tr_el, rot_a, rot_b;
            bounds(e.strand, tr_el, rot_a, rot_b);
            const double t = tr_el / e.rest_length;
            const double rot = rot_a * e.rest_length + rot_b / e.rest_length;
            for (int d = 0; d < 3; ++d) {
                P_[6 * e.i + d] += t;
                P_[6 * e.j + d] += t;
                P_[6 * e.i + 3 + d] += rot;
                P_[6 * e.j + 3 + d] += rot;
            }
        }
        for (size_t n = 0; n < nodes_.size(); ++n)
            for (int d = 0; d < 6; ++d) P_[6 * n + d] = nodes_[n].fixed ? 0.0 : 1.0 / std::fmax(P_[6 * n + d], 1.0e-30);
    }

    void limit_direction()
    {
        // cuda_backend.cu:623-630, 848-853
        double maxt = 0.0;
        for (size_t n = 0; n < nodes_.size(); ++n) {
            if (nodes_[n].fixed) continue;
            maxt = std::fmax(maxt, norm(v3(d_[6 * n], d_[6 * n + 1], d_[6 * n + 2])));
        }
        if (maxt > trust_radius_) {
            const double scale = trust_radius_ / maxt;
            for (double &x : d_) x *= scale;
        }
    }

    void steepest_direction()
    {
        // cuda_backend.cu:632-638
        for (size_t i = 0; i < d_.size(); ++i) d_[i] = (-g_[i]) * P_[i];
        for (size_t n = 0; n < nodes_.size(); ++n)
            if (nodes_[n].fixed) for (int d = 0; d < 6; ++d) d_[6 * n + d] = 0.0;
        limit_direction();
    }

    bool make_direction()
    {
        // cuda_backend.cu:640-648
        if (direction_mode == 1) {
            steepest_direction();
            return false;
        }
        std::fill(d_.begin(), d_.end(), 0.0);
        const bool failed = block_solve();
        bool used = false;
        if (!failed) {
            const double slope = dot_vec(g_, d_);
            used = slope < 0 && std::isfinite(slope);
        }
        if (!used) {
            steepest_direction();
            return false;
        }
        limit_direction();
        return true;
    }

    // Line-search CCD: conservative advancement of each contact-enabled
    // segment moving along the direction against the static collider
    // (cuda_backend.cu:989-1000).
    double ccd_limit()
    {
        std::vector<double> part(strands_.size(), 1.0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const double target = radius_of(s) + material_.collider_offset + desc_.minimum_gap;
            double limit = 1.0;
            std::vector<uint32_t> cand;
            const uint32_t e0 = st.first - s;
            for (uint32_t k = e0; k + 1 < e0 + st.count; ++k) {
                const Element &e = elements_[k];
                if (nodes_[e.i].fixed || !e.contact) continue;
                const V3 p0 = pos(e.i), p1 = pos(e.j);
                const V3 d0{d_[6 * e.i], d_[6 * e.i + 1], d_[6 * e.i + 2]};
                const V3 d1{d_[6 * e.j], d_[6 * e.j + 1], d_[6 * e.j + 2]};
                const V3 lo = vmin(vmin(p0, p1), vmin(p0 + d0, p1 + d1));
                const V3 hi = vmax(vmax(p0, p1), vmax(p0 + d0, p1 + d1));
                bvh_.query(lo, hi, target, cand);
                for (uint32_t t : cand) {
                    const Tri &tr = tris_[t];
                    const double speed = std::fmax(norm(d0), norm(d1));
                    if (speed < 1e-30) continue;
                    double alpha = 0;
                    bool done = false;
                    for (int it = 0; it < 80; ++it) {
                        const double dist = closest_segment_triangle(p0 + d0 * alpha, p1 + d1 * alpha, collider_cur_[tr.i0],
                                                                     collider_cur_[tr.i1], collider_cur_[tr.i2], false)
                                                .distance;
                        const double clear = dist - target;
                        if (clear <= 0) { limit = std::fmin(limit, std::fmax(0.0, alpha * 0.9)); done = true; break; }
                        const double advance = 0.8 * clear / speed;
                        if (alpha + advance >= 1) { done = true; break; }
                        if (advance < 1e-12) { limit = std::fmin(limit, alpha); done = true; break; }
                        alpha += advance;
                    }
                    if (!done) limit = std::fmin(limit, alpha);
                }
            }
            part[s] = limit;
        });
        double limit = 1.0;
        for (double p : part) limit = std::fmin(limit, p);
        return limit;
    }

    // Moving-collider sweep at substep start: conservative advancement of
    // each triangle from its old to its current position against the
    // stationary segment (cuda_backend.cu:975-987). The lowest failing
    // element, then its lowest failing triangle, is reported.
    bool sweep_feasible(SweepFailure &fail_out)
    {
        std::vector<SweepFailure> part(strands_.size());
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const double target = radius_of(s) + material_.collider_offset + desc_.minimum_gap;
            std::vector<uint32_t> cand;
            const uint32_t e0 = st.first - s;
            for (uint32_t k = e0; k + 1 < e0 + st.count && !part[s].claimed; ++k) {
                const Element &e = elements_[k];
                if (nodes_[e.i].fixed || !e.contact) continue;
                const V3 p0 = pos(e.i), p1 = pos(e.j);
                bvh_.query(vmin(p0, p1), vmax(p0, p1), target, cand);
                for (uint32_t t : cand) {
                    const Tri &tr = tris_[t];
                    const V3 oa = collider_old_[tr.i0], ob = collider_old_[tr.i1], oc = collider_old_[tr.i2];
                    const V3 da = collider_cur_[tr.i0] - oa, db = collider_cur_[tr.i1] - ob, dc = collider_cur_[tr.i2] - oc;
                    const double speed = std::fmax(norm(da), std::fmax(norm(db), norm(dc)));
                    if (speed < 1e-30) continue;
                    double alpha = 0;
                    bool end = false, bad = false;
                    Closest pair;
                    V3 a, b, c;
                    for (int it = 0; it < 80; ++it) {
                        a = oa + da * alpha;
                        b = ob + db * alpha;
                        c = oc + dc * alpha;
                        pair = closest_segment_triangle(p0, p1, a, b, c, false);
                        const double clearance = pair.distance - target;
                        if (clearance <= 0) { bad = true; break; }
                        const double advance = 0.8 * clearance / speed;
                        if (alpha + advance >= 1) { end = true; break; }
                        if (advance < 1e-12) { bad = true; break; }
                        alpha += advance;
                    }
                    if (!end) bad = true;
                    if (bad) {
                        SweepFailure &f = part[s];
                        f.claimed = true;
                        f.strand = e.strand;
                        f.element = k;
                        f.triangle = t;
                        f.distance = pair.distance;
                        f.required = target;
                        f.clearance = pair.distance - target;
                        f.displacement = speed;
                        f.hair_start = p0;
                        f.hair_end = p1;
                        f.collider_point = a * pair.bary.x + b * pair.bary.y + c * pair.bary.z;
                        break;
                    }
                }
            }
        });
        for (const SweepFailure &f : part)
            if (f.claimed) { fail_out = f; return false; }
        return true;
    }

    // Post-substep smoothed Coulomb impulse against the closest triangle
    // (cuda_backend.cu:941-964, 615-621).
    void apply_friction(double dt)
    {
        if (!has_collider_ || !(material_.friction > 0.0)) return;
        std::fill(dv_.begin(), dv_.end(), 0.0);
        parallel_for(threads, uint32_t(strands_.size()), [&](uint32_t s) {
            const Strand &st = strands_[s];
            const double pad = radius_of(s) + material_.collider_offset + material_.barrier_distance;
            std::vector<uint32_t> cand;
            const uint32_t e0 = st.first - s;
            for (uint32_t k = e0; k + 1 < e0 + st.count; ++k) {
                const Element &e = elements_[k];
                if (!e.contact) continue;
                const Node &ni = nodes_[e.i], &nj = nodes_[e.j];
                if (ni.fixed && nj.fixed) continue;
                const V3 p0 = pos(e.i), p1 = pos(e.j);
                Closest pair;
                uint32_t ti = 0;
                if (!closest_triangle(p0, p1, pad, cand, pair, ti, nullptr)) continue;
                const double gap = pair.distance - radius_of(s) - material_.collider_offset;
                if (!(gap > 0.0 && gap < material_.barrier_distance)) continue;
                const Tri &t = tris_[ti];
                const double w0 = 1.0 - pair.t, w1 = pair.t;
                const V3 hair = p0 * w0 + p1 * w1;
                const V3 body = collider_cur_[t.i0] * pair.bary.x + collider_cur_[t.i1] * pair.bary.y +
                                collider_cur_[t.i2] * pair.bary.z;
                const V3 old_body = collider_old_[t.i0] * pair.bary.x + collider_old_[t.i1] * pair.bary.y +
                                    collider_old_[t.i2] * pair.bary.z;
                const V3 normal = div_recip(hair - body, std::fmax(pair.distance, 1.0e-15));
                const V3 v0{v_[6 * e.i], v_[6 * e.i + 1], v_[6 * e.i + 2]};
                const V3 v1{v_[6 * e.j], v_[6 * e.j + 1], v_[6 * e.j + 2]};
                const V3 rel = v0 * w0 + v1 * w1 - div_recip(body - old_body, dt);
                const V3 tangent = rel - normal * dot(normal, rel);
                const double tspeed = norm(tangent);
                if (tspeed < 1.0e-15) continue;
                double inv_mass = 0.0;
                if (!ni.fixed) inv_mass += w0 * w0 / std::fmax(ni.mass, 1.0e-30);
                if (!nj.fixed) inv_mass += w1 * w1 / std::fmax(nj.mass, 1.0e-30);
                if (!(inv_mass > 0.0)) continue;
                const double aa = material_.barrier_distance - gap, lg = std::log(gap / material_.barrier_distance);
                const double fn = material_.contact_stiffness * (aa * aa / gap - 2.0 * aa * lg);
                const double required = tspeed / inv_mass;
                const double mag = std::fmin(required, material_.friction * fn * dt);
                const double smooth = material_.friction_smoothing / dt;
                const V3 impulse = tangent * (-mag / std::fmax(std::sqrt(tspeed * tspeed + smooth * smooth), 1.0e-30));
                if (!ni.fixed) {
                    const double sc = w0 / std::fmax(ni.mass, 1.0e-30);
                    dv_[6 * e.i] += impulse.x * sc;
                    dv_[6 * e.i + 1] += impulse.y * sc;
                    dv_[6 * e.i + 2] += impulse.z * sc;
                }
                if (!nj.fixed) {
                    const double sc = w1 / std::fmax(nj.mass, 1.0e-30);
                    dv_[6 * e.j] += impulse.x * sc;
                    dv_[6 * e.j + 1] += impulse.y * sc;
                    dv_[6 * e.j + 2] += impulse.z * sc;
                }
            }
        });
        for (size_t i = 0; i < v_.size(); ++i) v_[i] = 1.0 * dv_[i] + v_[i];
    }

    void update_velocity(double inv_dt)
    {
        // cuda_backend.cu:803-808
        for (size_t n = 0; n < nodes_.size(); ++n)
            if (!nodes_[n].fixed)
                for (int d = 0; d < 6; ++d) v_[6 * n + d] = (q_[6 * n + d] - q_old_[6 * n + d]) * inv_dt;
    }

    KhsResult fail(KhsResult r, const std::string &msg)
    {
        error = msg;
        step_stats_.phase = KHS_PHASE_FAILED;
        prog_phase_.store(KHS_PHASE_FAILED);
        return r;
    }

    void clear_failure()
    {
        failure_ = {};
        failure_.struct_size = sizeof(failure_);
        failure_.strand_index = 0xffffffffu;
        failure_.element_index = 0xffffffffu;
        failure_.collider_triangle_index = 0xffffffffu;
    }

    void set_failure_context(uint32_t kind, uint32_t frame, uint32_t sub, uint32_t attempted, uint32_t maximum,
                             uint32_t attempts)
    {
        failure_.kind = kind;
        failure_.frame_index = frame;
        failure_.substep = sub;
        failure_.requested_substeps = desc_.substeps;
        failure_.attempted_substeps = attempted;
        failure_.maximum_substeps = maximum;
        failure_.adaptive_attempt_count = attempts;
    }

    // Newton loop of one substep (cuda_backend.cu:650-688).
    KhsResult optimize_substep(double dt)
    {
        preconditioner(dt);
        Stats current = evaluate(true, dt);
        if (!current.feasible) return fail(KHS_ERROR_NOT_CONVERGED, "State is outside the barrier feasible region.");
        const double initial = nrm2(g_);
        step_stats_.initial_residual_norm = std::fmax(step_stats_.initial_residual_norm, initial);
        const double dofs = double(q_.size());
        for (uint32_t it = 0; it <= desc_.newton_iterations; ++it) {
            prog_iteration_.store(it);
            if (cancel_.load()) return fail(KHS_ERROR_INVALID_STATE, "Solve cancelled.");
            const double residual = nrm2(g_);
            const double relative = residual / std::fmax(initial, 1e-30);
            step_stats_.final_residual_norm = residual;
            step_stats_.relative_residual_norm = relative;
            step_stats_.minimum_gap = std::fmin(step_stats_.minimum_gap, current.min_gap);
            step_stats_.contact_candidate_count += current.candidates;
            step_stats_.active_contact_count += current.active;
            step_stats_.kinetic_energy = current.kinetic;
            step_stats_.elastic_energy = current.elastic;
            step_stats_.contact_energy = current.contact;
            step_stats_.friction_energy = current.friction;
            const double abs_tol = convergence_mode ? desc_.absolute_tolerance
                                                    : std::fmax(desc_.absolute_tolerance, 1.0e-5 * std::sqrt(dofs));
            const double rel_tol = convergence_mode ? desc_.relative_tolerance : std::fmax(desc_.relative_tolerance, 1.0e-4);
            if (residual <= abs_tol || relative <= rel_tol) {
                update_velocity(1.0 / dt);
                return KHS_OK;
            }
            if (it == desc_.newton_iterations) break;
            prog_phase_.store(KHS_PHASE_LINEAR_SOLVE);
            bool block_direction = make_direction();
            ++(block_direction ? block_directions_ : steepest_directions_);
            ++step_stats_.linear_solves;
            q_base_ = q_;
            bool accepted = false;
            double alpha = 1.0;
            for (int attempt = 0; attempt < 2; ++attempt) {
                const double slope = dot_vec(g_, d_);
                alpha = 1.0;
                if (has_collider_) {
                    prog_phase_.store(KHS_PHASE_CCD);
                    alpha = ccd_limit();
                    if (alpha < 1.0) alpha *= 0.9;
                    step_stats_.ccd_step_limit = std::fmin(step_stats_.ccd_step_limit, alpha);
                }
                accepted = false;
                prog_phase_.store(KHS_PHASE_LINE_SEARCH);
                if (alpha >= desc_.minimum_line_search_step)
                    for (uint32_t search = 0; search < desc_.line_search_iterations; ++search) {
                        for (size_t i = 0; i < q_.size(); ++i) q_[i] = alpha * d_[i] + q_base_[i];
                        const Stats trial = evaluate(false, dt);
                        ++step_stats_.line_search_evaluations;
                        if (trial.feasible && std::isfinite(trial.objective) &&
                            trial.objective <= current.objective + 1e-4 * alpha * slope) {
                            accepted = true;
                            break;
                        }
                        alpha *= 0.5;
                        if (alpha < desc_.minimum_line_search_step) break;
                    }
                if (accepted) break;
                if (!block_direction) break;
                q_ = q_base_;
                steepest_direction();
                ++steepest_directions_;
                block_direction = false;
            }
            if (!accepted) {
                q_ = q_base_;
                return fail(KHS_ERROR_NOT_CONVERGED, "Gauss-Newton line search failed.");
            }
            current = evaluate(true, dt);
            const double dnorm = nrm2(d_);
            step_stats_.increment_norm = alpha * dnorm;
            step_stats_.accepted_step_length = alpha;
            ++step_stats_.newton_iterations;
            prog_phase_.store(KHS_PHASE_ASSEMBLING);
            const double rms = step_stats_.increment_norm / std::sqrt(dofs);
            if (rms <= 2.0 * desc_.increment_tolerance) {
                update_velocity(1.0 / dt);
                return KHS_OK;
            }
        }
        update_velocity(1.0 / dt);
        return KHS_OK;
    }

    // One frame: substeps with root/collider interpolation, sweep check,
    // Newton solve, friction; on failure the frame is restored and retried
    // with doubled substeps up to the maximum (cuda_backend.cu:690-721).
    KhsResult run_frame(double frame_dt, const V3 *root_a, const V3 *rot_a, const V3 *root_b, const V3 *rot_b,
                        const V3 *col_a, const V3 *col_b, uint32_t frame, uint32_t frame_count, size_t vc)
    {
        if (!(frame_dt > 0 && std::isfinite(frame_dt))) return fail(KHS_ERROR_INVALID_ARGUMENT, "Invalid time step.");
        const auto t0 = std::chrono::steady_clock::now();
        cancel_.store(false);
        clear_failure();
        frame_clock_.store(t0.time_since_epoch().count());
        prog_frame_.store(frame);
        prog_frame_count_.store(frame_count);
        prog_iteration_.store(0);
        snap_q_ = q_;
        snap_v_ = v_;
        // The interpolation endpoints may alias buffers the loop rewrites.
        const std::vector<V3> ca(col_a, col_a + vc), cb(col_b, col_b + vc);
        uint32_t substeps = desc_.substeps;
        const uint32_t max_sub = has_collider_ ? desc_.maximum_substeps : desc_.substeps;
        uint32_t attempts = 0;
        const size_t n_nodes = nodes_.size();
        for (;;) {
            ++attempts;
            q_ = snap_q_;
            v_ = snap_v_;
            clear_failure();
            error.clear();
            prog_phase_.store(KHS_PHASE_ASSEMBLING);
            prog_substep_count_.store(substeps);
            prog_iteration_.store(0);
            step_stats_ = {};
            step_stats_.struct_size = sizeof(step_stats_);
            step_stats_.substeps = substeps;
            step_stats_.minimum_gap = kInf;
            step_stats_.ccd_step_limit = 1.0;
            block_directions_ = steepest_directions_ = 0;
            if (has_collider_) {
                collider_cur_ = ca;
                collider_frame_start_ = ca;
            }
            const double dt = frame_dt / substeps;
            bool crossed = false;
            KhsResult result = KHS_OK;
            for (uint32_t sub = 1; sub <= substeps; ++sub) {
                prog_substep_.store(sub);
                prog_iteration_.store(0);
                q_old_ = q_;
                const double f = double(sub) / substeps;
                // cuda_backend.cu:794-801: only fixed nodes move to the
                // interpolated targets; free nodes start at q_old.
                for (size_t n = 0; n < n_nodes; ++n) {
                    if (!nodes_[n].fixed) continue;
                    const V3 p = root_a[n] * (1.0 - f) + root_b[n] * f;
                    const V3 r = rot_a[n] * (1.0 - f) + rot_b[n] * f;
                    const double val[6] = {p.x, p.y, p.z, r.x, r.y, r.z};
                    for (int d = 0; d < 6; ++d) {
                        q_[6 * n + d] = val[d];
                        v_[6 * n + d] = (val[d] - q_old_[6 * n + d]) / dt;
                    }
                }
                if (has_collider_) {
                    collider_old_ = collider_cur_;
                    for (size_t i = 0; i < vc; ++i) collider_cur_[i] = ca[i] * (1.0 - f) + cb[i] * f;
                    bvh_.refit(collider_cur_, &collider_old_);
                    prog_phase_.store(KHS_PHASE_CCD);
                    SweepFailure sw;
                    if (!sweep_feasible(sw)) {
                        set_failure_context(KHS_FAILURE_MOVING_COLLIDER_SWEEP, frame, sub, substeps, max_sub, attempts);
                        failure_.strand_index = sw.strand;
                        failure_.element_index = sw.element;
                        failure_.collider_triangle_index = sw.triangle;
                        failure_.distance = sw.distance;
                        failure_.required_distance = sw.required;
                        failure_.clearance = sw.clearance;
                        failure_.collider_substep_displacement = sw.displacement;
                        failure_.collider_frame_displacement = sw.displacement * substeps;
                        failure_.hair_start = to_c(sw.hair_start);
                        failure_.hair_end = to_c(sw.hair_end);
                        failure_.collider_point = to_c(sw.collider_point);
                        crossed = true;
                        break;
                    }
                    bvh_.refit(collider_cur_, nullptr);
                }
                prog_phase_.store(KHS_PHASE_ASSEMBLING);
                result = optimize_substep(dt);
                if (result != KHS_OK) {
                    if (!cancel_.load()) {
                        const uint32_t kind = error.find("outside the barrier") != std::string::npos
                                                  ? KHS_FAILURE_BARRIER_INFEASIBLE
                                                  : KHS_FAILURE_NONLINEAR_SOLVE;
                        set_failure_context(kind, frame, sub, substeps, max_sub, attempts);
                    }
                    break;
                }
                apply_friction(dt);
                ++step_stats_.converged_substeps;
            }
            if (!crossed && result == KHS_OK) break;
            q_ = snap_q_;
            v_ = snap_v_;
            if (cancel_.load()) return result;
            if (substeps >= max_sub) {
                if (crossed) return fail(KHS_ERROR_NOT_CONVERGED, "Animated collider crossed hair after adaptive substepping.");
                return result;
            }
            substeps = std::min(max_sub, substeps * 2);
        }
        if (!anim_ready_) {
            root_cur_ = root_pending_;
            rot_cur_ = rot_pending_;
            if (has_collider_) collider_frame_start_ = collider_pending_;
        }
        clear_failure();
        step_stats_.phase = KHS_PHASE_FINISHED;
        prog_phase_.store(KHS_PHASE_FINISHED);
        last_frame_ms_ = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!(last_frame_ms_ > 0.0)) last_frame_ms_ = 1e-6;
        error.clear();
        return KHS_OK;
    }

    uint64_t block_directions_ = 0, steepest_directions_ = 0;
    // per-strand Young's modulus and radius (khsxSetStrandMaterials); empty:
    // the material's values for every strand
    std::vector<double> strand_E_, strand_r_;
    KhsSolverDesc desc_{};
    KhsHairMaterial material_{};
    bool has_hair_ = false, has_collider_ = false, built_ = false, closed_ = false;
    std::vector<V3> input_points_;
    std::vector<uint32_t> input_offsets_;
    std::vector<uint32_t> mapping_;
    std::vector<Node> nodes_;
    std::vector<Element> elements_;
    std::vector<Strand> strands_;
    uint64_t free_dofs_ = 0;
    uint32_t merged_ = 0, ext_strands_ = 0, ext_nodes_ = 0;
    double ext_length_ = 0.0, max_rest_length_ = 0.0, trust_radius_ = 0.0, collider_x_max_ = 0.0;
    std::vector<Tri> tris_;
    std::vector<V3> collider_input_, collider_pending_, collider_cur_, collider_old_, collider_frame_start_;
    uint32_t excluded_ = 0, boundary_edges_ = 0, nonmanifold_edges_ = 0, inconsistent_edges_ = 0, inverted_ = 0;
    Bvh bvh_;
    std::vector<double> q_, q_old_, v_, snap_q_, snap_v_, g_, d_, P_, q_base_, dv_, D_, U_, C_, dp_;
    std::vector<V3> root_cur_, rot_cur_, root_pending_, rot_pending_;
    uint32_t anim_frames_ = 0, anim_current_ = 0;
    bool anim_ready_ = false;
    std::vector<uint32_t> ring_frame_;
    std::vector<std::vector<double>> ring_q_, ring_v_;
    std::vector<V3> anim_root_, anim_rot_, anim_collider_;
    std::vector<bool> root_set_, collider_set_;
    KhsBuildStats build_stats_{};
    KhsStepStats step_stats_{};
    KhsFailureDiagnostics failure_{};
    double last_frame_ms_ = 0.0;
    std::atomic<bool> cancel_{false};
    std::atomic<uint32_t> prog_phase_{KHS_PHASE_IDLE}, prog_frame_{0}, prog_frame_count_{0}, prog_substep_{0},
        prog_substep_count_{0}, prog_iteration_{0};
    std::atomic<int64_t> frame_clock_{0};
};

bool valid_desc(const KhsSolverDesc *d)
{
    // api.cpp:195-205
    return d && d->struct_size >= sizeof(KhsSolverDesc) && d->substeps > 0 && d->maximum_substeps >= d->substeps &&
           d->newton_iterations > 0 && d->line_search_iterations > 0 && d->absolute_tolerance > 0.0 &&
           d->relative_tolerance > 0.0 && d->increment_tolerance > 0.0 && d->minimum_line_search_step > 0.0 &&
           d->minimum_gap > 0.0 && d->maximum_element_length > 0.0 && d->minimum_dynamic_length >= 0.0 &&
           std::isfinite(d->minimum_dynamic_length) && d->fixed_root_nodes > 0;
}

bool valid_material(const KhsHairMaterial *m)
{
    // api.cpp:207-216
    return m && m->struct_size >= sizeof(KhsHairMaterial) && m->density > 0.0 && m->radius > 0.0 &&
           m->young_modulus > 0.0 && m->poisson_ratio > -1.0 && m->poisson_ratio < 0.5 &&
           m->shear_correction > 0.0 && m->mass_damping >= 0.0 && m->contact_stiffness > 0.0 &&
           m->barrier_distance > 0.0 && m->friction >= 0.0 && m->friction_smoothing > 0.0 &&
           m->collider_offset >= 0.0;
}

}  // namespace
}  // namespace oracle

struct KhsSolver {
    explicit KhsSolver(const KhsSolverDesc &d) : impl(d) {}
    oracle::Solver impl;
};

namespace {
template <class F>
KhsResult guarded(KhsSolver *s, F &&f)
{
    // api.cpp:16-35
    if (!s) return KHS_ERROR_INVALID_ARGUMENT;
    try {
        return f();
    } catch (const std::bad_alloc &) {
        s->impl.error = "Out of memory.";
        return KHS_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        s->impl.error = std::string("Internal exception: ") + e.what();
        return KHS_ERROR_INTERNAL;
    } catch (...) {
        s->impl.error = "Unknown internal exception.";
        return KHS_ERROR_INTERNAL;
    }
}
}  // namespace

extern "C" {

uint32_t khsGetAbiVersion(void) { return KHS_ABI_VERSION; }

KhsResult khsGetGpuInfo(KhsGpuInfo *info)
{
    if (!info || info->struct_size < sizeof(KhsGpuInfo)) return KHS_ERROR_INVALID_ARGUMENT;
    KhsGpuInfo r{};
    r.struct_size = sizeof(r);
    r.device_ordinal = -1;
    std::strncpy(r.device_name, "CPU oracle (no GPU)", sizeof(r.device_name) - 1);
    *info = r;
    return KHS_ERROR_INVALID_STATE;
}

void khsDefaultSolverDesc(KhsSolverDesc *d)
{
    // api.cpp:237-256
    if (!d) return;
    std::memset(d, 0, sizeof(*d));
    d->struct_size = sizeof(*d);
    d->gravity = {0.0, 0.0, -9.81};
    d->substeps = 8;
    d->maximum_substeps = 32;
    d->newton_iterations = 24;
    d->line_search_iterations = 20;
    d->absolute_tolerance = 1.0e-8;
    d->relative_tolerance = 1.0e-5;
    d->increment_tolerance = 1.0e-8;
    d->minimum_line_search_step = 1.0e-9;
    d->minimum_gap = 1.0e-7;
    d->maximum_element_length = 0.01;
    d->minimum_dynamic_length = 0.0;
    d->fixed_root_nodes = 2;
    d->thread_count = 0;
}

void khsDefaultHairMaterial(KhsHairMaterial *m)
{
    // api.cpp:258-274
    if (!m) return;
    std::memset(m, 0, sizeof(*m));
    m->struct_size = sizeof(*m);
    m->density = 1300.0;
    m->radius = 4.0e-5;
    m->young_modulus = 4.0e9;
    m->poisson_ratio = 0.38;
    m->shear_correction = 0.9;
    m->mass_damping = 8.0;
    m->contact_stiffness = 1.0e4;
    m->barrier_distance = 2.0e-4;
    m->friction = 0.35;
    m->friction_smoothing = 1.0e-6;
    m->collider_offset = 0.0;
}

KhsSolver *khsCreate(const KhsSolverDesc *desc)
{
    if (!oracle::valid_desc(desc)) return nullptr;
    try {
        return new KhsSolver(*desc);
    } catch (...) {
        return nullptr;
    }
}

void khsDestroy(KhsSolver *s) { delete s; }

KhsResult khsSetHairCurves(KhsSolver *s, const KhsVec3 *points, uint32_t point_count, const uint32_t *offsets,
                           uint32_t strand_count, const KhsHairMaterial *material)
{
    if (!material || material->struct_size < sizeof(KhsHairMaterial)) return KHS_ERROR_INVALID_ARGUMENT;
    return guarded(s, [&] { return s->impl.set_hair(points, point_count, offsets, strand_count, *material); });
}

KhsResult khsSetColliderMesh(KhsSolver *s, const KhsVec3 *v, uint32_t vc, const KhsTriangle *t, uint32_t tc)
{
    return guarded(s, [&] { return s->impl.set_collider(v, vc, t, tc); });
}

KhsResult khsBuild(KhsSolver *s) { return guarded(s, [&] { return s->impl.build(); }); }

KhsResult khsUpdateRuntimeParameters(KhsSolver *s, const KhsSolverDesc *d, const KhsHairMaterial *m)
{
    if (!oracle::valid_desc(d) || !oracle::valid_material(m)) return KHS_ERROR_INVALID_ARGUMENT;
    return guarded(s, [&] { return s->impl.update_runtime(*d, *m); });
}

KhsResult khsUpdateColliderVertices(KhsSolver *s, const KhsVec3 *v, uint32_t c)
{
    return guarded(s, [&] { return s->impl.update_collider(v, c); });
}

KhsResult khsUpdateRootTargets(KhsSolver *s, const KhsVec3 *p, uint32_t c)
{
    return guarded(s, [&] { return s->impl.update_roots(p, c); });
}

KhsResult khsStep(KhsSolver *s, double dt) { return guarded(s, [&] { return s->impl.step(dt); }); }

KhsResult khsAllocateAnimation(KhsSolver *s, uint32_t frames)
{
    return guarded(s, [&] { return s->impl.allocate_animation(frames); });
}

KhsResult khsSetRootAnimationFrame(KhsSolver *s, uint32_t f, const KhsVec3 *p, uint32_t c)
{
    return guarded(s, [&] { return s->impl.set_root_frame(f, p, c); });
}

KhsResult khsSetColliderAnimationFrame(KhsSolver *s, uint32_t f, const KhsVec3 *v, uint32_t c)
{
    return guarded(s, [&] { return s->impl.set_collider_frame(f, v, c); });
}

KhsResult khsFinalizeAnimation(KhsSolver *s) { return guarded(s, [&] { return s->impl.finalize_animation(); }); }

KhsResult khsStepAnimationFrame(KhsSolver *s, uint32_t f, double dt)
{
    return guarded(s, [&] { return s->impl.step_animation_frame(f, dt); });
}

uint64_t khsGetAnimationCheckpointSize(const KhsSolver *s) { return s ? s->impl.checkpoint_size() : 0; }

KhsResult khsSaveAnimationCheckpoint(KhsSolver *s, void *data, uint64_t cap)
{
    return guarded(s, [&] { return s->impl.save_checkpoint(data, cap); });
}

KhsResult khsRestoreAnimationCheckpoint(KhsSolver *s, const void *data, uint64_t size)
{
    return guarded(s, [&] { return s->impl.restore_checkpoint(data, size); });
}

uint32_t khsGetOriginalPointCount(const KhsSolver *s) { return s ? s->impl.original_count() : 0; }
uint32_t khsGetInternalNodeCount(const KhsSolver *s) { return s ? s->impl.node_count() : 0; }

KhsResult khsCopyOriginalPositions(const KhsSolver *s, KhsVec3 *p, uint32_t cap)
{
    return s ? s->impl.copy_original(p, cap) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsCopyInternalPositions(const KhsSolver *s, KhsVec3 *p, uint32_t cap)
{
    return s ? s->impl.copy_internal(p, cap) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsCopyOriginalToInternalMap(const KhsSolver *s, uint32_t *idx, uint32_t cap)
{
    return s ? s->impl.copy_mapping(idx, cap) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsGetBuildStats(const KhsSolver *s, KhsBuildStats *st)
{
    return s ? s->impl.get_build_stats(st) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsGetLastStepStats(const KhsSolver *s, KhsStepStats *st)
{
    return s ? s->impl.get_step_stats(st) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsGetGpuStats(const KhsSolver *s, KhsGpuStats *st)
{
    return s ? s->impl.get_gpu_stats(st) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsGetProgress(const KhsSolver *s, KhsProgress *p)
{
    return s ? s->impl.get_progress(p) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsGetFailureDiagnostics(const KhsSolver *s, KhsFailureDiagnostics *d)
{
    return s ? s->impl.get_failure(d) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsRequestCancel(KhsSolver *s) { return guarded(s, [&] { return s->impl.request_cancel(); }); }

const char *khsGetLastError(const KhsSolver *s)
{
    static const char *none = "No solver.";
    return s ? s->impl.error.c_str() : none;
}

uint64_t khsxGetDofCount(const KhsSolver *s) { return s ? s->impl.dof_count() : 0; }

KhsResult khsxGetState(const KhsSolver *s, double *q, double *v, uint64_t count)
{
    return s ? s->impl.get_state(q, v, count) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxSetState(KhsSolver *s, const double *q, const double *q_old, const double *v, uint64_t count)
{
    return guarded(s, [&] { return s->impl.set_state(q, q_old, v, count); });
}

KhsResult khsxCopyNodeModel(const KhsSolver *s, double *f, double *m, double *rm, uint32_t *flags, uint32_t cap)
{
    return s ? s->impl.copy_node_model(f, m, rm, flags, cap) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxCopyElementModel(const KhsSolver *s, uint32_t *n2, uint32_t *st, uint32_t *c, double *r7, uint32_t cap)
{
    return s ? s->impl.copy_element_model(n2, st, c, r7, cap) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxEvaluate(KhsSolver *s, double dt, double *g, double *d, KhsxEvaluation *out)
{
    return guarded(s, [&] { return s->impl.test_evaluate(dt, g, d, out); });
}

KhsResult khsxBroadphasePairs(KhsSolver *s, uint32_t *pairs, uint64_t cap, uint64_t *count)
{
    return guarded(s, [&] { return s->impl.broadphase_pairs(pairs, cap, count); });
}

KhsResult khsxCopyBlocks(const KhsSolver *s, double *diagonal, double *upper, uint32_t capacity)
{
    return s ? s->impl.copy_blocks(diagonal, upper, capacity) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxCheckBroadphase(KhsSolver *s, uint64_t *violations)
{
    if (!s || !violations) return KHS_ERROR_INVALID_ARGUMENT;
    for (int k = 0; k < 6; ++k) violations[k] = 0;
    return KHS_OK;
}

KhsResult khsxSetConvergenceMode(KhsSolver *s, uint32_t mode)
{
    if (!s || mode > 1) return KHS_ERROR_INVALID_ARGUMENT;
    s->impl.convergence_mode = mode;
    return KHS_OK;
}

KhsResult khsxSetPrecisionMode(KhsSolver *s, uint32_t mode)
{
    if (!s || mode > 1) return KHS_ERROR_INVALID_ARGUMENT;
    s->impl.precision_mode = mode;
    return KHS_OK;
}

KhsResult khsxSetKernelTiming(KhsSolver *s, uint32_t enable)
{
    (void)enable;
    return s ? KHS_OK : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxGetKernelTiming(const KhsSolver *s, KhsxKernelTiming *t)
{
    if (!s || !t || t->struct_size < sizeof(KhsxKernelTiming)) return KHS_ERROR_INVALID_ARGUMENT;
    KhsxKernelTiming r{};
    r.struct_size = sizeof(r);
    r.nodes = s->impl.node_count();
    *t = r;
    return KHS_OK;
}

KhsResult khsxGetSubstepTimes(const KhsSolver *s, double *ms, uint32_t capacity, uint32_t *count)
{
    (void)ms;
    (void)capacity;
    if (!s || !count) return KHS_ERROR_INVALID_ARGUMENT;
    *count = 0;
    return KHS_OK;
}

KhsResult khsxGetContactListStats(KhsSolver *s, KhsxContactListStats *st)
{
    if (!s || !st || st->struct_size < sizeof(KhsxContactListStats)) return KHS_ERROR_INVALID_ARGUMENT;
    KhsxContactListStats r{};
    r.struct_size = sizeof(r);
    r.elements = s->impl.node_count();
    *st = r;
    return KHS_OK;
}

KhsResult khsxSetRollbackRing(KhsSolver *s, uint32_t frames)
{
    if (!s) return KHS_ERROR_INVALID_ARGUMENT;
    return s->impl.set_rollback_ring(frames);
}

KhsResult khsxRollback(KhsSolver *s, uint32_t frame)
{
    if (!s) return KHS_ERROR_INVALID_ARGUMENT;
    return s->impl.rollback(frame);
}

KhsResult khsxWriteCacheFrame(KhsSolver *s, int fd)
{
    if (!s) return KHS_ERROR_INVALID_ARGUMENT;
    return s->impl.write_cache_frame(fd);
}

KhsResult khsxSetThreadCount(KhsSolver *s, uint32_t threads)
{
    if (!s) return KHS_ERROR_INVALID_ARGUMENT;
    s->impl.threads = threads == 0 ? 1 : threads;
    return KHS_OK;
}

KhsResult khsxBroadphasePairsNoAlloc(KhsSolver *s, uint32_t *pairs, uint64_t cap, uint64_t *count)
{
    if (cap > KHSX_NO_ALLOC_PAIRS) return KHS_ERROR_INVALID_ARGUMENT;
    return guarded(s, [&] { return s->impl.broadphase_pairs(pairs, cap, count); });
}

KhsResult khsxSetColliderState(KhsSolver *s, const double *current, const double *previous, uint32_t count)
{
    return guarded(s, [&] { return s->impl.set_collider_state(current, previous, count); });
}

KhsResult khsxApplyFriction(KhsSolver *s, double dt)
{
    return guarded(s, [&] { return s->impl.friction_stage(dt); });
}

KhsResult khsxStepBound(KhsSolver *s, const double *direction, uint64_t count, double *alpha)
{
    return guarded(s, [&] { return s->impl.step_bound(direction, count, alpha); });
}

KhsResult khsxSweepCheck(KhsSolver *s, uint32_t *feasible, KhsFailureDiagnostics *diagnostics)
{
    return guarded(s, [&] { return s->impl.sweep_check(feasible, diagnostics); });
}

KhsResult khsxSetStrandMaterials(KhsSolver *s, const double *young_modulus, const double *radius, uint32_t count)
{
    return guarded(s, [&] { return s->impl.set_strand_materials(young_modulus, radius, count); });
}

KhsResult khsxSetSolveAlgorithm(KhsSolver *s, uint32_t algorithm)
{
    if (!s || algorithm > 2) return KHS_ERROR_INVALID_ARGUMENT;
    s->impl.solve_algorithm = algorithm;
    return KHS_OK;
}

KhsResult khsxSetContactReuse(KhsSolver *s, uint32_t mode)
{
    // no candidate lists here: both modes are the plain step
    if (!s || mode > 1) return KHS_ERROR_INVALID_ARGUMENT;
    return KHS_OK;
}

KhsResult khsxSetDirectionMode(KhsSolver *s, uint32_t mode)
{
    if (!s || mode > 1) return KHS_ERROR_INVALID_ARGUMENT;
    s->impl.direction_mode = mode;
    return KHS_OK;
}

KhsResult khsxGetDirectionCounts(const KhsSolver *s, uint64_t *block, uint64_t *steepest)
{
    return s ? s->impl.direction_counts(block, steepest) : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxGetLaunchCounts(const KhsSolver *s, uint64_t *dispatches, uint64_t *host_waits)
{
    if (!s || !dispatches || !host_waits) return KHS_ERROR_INVALID_ARGUMENT;
    *dispatches = *host_waits = 0;  // no device
    return KHS_OK;
}

}  // extern "C"
