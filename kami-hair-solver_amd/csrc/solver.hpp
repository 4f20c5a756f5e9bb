// The hair solver behind the C ABI: host model build plus the device-resident
// incremental-potential step on one MI355X (HIP, gfx950).
#pragma once

#include "kami_hair_solver.h"
#include "kami_hair_solver_ext.h"
#include "model.hpp"

#include <atomic>
#include <chrono>
#include <memory>
#include <string>
#include <vector>

namespace khs {

struct DeviceState;  // device buffers, defined in solver.hip
struct Partial;      // per-block reduction record of the kernels
struct StepParams;   // scalars of one substep passed to the kernels

class Solver {
public:
    explicit Solver(const KhsSolverDesc &desc);
    ~Solver();

    KhsResult set_hair(const KhsVec3 *points, uint32_t point_count, const uint32_t *offsets, uint32_t strand_count,
                       const KhsHairMaterial &material);
    KhsResult set_collider(const KhsVec3 *vertices, uint32_t vertex_count, const KhsTriangle *triangles,
                           uint32_t triangle_count);
    KhsResult build();
    KhsResult update_runtime_parameters(const KhsSolverDesc &desc, const KhsHairMaterial &material);
    KhsResult update_collider(const KhsVec3 *vertices, uint32_t vertex_count);
    KhsResult update_roots(const KhsVec3 *points, uint32_t point_count);
    KhsResult step(double frame_dt);
    KhsResult allocate_animation(uint32_t frame_count);
    KhsResult set_root_animation_frame(uint32_t frame, const KhsVec3 *points, uint32_t point_count);
    KhsResult set_collider_animation_frame(uint32_t frame, const KhsVec3 *vertices, uint32_t vertex_count);
    KhsResult finalize_animation();
    KhsResult step_animation_frame(uint32_t frame, double frame_dt);
    uint64_t checkpoint_size() const;
    KhsResult save_checkpoint(void *data, uint64_t capacity);
    KhsResult restore_checkpoint(const void *data, uint64_t size);

    uint32_t original_point_count() const { return uint32_t(points_.size()); }
    uint32_t internal_node_count() const { return built_ ? model_.node_count : 0; }
    KhsResult copy_original_positions(KhsVec3 *positions, uint32_t capacity) const;
    KhsResult copy_internal_positions(KhsVec3 *positions, uint32_t capacity) const;
    KhsResult copy_mapping(uint32_t *indices, uint32_t capacity) const;
    KhsResult build_stats(KhsBuildStats *stats) const;
    KhsResult step_stats(KhsStepStats *stats) const;
    KhsResult gpu_stats(KhsGpuStats *stats) const;
    KhsResult progress(KhsProgress *progress) const;
    KhsResult failure_diagnostics(KhsFailureDiagnostics *diagnostics) const;
    KhsResult request_cancel();

    // additive hooks (kami_hair_solver_ext.h)
    uint64_t dof_count() const { return built_ ? uint64_t(model_.node_count) * 6 : 0; }
    KhsResult get_state(double *q, double *v, uint64_t count) const;
    KhsResult set_state(const double *q, const double *q_old, const double *v, uint64_t count);
    KhsResult copy_node_model(double *frames9, double *mass, double *rotational_mass, uint32_t *flags,
                              uint32_t capacity) const;
    KhsResult copy_element_model(uint32_t *nodes2, uint32_t *strand, uint32_t *contact, double *rest7,
                                 uint32_t capacity) const;
    KhsResult test_evaluate(double dt, double *gradient, double *direction, KhsxEvaluation *out);
    KhsResult broadphase_pairs(uint32_t *pairs, uint64_t capacity, uint64_t *count);
    KhsResult broadphase_pairs_no_alloc(uint32_t *pairs, uint64_t capacity, uint64_t *count);
    KhsResult check_bvh(uint64_t *violations);
    KhsResult copy_blocks(double *diagonal, double *upper, uint32_t capacity) const;
    KhsResult set_convergence_mode(uint32_t mode);
    KhsResult set_precision_mode(uint32_t mode);
    KhsResult set_solve_algorithm(uint32_t algorithm);
    KhsResult set_contact_reuse(uint32_t mode);
    KhsResult set_kernel_timing(uint32_t enable);
    KhsResult kernel_timing(KhsxKernelTiming *out) const;
    KhsResult substep_times(double *ms, uint32_t capacity, uint32_t *count) const;
    KhsResult contact_list_stats(KhsxContactListStats *out);
    KhsResult set_rollback_ring(uint32_t frames);
    KhsResult rollback(uint32_t frame);
    KhsResult write_cache_frame(int fd);
    KhsResult set_collider_state(const double *current, const double *previous, uint32_t count);
    KhsResult set_strand_materials(const double *young_modulus, const double *radius, uint32_t count);
    KhsResult friction_stage(double dt);
    KhsResult step_bound(const double *direction, uint64_t count, double *alpha);
    KhsResult sweep_check(uint32_t *feasible, KhsFailureDiagnostics *diagnostics);
    KhsResult set_direction_mode(uint32_t mode);
    KhsResult direction_counts(uint64_t *block, uint64_t *steepest) const;
    KhsResult launch_counts(uint64_t *dispatches, uint64_t *host_waits) const;

    const std::string &last_error() const { return error_; }
    void set_error(std::string message) { error_ = std::move(message); }

private:
    struct Reduced;
    KhsResult fail(KhsResult result, const std::string &message);
    void clear_failure();
    void set_failure_context(uint32_t kind, uint32_t frame, uint32_t substep, uint32_t attempted, uint32_t maximum,
                             uint32_t attempts);
    KhsResult upload_model();
    bool device_root_targets(const std::vector<Vec3> &points, double *pos, double *rot);
    KhsResult initial_feasibility();
    void release_device();
    void restore_state(uint32_t frame);
    KhsResult build_model_on_device();
    void keep_in_ring(uint32_t frame);
    void rebuild_bvh(bool swept);
    void refit_bvh(bool swept);
    void refit_boxes();
    void refresh_contact_cache(int mode, bool full, bool swept_leaves = false, bool with_roots = false);
    void closest_pass(bool line_search, unsigned long long *block_candidates, bool with_roots = false);
    void swept_bound_async(const StepParams &cp, const unsigned *stop);
    // With `stop`, the gradient's reduction also sets *stop (device) when
    // |g| <= abs_tol or |g| / initial <= rel_tol (initial < 0: this one).
    void evaluate_async(bool gradient, double dt, bool line_search, unsigned *stop = nullptr, double abs_tol = 0.0,
                        double rel_tol = 0.0, double initial = -1.0);
    static Reduced unpack(const Partial &r);
    Reduced evaluate(bool gradient, double dt, bool line_search);
    void make_direction_async();
    uint32_t launch_block_solve(unsigned *failed, unsigned *next_failed, const unsigned *stop);
    // The direction of the next Newton iteration by the direction rule:
    // the block solve (mode 0) or forced steepest descent (mode 1).
    void direction_async();
    void launch_preconditioner(double dt);
    double max_radius() const;
    void apply_strand_materials();
    StepParams step_params(double dt) const;
    // The moving-collider sweep of a substep: swept tree, full list refresh
    // and the sweep kernels (reads nothing back).
    void enqueue_sweep(const StepParams &sp);
    void enqueue_friction(const StepParams &sp);
    void set_trial_step(int source, double value);
    void steepest_direction_async();
    KhsResult optimize_substep(double dt);
    KhsResult run_frame(double frame_dt, const double *root_a, const double *rot_a, const double *root_b,
                        const double *rot_b, const double *collider_a, const double *collider_b, uint32_t frame,
                        uint32_t frame_count);

    KhsSolverDesc desc_{};
    KhsHairMaterial material_{};
    bool has_hair_ = false, has_collider_ = false, built_ = false;
    std::string error_;
    std::vector<Vec3> points_;
    std::vector<uint32_t> offsets_;
    ColliderMesh collider_;
    std::vector<Vec3> collider_pending_host_;
    RodModel model_;
    KhsBuildStats build_stats_{};
    KhsStepStats step_stats_{};
    KhsGpuStats gpu_stats_{};
    KhsFailureDiagnostics failure_{};
    std::unique_ptr<DeviceState> dev_;
    uint32_t convergence_mode_ = 0;
    uint32_t precision_mode_ = 0;
    bool contact_reuse_ = true;     // khsxSetContactReuse
    uint32_t solve_algorithm_ = 0;  // 0 block Thomas, 1 twisted (FP64 only)
    uint32_t direction_mode_ = 0;
    uint64_t block_directions_ = 0, steepest_directions_ = 0;
    std::vector<double> strand_E_, strand_r_;  // per-strand materials (empty: the material's)
    double strand_r_max_ = 0.0;                 // the largest of strand_r_
    double trust_radius_ = 0.0;
    uint32_t animation_frames_ = 0, current_frame_ = 0;
    std::vector<uint32_t> ring_frame_;  // animation frame held by each rollback slot (~0u: empty)
    bool animation_ready_ = false;
    std::vector<bool> root_frame_set_, collider_frame_set_;
    std::atomic<bool> cancel_{false};
    std::atomic<uint32_t> phase_{KHS_PHASE_IDLE}, frame_{0}, frame_count_{0}, substep_{0}, substep_count_{0},
        iteration_{0};
    std::atomic<int64_t> frame_clock_{0};
};

bool query_device(KhsGpuInfo &info, std::string &error);

}  // namespace khs
