// C ABI v5 of the MI355X hair solver (include/kami_hair_solver.h) and the
// additive khsx entry points (include/kami_hair_solver_ext.h). Argument
// checks, defaults and exception mapping follow the reference shim
// (src/api.cpp:16-304).
#include "abi_layout.hpp"
#include "solver.hpp"

#include <cmath>
#include <cstring>
#include <new>

struct KhsSolver {
    explicit KhsSolver(const KhsSolverDesc &desc) : impl(desc) {}
    khs::Solver impl;
};

namespace {

template <class F>
KhsResult guarded(KhsSolver *solver, F &&f)
{
    if (!solver) return KHS_ERROR_INVALID_ARGUMENT;
    try {
        return f();
    } catch (const std::bad_alloc &) {
        solver->impl.set_error("Out of host memory.");
        return KHS_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        solver->impl.set_error(std::string("Internal error: ") + e.what());
        return KHS_ERROR_INTERNAL;
    } catch (...) {
        solver->impl.set_error("Unknown internal error.");
        return KHS_ERROR_INTERNAL;
    }
}

template <class F>
KhsResult guarded_const(const KhsSolver *solver, F &&f)
{
    if (!solver) return KHS_ERROR_INVALID_ARGUMENT;
    try {
        return f();
    } catch (...) {
        return KHS_ERROR_INTERNAL;
    }
}

bool valid_desc(const KhsSolverDesc *d)
{
    return d && d->struct_size >= sizeof(KhsSolverDesc) && d->substeps > 0 && d->maximum_substeps >= d->substeps &&
           d->newton_iterations > 0 && d->line_search_iterations > 0 && d->absolute_tolerance > 0.0 &&
           d->relative_tolerance > 0.0 && d->increment_tolerance > 0.0 && d->minimum_line_search_step > 0.0 &&
           d->minimum_gap > 0.0 && d->maximum_element_length > 0.0 && d->minimum_dynamic_length >= 0.0 &&
           std::isfinite(d->minimum_dynamic_length) && d->fixed_root_nodes > 0;
}

bool valid_material(const KhsHairMaterial *m)
{
    return m && m->struct_size >= sizeof(KhsHairMaterial) && m->density > 0.0 && m->radius > 0.0 &&
           m->young_modulus > 0.0 && m->poisson_ratio > -1.0 && m->poisson_ratio < 0.5 && m->shear_correction > 0.0 &&
           m->mass_damping >= 0.0 && m->contact_stiffness > 0.0 && m->barrier_distance > 0.0 && m->friction >= 0.0 &&
           m->friction_smoothing > 0.0 && m->collider_offset >= 0.0;
}

}  // namespace

extern "C" {

uint32_t khsGetAbiVersion(void) { return KHS_ABI_VERSION; }

KhsResult khsGetGpuInfo(KhsGpuInfo *info)
{
    if (!info || info->struct_size < sizeof(KhsGpuInfo)) return KHS_ERROR_INVALID_ARGUMENT;
    std::string error;
    KhsGpuInfo result{};
    const bool ok = khs::query_device(result, error);
    *info = result;
    return ok ? KHS_OK : KHS_ERROR_INVALID_STATE;
}

void khsDefaultSolverDesc(KhsSolverDesc *d)
{
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
    if (!valid_desc(desc)) return nullptr;
    try {
        return new KhsSolver(*desc);
    } catch (...) {
        return nullptr;
    }
}

void khsDestroy(KhsSolver *solver) { delete solver; }

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
    if (!valid_desc(d) || !valid_material(m)) return KHS_ERROR_INVALID_ARGUMENT;
    return guarded(s, [&] { return s->impl.update_runtime_parameters(*d, *m); });
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
    return guarded(s, [&] { return s->impl.set_root_animation_frame(f, p, c); });
}

KhsResult khsSetColliderAnimationFrame(KhsSolver *s, uint32_t f, const KhsVec3 *v, uint32_t c)
{
    return guarded(s, [&] { return s->impl.set_collider_animation_frame(f, v, c); });
}

KhsResult khsFinalizeAnimation(KhsSolver *s) { return guarded(s, [&] { return s->impl.finalize_animation(); }); }

KhsResult khsStepAnimationFrame(KhsSolver *s, uint32_t f, double dt)
{
    return guarded(s, [&] { return s->impl.step_animation_frame(f, dt); });
}

uint64_t khsGetAnimationCheckpointSize(const KhsSolver *s) { return s ? s->impl.checkpoint_size() : 0; }

KhsResult khsSaveAnimationCheckpoint(KhsSolver *s, void *data, uint64_t capacity)
{
    return guarded(s, [&] { return s->impl.save_checkpoint(data, capacity); });
}

KhsResult khsRestoreAnimationCheckpoint(KhsSolver *s, const void *data, uint64_t size)
{
    return guarded(s, [&] { return s->impl.restore_checkpoint(data, size); });
}

uint32_t khsGetOriginalPointCount(const KhsSolver *s) { return s ? s->impl.original_point_count() : 0; }
uint32_t khsGetInternalNodeCount(const KhsSolver *s) { return s ? s->impl.internal_node_count() : 0; }

KhsResult khsCopyOriginalPositions(const KhsSolver *s, KhsVec3 *p, uint32_t capacity)
{
    return guarded_const(s, [&] { return s->impl.copy_original_positions(p, capacity); });
}

KhsResult khsCopyInternalPositions(const KhsSolver *s, KhsVec3 *p, uint32_t capacity)
{
    return guarded_const(s, [&] { return s->impl.copy_internal_positions(p, capacity); });
}

KhsResult khsCopyOriginalToInternalMap(const KhsSolver *s, uint32_t *indices, uint32_t capacity)
{
    return guarded_const(s, [&] { return s->impl.copy_mapping(indices, capacity); });
}

KhsResult khsGetBuildStats(const KhsSolver *s, KhsBuildStats *stats)
{
    return guarded_const(s, [&] { return s->impl.build_stats(stats); });
}

KhsResult khsGetLastStepStats(const KhsSolver *s, KhsStepStats *stats)
{
    return guarded_const(s, [&] { return s->impl.step_stats(stats); });
}

KhsResult khsGetGpuStats(const KhsSolver *s, KhsGpuStats *stats)
{
    return guarded_const(s, [&] { return s->impl.gpu_stats(stats); });
}

KhsResult khsGetProgress(const KhsSolver *s, KhsProgress *progress)
{
    return guarded_const(s, [&] { return s->impl.progress(progress); });
}

KhsResult khsGetFailureDiagnostics(const KhsSolver *s, KhsFailureDiagnostics *d)
{
    return guarded_const(s, [&] { return s->impl.failure_diagnostics(d); });
}

KhsResult khsRequestCancel(KhsSolver *s) { return guarded(s, [&] { return s->impl.request_cancel(); }); }

const char *khsGetLastError(const KhsSolver *s)
{
    static const char *none = "No solver.";
    return s ? s->impl.last_error().c_str() : none;
}

uint64_t khsxGetDofCount(const KhsSolver *s) { return s ? s->impl.dof_count() : 0; }

KhsResult khsxGetState(const KhsSolver *s, double *q, double *v, uint64_t count)
{
    return guarded_const(s, [&] { return s->impl.get_state(q, v, count); });
}

KhsResult khsxSetState(KhsSolver *s, const double *q, const double *q_old, const double *v, uint64_t count)
{
    return guarded(s, [&] { return s->impl.set_state(q, q_old, v, count); });
}

KhsResult khsxCopyNodeModel(const KhsSolver *s, double *f, double *m, double *rm, uint32_t *flags, uint32_t cap)
{
    return guarded_const(s, [&] { return s->impl.copy_node_model(f, m, rm, flags, cap); });
}

KhsResult khsxCopyElementModel(const KhsSolver *s, uint32_t *n2, uint32_t *st, uint32_t *c, double *r7, uint32_t cap)
{
    return guarded_const(s, [&] { return s->impl.copy_element_model(n2, st, c, r7, cap); });
}

KhsResult khsxEvaluate(KhsSolver *s, double dt, double *g, double *d, KhsxEvaluation *out)
{
    return guarded(s, [&] { return s->impl.test_evaluate(dt, g, d, out); });
}

KhsResult khsxBroadphasePairs(KhsSolver *s, uint32_t *pairs, uint64_t capacity, uint64_t *count)
{
    return guarded(s, [&] { return s->impl.broadphase_pairs(pairs, capacity, count); });
}

KhsResult khsxBroadphasePairsNoAlloc(KhsSolver *s, uint32_t *pairs, uint64_t capacity, uint64_t *count)
{
    return guarded(s, [&] { return s->impl.broadphase_pairs_no_alloc(pairs, capacity, count); });
}

KhsResult khsxCopyBlocks(const KhsSolver *s, double *diagonal, double *upper, uint32_t capacity)
{
    return guarded_const(s, [&] { return s->impl.copy_blocks(diagonal, upper, capacity); });
}

KhsResult khsxCheckBroadphase(KhsSolver *s, uint64_t *violations)
{
    return guarded(s, [&] { return s->impl.check_bvh(violations); });
}

KhsResult khsxSetConvergenceMode(KhsSolver *s, uint32_t mode)
{
    return guarded(s, [&] { return s->impl.set_convergence_mode(mode); });
}

KhsResult khsxSetPrecisionMode(KhsSolver *s, uint32_t mode)
{
    return guarded(s, [&] { return s->impl.set_precision_mode(mode); });
}

KhsResult khsxSetSolveAlgorithm(KhsSolver *s, uint32_t algorithm)
{
    return guarded(s, [&] { return s->impl.set_solve_algorithm(algorithm); });
}

KhsResult khsxSetContactReuse(KhsSolver *s, uint32_t mode)
{
    return guarded(s, [&] { return s->impl.set_contact_reuse(mode); });
}

KhsResult khsxSetKernelTiming(KhsSolver *s, uint32_t enable)
{
    return guarded(s, [&] { return s->impl.set_kernel_timing(enable); });
}

KhsResult khsxGetKernelTiming(const KhsSolver *s, KhsxKernelTiming *t)
{
    return guarded_const(s, [&] { return s->impl.kernel_timing(t); });
}

KhsResult khsxGetSubstepTimes(const KhsSolver *s, double *ms, uint32_t capacity, uint32_t *count)
{
    return guarded_const(s, [&] { return s->impl.substep_times(ms, capacity, count); });
}

KhsResult khsxGetContactListStats(KhsSolver *s, KhsxContactListStats *st)
{
    return guarded(s, [&] { return s->impl.contact_list_stats(st); });
}

KhsResult khsxSetRollbackRing(KhsSolver *s, uint32_t frames)
{
    return guarded(s, [&] { return s->impl.set_rollback_ring(frames); });
}

KhsResult khsxRollback(KhsSolver *s, uint32_t frame)
{
    return guarded(s, [&] { return s->impl.rollback(frame); });
}

KhsResult khsxWriteCacheFrame(KhsSolver *s, int fd)
{
    return guarded(s, [&] { return s->impl.write_cache_frame(fd); });
}

KhsResult khsxSetThreadCount(KhsSolver *s, uint32_t threads)
{
    (void)threads;
    return s ? KHS_OK : KHS_ERROR_INVALID_ARGUMENT;
}

KhsResult khsxSetStrandMaterials(KhsSolver *s, const double *young_modulus, const double *radius, uint32_t count)
{
    return guarded(s, [&] { return s->impl.set_strand_materials(young_modulus, radius, count); });
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

KhsResult khsxSetDirectionMode(KhsSolver *s, uint32_t mode)
{
    return guarded(s, [&] { return s->impl.set_direction_mode(mode); });
}

KhsResult khsxGetDirectionCounts(const KhsSolver *s, uint64_t *block, uint64_t *steepest)
{
    return guarded_const(s, [&] { return s->impl.direction_counts(block, steepest); });
}

KhsResult khsxGetLaunchCounts(const KhsSolver *s, uint64_t *dispatches, uint64_t *host_waits)
{
    return guarded_const(s, [&] { return s->impl.launch_counts(dispatches, host_waits); });
}

}  // extern "C"
