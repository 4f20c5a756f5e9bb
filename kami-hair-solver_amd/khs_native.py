"""ctypes binding of the kami hair solver C ABI v5 plus the additive khsx hooks.

The struct mirrors follow include/kami_hair_solver.h (the reference ABI,
include/kami_hair_solver.h:25-204 of ysk424/kami-hair-solver) and the
reference's own binding (blender_extension/kami_hair_solver/native.py:13-171);
the HairSolver class offers the same operations as the reference's
HairSolver (native.py:180-413) with numpy arrays in and out.

One binding serves two libraries with the same exports: the HIP library
(kami-hair-solver_amd/libkami_hair_solver.so) and the CPU oracle used by
the tests as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PACKAGE_DIR)
GPU_LIBRARY = os.path.join(PACKAGE_DIR, "libkami_hair_solver.so")

ABI_VERSION = 5

KHS_OK = 0
KHS_ERROR_INVALID_ARGUMENT = 1
KHS_ERROR_INVALID_STATE = 2
KHS_ERROR_INVALID_HAIR = 3
KHS_ERROR_INVALID_COLLIDER = 4
KHS_ERROR_INITIAL_INTERSECTION = 5
KHS_ERROR_NOT_CONVERGED = 6
KHS_ERROR_NUMERICAL_FAILURE = 7
KHS_ERROR_OUT_OF_MEMORY = 8
KHS_ERROR_INTERNAL = 9

PHASE_IDLE, PHASE_PREPARING, PHASE_ASSEMBLING, PHASE_LINEAR_SOLVE = 0, 1, 2, 3
PHASE_CCD, PHASE_LINE_SEARCH, PHASE_FINISHED, PHASE_FAILED = 4, 5, 6, 7

FAILURE_NONE, FAILURE_SWEEP, FAILURE_BARRIER_INFEASIBLE, FAILURE_NONLINEAR = 0, 1, 2, 3

KHSX_NO_ALLOC_PAIRS = 1048576  # staging limit of khsxBroadphasePairsNoAlloc


class KhsVec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class KhsTriangle(C.Structure):
    _fields_ = [("i0", C.c_uint32), ("i1", C.c_uint32), ("i2", C.c_uint32)]


class KhsSolverDesc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("gravity", KhsVec3),
        ("substeps", C.c_uint32),
        ("maximum_substeps", C.c_uint32),
        ("newton_iterations", C.c_uint32),
        ("line_search_iterations", C.c_uint32),
        ("absolute_tolerance", C.c_double),
        ("relative_tolerance", C.c_double),
        ("increment_tolerance", C.c_double),
        ("minimum_line_search_step", C.c_double),
        ("minimum_gap", C.c_double),
        ("maximum_element_length", C.c_double),
        ("minimum_dynamic_length", C.c_double),
        ("fixed_root_nodes", C.c_uint32),
        ("thread_count", C.c_uint32),
    ]


class KhsHairMaterial(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("density", C.c_double),
        ("radius", C.c_double),
        ("young_modulus", C.c_double),
        ("poisson_ratio", C.c_double),
        ("shear_correction", C.c_double),
        ("mass_damping", C.c_double),
        ("contact_stiffness", C.c_double),
        ("barrier_distance", C.c_double),
        ("friction", C.c_double),
        ("friction_smoothing", C.c_double),
        ("collider_offset", C.c_double),
    ]


class KhsBuildStats(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("strand_count", C.c_uint32),
        ("original_point_count", C.c_uint32),
        ("internal_node_count", C.c_uint32),
        ("element_count", C.c_uint32),
        ("fixed_node_count", C.c_uint32),
        ("collider_vertex_count", C.c_uint32),
        ("collider_triangle_count", C.c_uint32),
        ("excluded_collider_triangle_count", C.c_uint32),
        ("collider_boundary_edge_count", C.c_uint32),
        ("collider_nonmanifold_edge_count", C.c_uint32),
        ("collider_inconsistent_edge_count", C.c_uint32),
        ("collider_inverted_closed_component_count", C.c_uint32),
        ("merged_zero_length_segment_count", C.c_uint32),
        ("virtual_extension_strand_count", C.c_uint32),
        ("virtual_extension_node_count", C.c_uint32),
        ("virtual_extension_rest_length", C.c_double),
        ("degree_of_freedom_count", C.c_uint64),
        ("estimated_bytes", C.c_uint64),
        ("initial_minimum_gap", C.c_double),
    ]


class KhsStepStats(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("substeps", C.c_uint32),
        ("converged_substeps", C.c_uint32),
        ("newton_iterations", C.c_uint32),
        ("linear_solves", C.c_uint32),
        ("line_search_evaluations", C.c_uint32),
        ("contact_candidate_count", C.c_uint64),
        ("active_contact_count", C.c_uint64),
        ("initial_residual_norm", C.c_double),
        ("final_residual_norm", C.c_double),
        ("relative_residual_norm", C.c_double),
        ("increment_norm", C.c_double),
        ("objective_change", C.c_double),
        ("minimum_gap", C.c_double),
        ("accepted_step_length", C.c_double),
        ("ccd_step_limit", C.c_double),
        ("kinetic_energy", C.c_double),
        ("elastic_energy", C.c_double),
        ("contact_energy", C.c_double),
        ("friction_energy", C.c_double),
        ("phase", C.c_int),
    ]


class KhsGpuInfo(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("available", C.c_uint32),
        ("device_ordinal", C.c_int32),
        ("compute_capability_major", C.c_uint32),
        ("compute_capability_minor", C.c_uint32),
        ("total_vram_bytes", C.c_uint64),
        ("device_name", C.c_char * 128),
    ]


class KhsGpuStats(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("animation_frame_count", C.c_uint32),
        ("resident_bytes", C.c_uint64),
        ("peak_temporary_bytes", C.c_uint64),
        ("last_frame_milliseconds", C.c_double),
        ("last_assembly_milliseconds", C.c_double),
        ("last_collision_milliseconds", C.c_double),
        ("last_optimization_milliseconds", C.c_double),
    ]


class KhsProgress(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("phase", C.c_uint32),
        ("frame_index", C.c_uint32),
        ("frame_count", C.c_uint32),
        ("substep", C.c_uint32),
        ("substep_count", C.c_uint32),
        ("nonlinear_iteration", C.c_uint32),
        ("nonlinear_iteration_limit", C.c_uint32),
        ("cancelled", C.c_uint32),
        ("frame_elapsed_seconds", C.c_double),
    ]


class KhsFailureDiagnostics(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("kind", C.c_uint32),
        ("frame_index", C.c_uint32),
        ("substep", C.c_uint32),
        ("requested_substeps", C.c_uint32),
        ("attempted_substeps", C.c_uint32),
        ("maximum_substeps", C.c_uint32),
        ("adaptive_attempt_count", C.c_uint32),
        ("strand_index", C.c_uint32),
        ("element_index", C.c_uint32),
        ("collider_triangle_index", C.c_uint32),
        ("reserved", C.c_uint32),
        ("distance", C.c_double),
        ("required_distance", C.c_double),
        ("clearance", C.c_double),
        ("collider_substep_displacement", C.c_double),
        ("collider_frame_displacement", C.c_double),
        ("hair_start", KhsVec3),
        ("hair_end", KhsVec3),
        ("collider_point", KhsVec3),
    ]


class KhsxEvaluation(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("feasible", C.c_uint32),
        ("solve_failed", C.c_uint32),
        ("reserved", C.c_uint32),
        ("candidate_pairs", C.c_uint64),
        ("active_contacts", C.c_uint64),
        ("objective", C.c_double),
        ("elastic", C.c_double),
        ("contact", C.c_double),
        ("kinetic", C.c_double),
        ("minimum_gap", C.c_double),
    ]


class KhsxKernelTiming(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("substep_count", C.c_uint32),
        ("assembly_launches", C.c_uint64),
        ("assembly_ms", C.c_double),
        ("energy_launches", C.c_uint64),
        ("energy_ms", C.c_double),
        ("solve_launches", C.c_uint64),
        ("solve_ms", C.c_double),
        ("collision_launches", C.c_uint64),
        ("collision_ms", C.c_double),
        ("nodes", C.c_uint64),
    ]


class KhsxContactListStats(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("capacity", C.c_uint32),
        ("elements", C.c_uint64),
        ("listed", C.c_uint64),
        ("nonempty", C.c_uint64),
        ("overflowed", C.c_uint64),
        ("entries", C.c_uint64),
        ("longest", C.c_uint64),
        ("length_histogram", C.c_uint64 * 9),
        ("root_nonempty", C.c_uint64),
        ("root_entries", C.c_uint64),
    ]


# Every exported function with its ctypes signature. The first 33 are the
# reference ABI (include/kami_hair_solver.h:206-298); khsx* are additive.
_P = C.c_void_p
_D = C.POINTER
ABI_FUNCTIONS = {
    "khsGetAbiVersion": (C.c_uint32, []),
    "khsGetGpuInfo": (C.c_int, [_D(KhsGpuInfo)]),
    "khsDefaultSolverDesc": (None, [_D(KhsSolverDesc)]),
    "khsDefaultHairMaterial": (None, [_D(KhsHairMaterial)]),
    "khsCreate": (_P, [_D(KhsSolverDesc)]),
    "khsDestroy": (None, [_P]),
    "khsSetHairCurves": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32, _D(KhsHairMaterial)]),
    "khsSetColliderMesh": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_uint32]),
    "khsBuild": (C.c_int, [_P]),
    "khsUpdateRuntimeParameters": (C.c_int, [_P, _D(KhsSolverDesc), _D(KhsHairMaterial)]),
    "khsUpdateColliderVertices": (C.c_int, [_P, _P, C.c_uint32]),
    "khsUpdateRootTargets": (C.c_int, [_P, _P, C.c_uint32]),
    "khsStep": (C.c_int, [_P, C.c_double]),
    "khsAllocateAnimation": (C.c_int, [_P, C.c_uint32]),
    "khsSetRootAnimationFrame": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "khsSetColliderAnimationFrame": (C.c_int, [_P, C.c_uint32, _P, C.c_uint32]),
    "khsFinalizeAnimation": (C.c_int, [_P]),
    "khsStepAnimationFrame": (C.c_int, [_P, C.c_uint32, C.c_double]),
    "khsGetAnimationCheckpointSize": (C.c_uint64, [_P]),
    "khsSaveAnimationCheckpoint": (C.c_int, [_P, _P, C.c_uint64]),
    "khsRestoreAnimationCheckpoint": (C.c_int, [_P, _P, C.c_uint64]),
    "khsGetOriginalPointCount": (C.c_uint32, [_P]),
    "khsGetInternalNodeCount": (C.c_uint32, [_P]),
    "khsCopyOriginalPositions": (C.c_int, [_P, _P, C.c_uint32]),
    "khsCopyInternalPositions": (C.c_int, [_P, _P, C.c_uint32]),
    "khsCopyOriginalToInternalMap": (C.c_int, [_P, _P, C.c_uint32]),
    "khsGetBuildStats": (C.c_int, [_P, _D(KhsBuildStats)]),
    "khsGetLastStepStats": (C.c_int, [_P, _D(KhsStepStats)]),
    "khsGetGpuStats": (C.c_int, [_P, _D(KhsGpuStats)]),
    "khsGetProgress": (C.c_int, [_P, _D(KhsProgress)]),
    "khsGetFailureDiagnostics": (C.c_int, [_P, _D(KhsFailureDiagnostics)]),
    "khsRequestCancel": (C.c_int, [_P]),
    "khsGetLastError": (C.c_char_p, [_P]),
}
EXT_FUNCTIONS = {
    "khsxGetDofCount": (C.c_uint64, [_P]),
    "khsxGetState": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "khsxSetState": (C.c_int, [_P, _P, _P, _P, C.c_uint64]),
    "khsxCopyNodeModel": (C.c_int, [_P, _P, _P, _P, _P, C.c_uint32]),
    "khsxCopyElementModel": (C.c_int, [_P, _P, _P, _P, _P, C.c_uint32]),
    "khsxEvaluate": (C.c_int, [_P, C.c_double, _P, _P, _D(KhsxEvaluation)]),
    "khsxBroadphasePairs": (C.c_int, [_P, _P, C.c_uint64, _D(C.c_uint64)]),
    "khsxBroadphasePairsNoAlloc": (C.c_int, [_P, _P, C.c_uint64, _D(C.c_uint64)]),
    "khsxCheckBroadphase": (C.c_int, [_P, _D(C.c_uint64)]),
    "khsxCopyBlocks": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "khsxSetConvergenceMode": (C.c_int, [_P, C.c_uint32]),
    "khsxSetPrecisionMode": (C.c_int, [_P, C.c_uint32]),
    "khsxSetThreadCount": (C.c_int, [_P, C.c_uint32]),
    "khsxSetKernelTiming": (C.c_int, [_P, C.c_uint32]),
    "khsxGetKernelTiming": (C.c_int, [_P, _D(KhsxKernelTiming)]),
    "khsxGetSubstepTimes": (C.c_int, [_P, _P, C.c_uint32, _D(C.c_uint32)]),
    "khsxGetContactListStats": (C.c_int, [_P, _D(KhsxContactListStats)]),
    "khsxSetRollbackRing": (C.c_int, [_P, C.c_uint32]),
    "khsxRollback": (C.c_int, [_P, C.c_uint32]),
    "khsxWriteCacheFrame": (C.c_int, [_P, C.c_int]),
    "khsxSetStrandMaterials": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "khsxSetColliderState": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "khsxApplyFriction": (C.c_int, [_P, C.c_double]),
    "khsxStepBound": (C.c_int, [_P, _P, C.c_uint64, _D(C.c_double)]),
    "khsxSweepCheck": (C.c_int, [_P, _D(C.c_uint32), _D(KhsFailureDiagnostics)]),
    "khsxSetDirectionMode": (C.c_int, [_P, C.c_uint32]),
    "khsxSetSolveAlgorithm": (C.c_int, [_P, C.c_uint32]),
    "khsxSetContactReuse": (C.c_int, [_P, C.c_uint32]),
    "khsxGetDirectionCounts": (C.c_int, [_P, _D(C.c_uint64), _D(C.c_uint64)]),
    "khsxGetLaunchCounts": (C.c_int, [_P, _D(C.c_uint64), _D(C.c_uint64)]),
}


def load_library(path: str) -> C.CDLL:
    """Loads a library exporting the ABI and binds every signature."""
    lib = C.CDLL(path)
    for name, (restype, argtypes) in {**ABI_FUNCTIONS, **EXT_FUNCTIONS}.items():
        fn = getattr(lib, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def default_desc(lib) -> KhsSolverDesc:
    d = KhsSolverDesc()
    lib.khsDefaultSolverDesc(C.byref(d))
    return d


def default_material(lib) -> KhsHairMaterial:
    m = KhsHairMaterial()
    lib.khsDefaultHairMaterial(C.byref(m))
    return m


class SolverError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"khs error {code}: {message}")
        self.code = code


class HairSolver:
    """Owns one KhsSolver handle of a loaded library."""

    def __init__(self, lib, desc: KhsSolverDesc):
        self.lib = lib
        self.desc = desc
        self.handle = lib.khsCreate(C.byref(desc))
        if not self.handle:
            raise SolverError(KHS_ERROR_INVALID_ARGUMENT, "khsCreate rejected the descriptor")

    def close(self):
        if self.handle:
            self.lib.khsDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ helpers
    def last_error(self) -> str:
        msg = self.lib.khsGetLastError(self.handle)
        return msg.decode("utf-8", "replace") if msg else ""

    def check(self, code: int) -> int:
        if code != KHS_OK:
            raise SolverError(code, self.last_error())
        return code

    # -------------------------------------------------------------- setup
    def set_hair(self, points, offsets, material: KhsHairMaterial) -> int:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        o = np.ascontiguousarray(offsets, dtype=np.uint32)
        self._points = p
        return self.lib.khsSetHairCurves(self.handle, _ptr(p), p.shape[0], _ptr(o), o.size - 1, C.byref(material))

    def set_collider(self, vertices, triangles) -> int:
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        return self.lib.khsSetColliderMesh(self.handle, _ptr(v), v.shape[0], _ptr(t), t.shape[0])

    def build(self) -> int:
        return self.lib.khsBuild(self.handle)

    def update_runtime(self, desc: KhsSolverDesc, material: KhsHairMaterial) -> int:
        return self.lib.khsUpdateRuntimeParameters(self.handle, C.byref(desc), C.byref(material))

    def update_collider(self, vertices) -> int:
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        return self.lib.khsUpdateColliderVertices(self.handle, _ptr(v), v.shape[0])

    def update_roots(self, points) -> int:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        return self.lib.khsUpdateRootTargets(self.handle, _ptr(p), p.shape[0])

    def step(self, frame_dt: float) -> int:
        return self.lib.khsStep(self.handle, frame_dt)

    # ---------------------------------------------------------- animation
    def allocate_animation(self, frames: int) -> int:
        return self.lib.khsAllocateAnimation(self.handle, frames)

    def set_root_frame(self, frame: int, points) -> int:
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        return self.lib.khsSetRootAnimationFrame(self.handle, frame, _ptr(p), p.shape[0])

    def set_collider_frame(self, frame: int, vertices) -> int:
        if vertices is None:
            return self.lib.khsSetColliderAnimationFrame(self.handle, frame, None, 0)
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        return self.lib.khsSetColliderAnimationFrame(self.handle, frame, _ptr(v), v.shape[0])

    def finalize_animation(self) -> int:
        return self.lib.khsFinalizeAnimation(self.handle)

    def step_frame(self, frame: int, frame_dt: float) -> int:
        return self.lib.khsStepAnimationFrame(self.handle, frame, frame_dt)

    def checkpoint_size(self) -> int:
        return int(self.lib.khsGetAnimationCheckpointSize(self.handle))

    def save_checkpoint(self) -> np.ndarray:
        buf = np.zeros(self.checkpoint_size(), dtype=np.uint8)
        self.check(self.lib.khsSaveAnimationCheckpoint(self.handle, _ptr(buf), buf.size))
        return buf

    def restore_checkpoint(self, buf: np.ndarray) -> int:
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        return self.lib.khsRestoreAnimationCheckpoint(self.handle, _ptr(b), b.size)

    def set_rollback_ring(self, frames: int) -> int:
        return self.lib.khsxSetRollbackRing(self.handle, frames)

    def rollback(self, frame: int) -> int:
        return self.lib.khsxRollback(self.handle, frame)

    def write_cache_frame(self, fd: int) -> int:
        return self.lib.khsxWriteCacheFrame(self.handle, fd)

    # ------------------------------------------------------------- output
    def original_count(self) -> int:
        return int(self.lib.khsGetOriginalPointCount(self.handle))

    def node_count(self) -> int:
        return int(self.lib.khsGetInternalNodeCount(self.handle))

    def original_positions(self) -> np.ndarray:
        out = np.zeros((self.original_count(), 3))
        self.check(self.lib.khsCopyOriginalPositions(self.handle, _ptr(out), out.shape[0]))
        return out

    def internal_positions(self) -> np.ndarray:
        out = np.zeros((self.node_count(), 3))
        self.check(self.lib.khsCopyInternalPositions(self.handle, _ptr(out), out.shape[0]))
        return out

    def mapping(self) -> np.ndarray:
        out = np.zeros(self.original_count(), dtype=np.uint32)
        self.check(self.lib.khsCopyOriginalToInternalMap(self.handle, _ptr(out), out.size))
        return out

    def _get(self, fn, cls):
        s = cls()
        s.struct_size = C.sizeof(cls)
        self.check(fn(self.handle, C.byref(s)))
        return s

    def build_stats(self) -> KhsBuildStats:
        return self._get(self.lib.khsGetBuildStats, KhsBuildStats)

    def step_stats(self) -> KhsStepStats:
        return self._get(self.lib.khsGetLastStepStats, KhsStepStats)

    def gpu_stats(self) -> KhsGpuStats:
        return self._get(self.lib.khsGetGpuStats, KhsGpuStats)

    def progress(self) -> KhsProgress:
        return self._get(self.lib.khsGetProgress, KhsProgress)

    def failure(self) -> KhsFailureDiagnostics:
        return self._get(self.lib.khsGetFailureDiagnostics, KhsFailureDiagnostics)

    def request_cancel(self) -> int:
        return self.lib.khsRequestCancel(self.handle)

    # ---------------------------------------------------------- test hooks
    def dof_count(self) -> int:
        return int(self.lib.khsxGetDofCount(self.handle))

    def get_state(self):
        n = self.dof_count()
        q = np.zeros(n)
        v = np.zeros(n)
        self.check(self.lib.khsxGetState(self.handle, _ptr(q), _ptr(v), n))
        return q, v

    def set_state(self, q, v, q_old=None) -> int:
        q = np.ascontiguousarray(q, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        qo = None if q_old is None else np.ascontiguousarray(q_old, dtype=np.float64)
        return self.lib.khsxSetState(self.handle, _ptr(q), None if qo is None else _ptr(qo), _ptr(v), q.size)

    def node_model(self):
        n = self.node_count()
        frames = np.zeros((n, 9))
        mass = np.zeros(n)
        rot = np.zeros(n)
        flags = np.zeros(n, dtype=np.uint32)
        self.check(self.lib.khsxCopyNodeModel(self.handle, _ptr(frames), _ptr(mass), _ptr(rot), _ptr(flags), n))
        return frames, mass, rot, flags

    def element_model(self):
        e = self.build_stats().element_count
        nodes2 = np.zeros((e, 2), dtype=np.uint32)
        strand = np.zeros(e, dtype=np.uint32)
        contact = np.zeros(e, dtype=np.uint32)
        rest = np.zeros((e, 7))
        self.check(self.lib.khsxCopyElementModel(self.handle, _ptr(nodes2), _ptr(strand), _ptr(contact), _ptr(rest), e))
        return nodes2, strand, contact, rest

    def evaluate(self, dt: float, direction: bool = True):
        n = self.dof_count()
        g = np.zeros(n)
        d = np.zeros(n) if direction else None
        ev = KhsxEvaluation()
        ev.struct_size = C.sizeof(KhsxEvaluation)
        self.check(self.lib.khsxEvaluate(self.handle, dt, _ptr(g), None if d is None else _ptr(d), C.byref(ev)))
        return ev, g, d

    def broadphase_pairs(self) -> np.ndarray:
        count = C.c_uint64(0)
        self.check(self.lib.khsxBroadphasePairs(self.handle, None, 0, C.byref(count)))
        out = np.zeros((count.value, 2), dtype=np.uint32)
        if count.value:
            self.check(self.lib.khsxBroadphasePairs(self.handle, _ptr(out), count.value, C.byref(count)))
        return out

    def broadphase_pairs_no_alloc(self, capacity=None):
        """The allocation-free broad phase; capacity defaults to the count."""
        count = C.c_uint64(0)
        self.check(self.lib.khsxBroadphasePairsNoAlloc(self.handle, None, 0, C.byref(count)))
        cap = count.value if capacity is None else capacity
        out = np.zeros((cap, 2), dtype=np.uint32)
        if cap:
            self.check(self.lib.khsxBroadphasePairsNoAlloc(self.handle, _ptr(out), cap, C.byref(count)))
        return out[:count.value]

    def blocks(self):
        n = self.node_count()
        D = np.zeros((n, 6, 6))
        U = np.zeros((n, 6, 6))
        self.check(self.lib.khsxCopyBlocks(self.handle, _ptr(D), _ptr(U), n))
        return D, U

    def check_broadphase(self):
        v = (C.c_uint64 * 6)()
        self.check(self.lib.khsxCheckBroadphase(self.handle, v))
        return list(v)

    def set_convergence_mode(self, mode: int) -> int:
        return self.lib.khsxSetConvergenceMode(self.handle, mode)

    def set_precision_mode(self, mode: int) -> int:
        return self.lib.khsxSetPrecisionMode(self.handle, mode)

    def set_threads(self, threads: int) -> int:
        return self.lib.khsxSetThreadCount(self.handle, threads)

    def set_kernel_timing(self, enable: bool) -> int:
        return self.lib.khsxSetKernelTiming(self.handle, 1 if enable else 0)

    def kernel_timing(self) -> KhsxKernelTiming:
        return self._get(self.lib.khsxGetKernelTiming, KhsxKernelTiming)

    def contact_list_stats(self) -> KhsxContactListStats:
        return self._get(self.lib.khsxGetContactListStats, KhsxContactListStats)

    def set_strand_materials(self, young_modulus=None, radius=None) -> int:
        if young_modulus is None and radius is None:
            return self.lib.khsxSetStrandMaterials(self.handle, None, None, 0)
        E = np.ascontiguousarray(young_modulus, dtype=np.float64)
        r = np.ascontiguousarray(radius, dtype=np.float64)
        return self.lib.khsxSetStrandMaterials(self.handle, _ptr(E), _ptr(r), E.size)

    def set_collider_state(self, current, previous=None) -> int:
        cur = np.ascontiguousarray(current, dtype=np.float64)
        prev = None if previous is None else np.ascontiguousarray(previous, dtype=np.float64)
        return self.lib.khsxSetColliderState(self.handle, _ptr(cur), None if prev is None else _ptr(prev),
                                             cur.size // 3)

    def apply_friction(self, dt: float) -> int:
        return self.lib.khsxApplyFriction(self.handle, dt)

    def step_bound(self, direction) -> float:
        d = np.ascontiguousarray(direction, dtype=np.float64)
        alpha = C.c_double(0.0)
        self.check(self.lib.khsxStepBound(self.handle, _ptr(d), d.size, C.byref(alpha)))
        return alpha.value

    def sweep_check(self):
        feasible = C.c_uint32(0)
        diag = KhsFailureDiagnostics()
        diag.struct_size = C.sizeof(KhsFailureDiagnostics)
        self.check(self.lib.khsxSweepCheck(self.handle, C.byref(feasible), C.byref(diag)))
        return bool(feasible.value), diag

    def set_solve_algorithm(self, algorithm: int) -> int:
        return self.lib.khsxSetSolveAlgorithm(self.handle, algorithm)

    def set_contact_reuse(self, mode: int) -> int:
        return self.lib.khsxSetContactReuse(self.handle, mode)

    def set_direction_mode(self, mode: int) -> int:
        return self.lib.khsxSetDirectionMode(self.handle, mode)

    def launch_counts(self):
        """(dispatches, host waits) enqueued on the solver's stream so far."""
        dispatches, waits = C.c_uint64(0), C.c_uint64(0)
        self.check(self.lib.khsxGetLaunchCounts(self.handle, C.byref(dispatches), C.byref(waits)))
        return dispatches.value, waits.value

    def direction_counts(self):
        block, steepest = C.c_uint64(0), C.c_uint64(0)
        self.check(self.lib.khsxGetDirectionCounts(self.handle, C.byref(block), C.byref(steepest)))
        return block.value, steepest.value

    def substep_times(self) -> np.ndarray:
        count = C.c_uint32(0)
        self.check(self.lib.khsxGetSubstepTimes(self.handle, None, 0, C.byref(count)))
        out = np.zeros(count.value)
        if count.value:
            self.check(self.lib.khsxGetSubstepTimes(self.handle, _ptr(out), count.value, C.byref(count)))
        return out


def gpu_info(lib) -> KhsGpuInfo:
    info = KhsGpuInfo()
    info.struct_size = C.sizeof(KhsGpuInfo)
    lib.khsGetGpuInfo(C.byref(info))
    return info
