"""The reference's scenario tests, restated over the ctypes binding.

Each function takes a loaded library (the CPU oracle or the HIP library) and
asserts what the matching reference test asserts
(tests/test_solver.cpp:18-343 and tests/test_c_api.c:11-30 of the
reference); EXTENSIONS holds scenarios of the extension API, which has no
reference counterpart. The test modules run them against the oracle on the
CPU and against the HIP library on the GPU.
"""
import ctypes as C
import math

import numpy as np

import khs_native as kn


def _solver(lib, **desc_fields):
    desc = kn.default_desc(lib)
    for k, v in desc_fields.items():
        if k == "gravity":
            desc.gravity = kn.KhsVec3(*v)
        else:
            setattr(desc, k, v)
    return kn.HairSolver(lib, desc), desc


def c_api_defaults(lib):
    """tests/test_c_api.c:11-24 (the GPU-info part is separate)."""
    desc = kn.default_desc(lib)
    mat = kn.default_material(lib)
    assert lib.khsGetAbiVersion() == kn.ABI_VERSION
    assert desc.struct_size == C.sizeof(kn.KhsSolverDesc)
    assert mat.struct_size == C.sizeof(kn.KhsHairMaterial)
    assert desc.minimum_dynamic_length == 0.0
    assert desc.maximum_substeps == 32
    desc.maximum_substeps = desc.substeps - 1
    assert not lib.khsCreate(C.byref(desc))
    desc.maximum_substeps = 32
    desc.minimum_dynamic_length = -1.0
    assert not lib.khsCreate(C.byref(desc))
    desc.minimum_dynamic_length = 0.0
    h = lib.khsCreate(C.byref(desc))
    assert h
    lib.khsDestroy(h)


def rest_state_and_subdivision(lib):
    """tests/test_solver.cpp:18-60."""
    s, desc = _solver(lib, gravity=(0, 0, 0), substeps=1, fixed_root_nodes=1, maximum_element_length=0.006)
    mat = kn.default_material(lib)
    mat.mass_damping = 0.0
    pts = np.array([[0, 0, 0], [0, 0, -0.01], [0, 0, -0.02]], float)
    s.check(s.set_hair(pts, [0, 3], mat))
    s.check(s.build())
    assert s.node_count() == 5
    desc.substeps = 2
    desc.newton_iterations = 32
    mat.density = 1400.0
    mat.radius = 5.0e-5
    mat.young_modulus = 2.0e9
    s.check(s.update_runtime(desc, mat))
    changed = kn.KhsSolverDesc.from_buffer_copy(desc)
    changed.maximum_element_length = 0.007
    assert s.update_runtime(changed, mat) == kn.KHS_ERROR_INVALID_ARGUMENT
    s.check(s.step(1.0 / 24.0))
    out = s.original_positions()
    assert np.all(np.abs(out - pts) < 1.0e-10)
    st = s.step_stats()
    assert st.phase == kn.PHASE_FINISHED
    assert st.final_residual_norm < 1.0e-5
    s.close()


def minimum_dynamic_length(lib):
    """tests/test_solver.cpp:62-107."""
    s, desc = _solver(lib, gravity=(0, 0, 0), substeps=1, fixed_root_nodes=1, maximum_element_length=0.06,
                      minimum_dynamic_length=0.30)
    mat = kn.default_material(lib)
    mat.mass_damping = 0.0
    pts = np.array([[0, 0, 0], [0.05, 0, 0], [0.10, 0, 0]], float)
    s.check(s.set_hair(pts, [0, 3], mat))
    s.check(s.build())
    b = s.build_stats()
    assert b.original_point_count == 3
    assert b.internal_node_count == 7
    assert b.element_count == 6
    assert b.virtual_extension_strand_count == 1
    assert b.virtual_extension_node_count == 4
    assert abs(b.virtual_extension_rest_length - 0.20) < 1.0e-12
    internal = s.internal_positions()
    assert abs(internal[-1, 0] - 0.30) < 1.0e-12
    assert abs(internal[-1, 1]) < 1.0e-12
    assert abs(internal[-1, 2]) < 1.0e-12
    s.check(s.step(1.0 / 24.0))
    out = s.original_positions()
    assert np.all(np.abs(out - pts) < 1.0e-10)
    s.close()


def virtual_extension_ignores_collider(lib):
    """tests/test_solver.cpp:109-170."""
    kw = dict(gravity=(0, 0, 0), substeps=1, fixed_root_nodes=1, maximum_element_length=0.05,
              minimum_dynamic_length=0.30)
    mat = kn.default_material(lib)
    mat.mass_damping = 0.0
    pts = np.array([[0, 0, 0], [0.05, 0, 0]], float)
    tri = np.array([[0, 1, 2]], np.uint32)
    s, _ = _solver(lib, **kw)
    collider = np.array([[0.10, -1.0, 0.0], [0.40, -1.0, 0.0], [0.25, 1.0, 0.0]])
    s.check(s.set_hair(pts, [0, 2], mat))
    s.check(s.set_collider(collider, tri))
    s.check(s.build())
    b = s.build_stats()
    assert b.virtual_extension_strand_count == 1
    assert math.isinf(b.initial_minimum_gap)
    s.check(s.step(1.0 / 24.0))
    st = s.step_stats()
    assert st.contact_candidate_count == 0
    assert st.active_contact_count == 0
    s.close()

    s, _ = _solver(lib, **kw)
    below = np.array([[0.10, -1.0, -0.1], [0.40, -1.0, -0.1], [0.25, 1.0, -0.1]])
    above = np.array([[0.10, -1.0, 0.1], [0.40, -1.0, 0.1], [0.25, 1.0, 0.1]])
    s.check(s.set_hair(pts, [0, 2], mat))
    s.check(s.set_collider(below, tri))
    s.check(s.build())
    s.check(s.update_collider(above))
    s.check(s.step(1.0 / 24.0))
    st = s.step_stats()
    assert st.contact_candidate_count == 0
    assert st.active_contact_count == 0
    s.close()


def duplicate_mapping(lib):
    """tests/test_solver.cpp:172-195."""
    s, _ = _solver(lib, gravity=(0, 0, 0), maximum_element_length=1.0, fixed_root_nodes=1)
    mat = kn.default_material(lib)
    pts = np.array([[0, 0, 0], [0, 0, 0], [0, 0, 0.1]], float)
    s.check(s.set_hair(pts, [0, 3], mat))
    s.check(s.build())
    m = s.mapping()
    assert m[0] == m[1]
    assert m[2] != m[1]
    assert s.build_stats().merged_zero_length_segment_count == 1
    s.close()


def initial_intersection_rejected(lib):
    """tests/test_solver.cpp:197-216."""
    s, _ = _solver(lib, maximum_element_length=1.0, fixed_root_nodes=1)
    mat = kn.default_material(lib)
    mat.radius = 1.0e-3
    mat.barrier_distance = 2.0e-3
    pts = np.array([[-0.2, 0, 0.2], [-0.1, 0, 0.1], [0.1, 0, -0.1]])
    collider = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 1.0, 0.0]])
    s.check(s.set_hair(pts, [0, 3], mat))
    s.check(s.set_collider(collider, [[0, 1, 2]]))
    assert s.build() == kn.KHS_ERROR_INITIAL_INTERSECTION
    s.close()


def barrier_contact_remains_feasible(lib):
    """tests/test_solver.cpp:218-251."""
    s, desc = _solver(lib, substeps=4, fixed_root_nodes=1, maximum_element_length=0.02, newton_iterations=40)
    mat = kn.default_material(lib)
    mat.radius = 2.0e-4
    mat.young_modulus = 1.0e6
    mat.contact_stiffness = 100.0
    mat.barrier_distance = 2.0e-3
    mat.mass_damping = 2.0
    pts = np.array([[0, 0, 0.03], [0, 0, 0.02], [0, 0, 0.01], [0, 0, 0.001]])
    collider = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]])
    s.check(s.set_hair(pts, [0, 4], mat))
    s.check(s.set_collider(collider, [[0, 1, 2], [0, 2, 3]]))
    s.check(s.build())
    for _ in range(3):
        s.check(s.step(1.0 / 60.0))
    st = s.step_stats()
    assert st.active_contact_count > 0
    assert st.minimum_gap > desc.minimum_gap
    assert st.phase == kn.PHASE_FINISHED
    s.close()


def moving_collider_tunnelling_rejected(lib):
    """tests/test_solver.cpp:253-288."""
    s, _ = _solver(lib, substeps=1, maximum_substeps=4, fixed_root_nodes=1, maximum_element_length=1.0)
    mat = kn.default_material(lib)
    mat.radius = 1.0e-3
    pts = np.array([[-0.2, 0, 0], [-0.1, 0, 0], [0.1, 0, 0]])
    below = np.array([[-1.0, -1.0, -0.1], [1.0, -1.0, -0.1], [0.0, 1.0, -0.1]])
    above = np.array([[-1.0, -1.0, 0.1], [1.0, -1.0, 0.1], [0.0, 1.0, 0.1]])
    s.check(s.set_hair(pts, [0, 3], mat))
    s.check(s.set_collider(below, [[0, 1, 2]]))
    s.check(s.build())
    s.check(s.update_collider(above))
    assert s.step(1.0 / 24.0) == kn.KHS_ERROR_NOT_CONVERGED
    d = s.failure()
    assert d.kind == kn.FAILURE_SWEEP
    assert d.requested_substeps == 1
    assert d.attempted_substeps == 4
    assert d.maximum_substeps == 4
    assert d.adaptive_attempt_count == 3
    assert d.strand_index == 0
    assert d.collider_triangle_index == 0
    assert d.distance <= d.required_distance + 1.0e-12
    assert abs(d.collider_frame_displacement - 0.2) < 1.0e-12
    s.close()


def resident_animation_and_checkpoint(lib):
    """tests/test_solver.cpp:290-343."""
    s, _ = _solver(lib, gravity=(0, 0, 0), substeps=1, newton_iterations=100, fixed_root_nodes=1,
                   maximum_element_length=1.0)
    mat = kn.default_material(lib)
    first = np.array([[0, 0, 0], [0, 0, -0.1], [0, 0, -0.2]], float)
    second = np.array([[0.001, 0, 0], [0.001, 0, -0.1], [0.001, 0, -0.2]])
    s.check(s.set_hair(first, [0, 3], mat))
    s.check(s.build())
    s.check(s.allocate_animation(2))
    s.check(s.set_root_frame(0, first))
    s.check(s.set_root_frame(1, second))
    s.check(s.set_collider_frame(0, None))
    s.check(s.set_collider_frame(1, None))
    s.check(s.finalize_animation())
    assert s.checkpoint_size() > 0
    cp = s.save_checkpoint()
    s.check(s.step_frame(1, 1.0 / 24.0))
    r1 = s.original_positions()
    s.check(s.restore_checkpoint(cp))
    s.check(s.step_frame(1, 1.0 / 24.0))
    r2 = s.original_positions()
    assert np.all(np.abs(r1 - r2) < 1.0e-14)
    g = s.gpu_stats()
    assert g.animation_frame_count == 2
    assert g.resident_bytes > 0
    assert g.last_frame_milliseconds > 0.0
    assert s.progress().phase == kn.PHASE_FINISHED
    s.close()


def rollback_ring(lib):
    """The rollback ring of the extension API (no counterpart in the
    reference): the state after each stepped frame is kept in slot
    frame % size; rolling back to a kept frame and stepping on reproduces the
    straight run bit for bit, and a frame the ring no longer holds is
    refused."""
    s, _ = _solver(lib, gravity=(0, -9.81, 0), fixed_root_nodes=1)
    mat = kn.default_material(lib)
    strand = np.array([[0, 0, 0], [0, -0.02, 0], [0, -0.04, 0], [0, -0.06, 0]], float)
    points = np.concatenate([strand, strand + [0.01, 0, 0.003]])
    s.check(s.set_hair(points, [0, 4, 8], mat))
    s.check(s.build())
    assert s.rollback(1) == kn.KHS_ERROR_INVALID_STATE
    frames, dt = 7, 1.0 / 24.0
    s.check(s.allocate_animation(frames))
    for k in range(frames):
        s.check(s.set_root_frame(k, points + [0.004 * math.sin(0.7 * k), 0.0, 0.003 * (1.0 - math.cos(0.7 * k))]))
        s.check(s.set_collider_frame(k, None))
    s.check(s.finalize_animation())
    s.check(s.set_rollback_ring(3))
    straight = {}
    for k in range(1, frames):
        s.check(s.step_frame(k, dt))
        straight[k] = s.original_positions()
    # frames 4, 5 and 6 are held; 3 has been evicted by 6
    assert s.rollback(3) == kn.KHS_ERROR_INVALID_ARGUMENT
    assert s.rollback(frames + 1) == kn.KHS_ERROR_INVALID_ARGUMENT
    s.check(s.rollback(4))
    assert np.array_equal(s.original_positions(), straight[4])
    assert s.progress().phase == kn.PHASE_FINISHED
    assert s.step_frame(6, dt) == kn.KHS_ERROR_INVALID_STATE
    for k in (5, 6):
        s.check(s.step_frame(k, dt))
        assert np.array_equal(s.original_positions(), straight[k]), k
    # rolling back further than the last rollback, then stepping on again
    s.check(s.rollback(5))
    s.check(s.step_frame(6, dt))
    assert np.array_equal(s.original_positions(), straight[6])
    assert not np.array_equal(straight[4], straight[6])
    # a zero-sized ring keeps nothing
    s.check(s.set_rollback_ring(0))
    assert s.rollback(6) == kn.KHS_ERROR_INVALID_ARGUMENT
    s.close()


def every_node_fixed(lib):
    """fixed_root_nodes at or above a strand's node count fixes the whole
    strand (its count is min(fixed_root_nodes, nodes), reference
    src/solver.cpp:227-318): that strand has no degrees of freedom and
    follows its root targets, while a longer strand beside it moves freely."""
    s, _ = _solver(lib, gravity=(0, 0, -9.81), fixed_root_nodes=6, maximum_element_length=0.05)
    short = np.array([[0, 0, 0], [0, 0, -0.01], [0, 0, -0.02], [0, 0, -0.03]], float)
    long = np.array([[0.01, 0, -0.01 * k] for k in range(12)], float)
    points = np.concatenate([short, long])
    s.check(s.set_hair(points, [0, 4, 16], kn.default_material(lib)))
    s.check(s.build())
    b = s.build_stats()
    assert b.internal_node_count == 16 and b.fixed_node_count == 4 + 6
    assert b.degree_of_freedom_count == 6 * 6
    s.check(s.allocate_animation(3))
    for k in range(3):
        s.check(s.set_root_frame(k, points + [0.002 * k, 0.0, 0.001 * k]))
        s.check(s.set_collider_frame(k, None))
    s.check(s.finalize_animation())
    start = s.original_positions()
    for k in (1, 2):
        s.check(s.step_frame(k, 1.0 / 24.0))
    out = s.original_positions()
    target = points + [0.004, 0.0, 0.002]
    assert np.max(np.abs(out[:4] - target[:4])) <= 1e-15
    assert np.max(np.abs(out[:6 + 4] - target[:6 + 4])) <= 1e-15
    assert np.max(np.abs(out[10:] - start[10:])) > 1e-4  # the free part of the long strand fell
    s.close()


def cancel_mid_frame(lib):
    """khsRequestCancel from another thread while a frame runs: the step
    returns KHS_ERROR_INVALID_STATE at its next Newton iteration, the state is
    back at the frame's start, progress reports the cancel and records no
    failure, and the next step runs normally (reference
    src/cuda_backend.cu:556, 657, 693, 711-714)."""
    import threading
    import time

    scene_points = []
    for k in range(16):
        x = 0.004 * (k % 4)
        y = 0.004 * (k // 4)
        scene_points.append(np.array([[x + 0.005 * j, y, 0.0] for j in range(24)], float))  # level, so it sags
    points = np.concatenate(scene_points)
    offsets = list(range(0, len(points) + 1, 24))
    s, _ = _solver(lib, gravity=(0, 0, -9.81), substeps=32, maximum_substeps=32, fixed_root_nodes=1,
                   newton_iterations=100, relative_tolerance=1e-9, absolute_tolerance=1e-30)
    s.check(s.set_hair(points, offsets, kn.default_material(lib)))
    s.check(s.build())
    s.check(s.set_convergence_mode(1))  # thin hair would meet the default floors at once
    before = s.internal_positions()
    result = {}
    worker = threading.Thread(target=lambda: result.setdefault("rc", s.step(1.0 / 24.0)))
    worker.start()
    deadline = time.time() + 60.0
    while time.time() < deadline and "rc" not in result:
        p = s.progress()
        if p.substep >= 2:
            s.check(s.request_cancel())
            break
    worker.join()
    assert result["rc"] == kn.KHS_ERROR_INVALID_STATE, result["rc"]
    assert "cancel" in s.last_error().lower()
    assert s.progress().cancelled == 1
    assert np.array_equal(s.internal_positions(), before)
    assert s.failure().kind == 0
    s.check(s.step(1.0 / 24.0))
    assert s.progress().cancelled == 0
    assert not np.array_equal(s.internal_positions(), before)
    s.close()


EXTENSIONS = [rollback_ring, every_node_fixed, cancel_mid_frame]

ALL = [
    c_api_defaults,
    rest_state_and_subdivision,
    minimum_dynamic_length,
    virtual_extension_ignores_collider,
    duplicate_mapping,
    initial_intersection_rejected,
    barrier_contact_remains_feasible,
    moving_collider_tunnelling_rejected,
    resident_animation_and_checkpoint,
]
