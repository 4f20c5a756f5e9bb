"""Parity of the device's other block-solve algorithms with the oracle's,
which restate them operation for operation: 1, the twisted factorization;
2, the fused elimination (factors straight out of Gauss-Jordan on
[S | U | rhs], no inverse, no row exchanges, one DPP fused multiply-add
per term).

Bars as for block Thomas: the direction by its backward error in the
oracle's own system (no worse than 10x the oracle's, or 1e-10), zero on
fixed nodes; positions after a frame stepped with the algorithm on both
libraries within 1e-7 m. Short strands of every free count from 0 to 11 and
39 put strands of different lengths into one wavefront (each row or side
stops at its own length) and cover the twisted middle node without a root
side (one or two free nodes) and without any free node."""
import numpy as np
import pytest

import grooms
import khs_native as kn
from linear_ref import build_short, newton_residual, perturbed_state, short_strands

pytestmark = pytest.mark.gpu

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)
ALGORITHMS = [1, 2]


def _evaluate(solvers, dt, algorithm):
    out = []
    for s in solvers:
        s.check(s.set_solve_algorithm(algorithm))
        e, g, d = s.evaluate(dt)
        assert e.solve_failed == 0
        out.append((g, d))
    return out


@pytest.mark.parametrize("algorithm", ALGORITHMS)
@pytest.mark.parametrize("seed", [0, 1])
def test_direction_matches_oracle(oracle_lib, gpu_lib, seed, algorithm):
    print(f"[gpu] solve algorithm {algorithm} direction, small sphere scene, seed {seed}", flush=True)
    scene = grooms.small_sphere_scene()
    o = grooms.build_solver(oracle_lib, scene)
    g = grooms.build_solver(gpu_lib, scene)
    q_old, q, v = perturbed_state(o, seed)
    for s in (o, g):
        s.check(s.set_state(q, v, q_old))
    (go, do), (gg, dg) = _evaluate((o, g), scene.frame_dt / 4, algorithm)
    D, U = o.blocks()
    res_o = newton_residual(D, U, go, do, o)
    res_g = newton_residual(D, U, go, dg, o)
    print(f"backward error oracle {res_o:.3e} gpu {res_g:.3e}; direction rel {np.linalg.norm(dg - do) / np.linalg.norm(do):.3e}")
    assert res_g <= max(10 * res_o, 1e-10)


@pytest.mark.parametrize("algorithm", ALGORITHMS)
@pytest.mark.parametrize("fixed", [1, 2])
def test_short_strands_match_oracle(oracle_lib, gpu_lib, fixed, algorithm):
    print(f"[gpu] solve algorithm {algorithm} direction, short strands, {fixed} fixed root nodes", flush=True)
    o, g = build_short(oracle_lib, fixed), build_short(gpu_lib, fixed)
    q_old, q, v = perturbed_state(o, 5, dx=2e-4, du=0.1)
    for s in (o, g):
        s.check(s.set_state(q, v, q_old))
    (go, do), (gg, dg) = _evaluate((o, g), 1.0 / 240, algorithm)
    D, U = o.blocks()
    res_o = newton_residual(D, U, go, do, o)
    res_g = newton_residual(D, U, go, dg, o)
    print(f"backward error oracle {res_o:.3e} gpu {res_g:.3e}; direction rel {np.linalg.norm(dg - do) / np.linalg.norm(do):.3e}")
    assert res_g <= max(10 * res_o, 1e-12)
    _, _, _, flags = o.node_model()
    assert np.all(dg.reshape(-1, 6)[(flags & 1) == 1] == 0.0)


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_frames_match_oracle(oracle_lib, gpu_lib, algorithm):
    print(f"[gpu] solve algorithm {algorithm}, one frame, small sphere scene", flush=True)
    scene = grooms.small_sphere_scene()
    out = []
    for lib in (oracle_lib, gpu_lib):
        s = grooms.build_solver(lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_solve_algorithm(algorithm))
        s.check(s.step(scene.frame_dt))
        out.append((s.internal_positions(), s.step_stats(), s.direction_counts()))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    print(f"max position error {err:.3e} m; newton oracle {out[0][1].newton_iterations} gpu {out[1][1].newton_iterations}; "
          f"directions oracle {out[0][2]} gpu {out[1][2]}")
    assert err <= 1e-7
    assert out[1][1].phase == kn.PHASE_FINISHED
    assert out[1][2][0] > 0


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_short_strand_frames_match_oracle(oracle_lib, gpu_lib, algorithm):
    print(f"[gpu] solve algorithm {algorithm}, three frames, short strands", flush=True)
    out = []
    for lib in (oracle_lib, gpu_lib):
        s = build_short(lib, 1, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_solve_algorithm(algorithm))
        if not out:
            state = perturbed_state(s, 5, dx=2e-4, du=0.1, dv=0.05)
        q_old, q, v = state
        s.check(s.set_state(q, v, q_old))
        for _ in range(3):
            s.check(s.step(1.0 / 60))
        out.append((s.internal_positions(), s.direction_counts()))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    print(f"max position error {err:.3e} m; directions oracle {out[0][1]} gpu {out[1][1]}")
    assert err <= 1e-7
    assert out[1][1][0] > 10


def test_fused_solve_failure_matches_oracle(oracle_lib, gpu_lib):
    """Indefinite node blocks (large negative mass damping): the fused
    elimination reports the failure on both libraries, block Thomas with
    pivoting solves on both."""
    print("[gpu] fused solve on indefinite blocks", flush=True)
    pts, offs = short_strands()
    out = {}
    for name, lib in (("oracle", oracle_lib), ("gpu", gpu_lib)):
        mat = kn.default_material(lib)
        mat.mass_damping = -1.0e5
        desc = kn.default_desc(lib)
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.02
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, offs, mat))
        s.check(s.build())
        if not out:
            state = perturbed_state(s, 5, dx=2e-4, du=0.1)
        q_old, q, v = state
        s.check(s.set_state(q, v, q_old))
        for a in (0, 2):
            s.check(s.set_solve_algorithm(a))
            e, _, _ = s.evaluate(1.0 / 240)
            out[(name, a)] = e.solve_failed
    print(out)
    assert out[("oracle", 0)] == out[("gpu", 0)] == 0
    assert out[("oracle", 2)] == out[("gpu", 2)] == 1


def test_fused_frames_match_reference_algorithm(oracle_lib, gpu_lib):
    """The device's fused elimination against the oracle running the
    reference's own block Thomas (Gauss-Jordan inverse, partial pivoting):
    one frame of the sphere scene lands within the position bar with the
    same Newton count."""
    print("[gpu] fused solve on the device vs block Thomas on the oracle, one frame", flush=True)
    scene = grooms.small_sphere_scene()
    out = []
    for lib, algorithm in ((oracle_lib, 0), (gpu_lib, 2)):
        s = grooms.build_solver(lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_solve_algorithm(algorithm))
        s.check(s.step(scene.frame_dt))
        out.append((s.internal_positions(), s.step_stats()))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    print(f"max position error {err:.3e} m; newton oracle {out[0][1].newton_iterations} gpu {out[1][1].newton_iterations}")
    assert err <= 1e-7
    assert out[0][1].newton_iterations == out[1][1].newton_iterations
