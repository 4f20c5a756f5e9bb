"""The block-tridiagonal solve's other algorithms on the CPU oracle
(khsxSetSolveAlgorithm): 1, the twisted factorization; 2, the fused
elimination (factors straight out of Gauss-Jordan on [S | U | rhs], no
inverse, no row exchanges). Each is pinned by what the mathematics fixes:
its direction solves the assembled Newton system H d = -g with a backward
error as small as the reference's block Thomas (a transposed coupling
block, a dropped Schur term, a wrong middle node or a wrong elimination
leaves an O(1) residual), on the sphere groom and on short strands of every
free count from 0 to 11 and 39; and a frame stepped with it follows the
block Thomas frame to well under the position bar."""
import numpy as np
import pytest

import grooms
import khs_native as kn
from linear_ref import build_short, newton_residual, perturbed_state

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)
ALGORITHMS = [1, 2]


def _directions(solver, dt, algorithm):
    out = []
    for a in (0, algorithm):
        solver.check(solver.set_solve_algorithm(a))
        e, g, d = solver.evaluate(dt)
        assert e.solve_failed == 0
        out.append((g, d))
    solver.check(solver.set_solve_algorithm(0))
    return out


@pytest.mark.parametrize("algorithm", ALGORITHMS)
@pytest.mark.parametrize("seed", [0, 1])
def test_direction_solves_sphere_system(oracle_lib, seed, algorithm):
    scene = grooms.small_sphere_scene()
    o = grooms.build_solver(oracle_lib, scene)
    q_old, q, v = perturbed_state(o, seed)
    o.check(o.set_state(q, v, q_old))
    (g, d_ref), (g2, d) = _directions(o, scene.frame_dt / 4, algorithm)
    D, U = o.blocks()
    assert np.array_equal(g, g2)
    res_ref = newton_residual(D, U, g, d_ref, o)
    res = newton_residual(D, U, g, d, o)
    rel = np.linalg.norm(d - d_ref) / np.linalg.norm(d_ref)
    print(f"algorithm {algorithm}: backward error thomas {res_ref:.3e}, this {res:.3e}; forward difference {rel:.3e}")
    assert res <= max(10 * res_ref, 1e-10)
    # the same direction up to the system's conditioning (kappa ~ 3e12: the
    # soft twist modes take forward differences up to about 1e-2)
    assert rel < 0.05
    assert not np.array_equal(d, d_ref)  # rounded differently: the other algorithm ran


@pytest.mark.parametrize("algorithm", ALGORITHMS)
@pytest.mark.parametrize("fixed", [1, 2])
def test_direction_short_strands(oracle_lib, fixed, algorithm):
    o = build_short(oracle_lib, fixed)
    q_old, q, v = perturbed_state(o, 5, dx=2e-4, du=0.1)
    o.check(o.set_state(q, v, q_old))
    (g, d_ref), (_, d) = _directions(o, 1.0 / 240, algorithm)
    D, U = o.blocks()
    res_ref = newton_residual(D, U, g, d_ref, o)
    res = newton_residual(D, U, g, d, o)
    print(f"algorithm {algorithm}, {fixed} fixed: backward error thomas {res_ref:.3e}, this {res:.3e}")
    assert res <= max(10 * res_ref, 1e-12)
    # fixed nodes get a zero direction
    _, _, _, flags = o.node_model()
    assert np.all(d.reshape(-1, 6)[(flags & 1) == 1] == 0.0)


@pytest.mark.parametrize("algorithm", ALGORITHMS)
def test_frame_follows_thomas(oracle_lib, algorithm):
    scene = grooms.small_sphere_scene()
    out = []
    for a in (0, algorithm):
        s = grooms.build_solver(oracle_lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_solve_algorithm(a))
        s.check(s.step(scene.frame_dt))
        st = s.step_stats()
        assert st.phase == kn.PHASE_FINISHED
        out.append((s.internal_positions(), st))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    print(f"algorithm {algorithm} vs thomas frame: {err:.3e} m, newton {out[0][1].newton_iterations} / "
          f"{out[1][1].newton_iterations}")
    assert err <= 1e-7


def test_fused_solve_fails_on_indefinite_block(oracle_lib):
    """Without row exchanges a non-positive pivot fails the solve (the host
    then takes steepest descent), as a singular pivot fails the reference's
    solve. A large negative mass damping makes the node blocks indefinite:
    block Thomas with partial pivoting still solves, the fused elimination
    reports the failure."""
    import khs_native as kn
    from linear_ref import short_strands

    pts, offs = short_strands()
    out = {}
    for damping in (0.0, -1.0e5):
        mat = kn.default_material(oracle_lib)
        mat.mass_damping = damping
        desc = kn.default_desc(oracle_lib)
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.02
        s = kn.HairSolver(oracle_lib, desc)
        s.check(s.set_hair(pts, offs, mat))
        s.check(s.build())
        q_old, q, v = perturbed_state(s, 5, dx=2e-4, du=0.1)
        s.check(s.set_state(q, v, q_old))
        for a in (0, 2):
            s.check(s.set_solve_algorithm(a))
            e, _, _ = s.evaluate(1.0 / 240)
            out[(damping, a)] = e.solve_failed
        D, _ = s.blocks()
        _, _, _, flags = s.node_model()
        out[(damping, "min eigenvalue")] = min(np.linalg.eigvalsh(D[i]).min() for i in np.flatnonzero((flags & 1) == 0))
    print(out)
    assert out[(0.0, "min eigenvalue")] > 0 and out[(-1.0e5, "min eigenvalue")] < 0
    assert out[(0.0, 0)] == out[(0.0, 2)] == 0
    assert out[(-1.0e5, 0)] == 0 and out[(-1.0e5, 2)] == 1


def test_solve_algorithm_argument(oracle_lib):
    s = build_short(oracle_lib, 1)
    assert s.set_solve_algorithm(3) == kn.KHS_ERROR_INVALID_ARGUMENT
    for a in (0, 1, 2):
        assert s.set_solve_algorithm(a) == kn.KHS_OK
