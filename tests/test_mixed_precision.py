"""The mixed-precision step (khsxSetPrecisionMode(1): FP32 blocks and block
solve, FP64 residual, energies, barrier, step bound and convergence test)
measured on the oracle, which restates the same FP32 block Thomas.

On the small-sphere scene (64 strands of 32 points, E = 3 GPa, r = 40 um)
the per-strand Newton systems have condition numbers near 3e12 (stretch
EA/L ~ 1e3 against twist GJ/L ~ 1e-7). Measured here:
  - the FP32 block Thomas leaves relative linear residuals of order 1e-3
    (FP64: 1e-10), with or without symmetric Jacobi scaling of the DOFs;
  - FP64 iterative refinement that reuses the FP32 factors diverges: each
    sweep multiplies the residual (the FP32 factors' error exceeds the
    system's own scale, kappa * 6e-8 >> 1);
  - as an inexact Newton direction refined by the FP64 Newton loop itself,
    it stalls: one frame to the 1e-6 residual takes the FP64 step 58 Newton
    iterations and the FP32 step hits the cap (24 per substep) with the
    residual still near 1, several times the line-search evaluations and a
    large share of steepest-descent fallbacks.
The CPU oracle is the checker here; nothing of the device library runs.
"""
import numpy as np

import grooms


def _strand_systems(solver, dt):
    q, _ = solver.get_state()
    rng = np.random.default_rng(0)
    n = q.size // 6
    dq = np.zeros((n, 6))
    dq[:, :3] = rng.uniform(-1e-4, 1e-4, (n, 3))
    dq[:, 3:] = rng.uniform(-0.05, 0.05, (n, 3))
    solver.check(solver.set_state(q + dq.ravel(), rng.uniform(-1e-3, 1e-3, 6 * n), q))
    _, g, _ = solver.evaluate(dt, direction=False)
    D, U = solver.blocks()
    _, _, _, flags = solver.node_model()
    nodes2, strand, _, _ = solver.element_model()
    g = g.reshape(n, 6)
    for s in np.unique(strand):
        nodes = np.unique(nodes2[strand == s].ravel())
        free = [int(x) for x in nodes if not flags[x] & 1]
        yield [D[x] for x in free], [U[x] for x in free], [-g[x] for x in free]


def _block_thomas(Dl, Ul, b, dtype):
    m = len(Dl)
    C, dd = [None] * m, [None] * m
    for i in range(m):
        S = Dl[i].astype(dtype)
        r = b[i].astype(dtype)
        if i:
            S = S - Ul[i - 1].astype(dtype).T @ C[i - 1]
            r = r - Ul[i - 1].astype(dtype).T @ dd[i - 1]
        shift = max(1e-12 * float(np.max(np.abs(np.diag(S)))), 1e-24)
        inv = np.linalg.inv(S + shift * np.eye(6, dtype=dtype)).astype(dtype)
        C[i] = inv @ Ul[i].astype(dtype)
        dd[i] = inv @ r
    x = [None] * m
    for i in range(m - 1, -1, -1):
        x[i] = dd[i] - (C[i] @ x[i + 1] if i + 1 < m else 0)
    return np.concatenate(x).astype(np.float64)


def _residual(Dl, Ul, b, x):
    m = len(Dl)
    Ax = np.zeros(6 * m)
    for i in range(m):
        Ax[6 * i:6 * i + 6] += Dl[i] @ x[6 * i:6 * i + 6]
        if i + 1 < m:
            Ax[6 * i:6 * i + 6] += Ul[i] @ x[6 * i + 6:6 * i + 12]
            Ax[6 * i + 6:6 * i + 12] += Ul[i].T @ x[6 * i:6 * i + 6]
    bb = np.concatenate(b)
    return float(np.linalg.norm(bb - Ax) / np.linalg.norm(bb))


def _dense(Dl, Ul):
    m = len(Dl)
    A = np.zeros((6 * m, 6 * m))
    for i in range(m):
        A[6 * i:6 * i + 6, 6 * i:6 * i + 6] = Dl[i]
        if i + 1 < m:
            A[6 * i:6 * i + 6, 6 * i + 6:6 * i + 12] = Ul[i]
            A[6 * i + 6:6 * i + 12, 6 * i:6 * i + 6] = Ul[i].T
    return A


def test_fp32_block_thomas_is_too_inaccurate(oracle_lib):
    scene = grooms.small_sphere_scene()
    solver = grooms.build_solver(oracle_lib, scene)
    r64, r32, conds = [], [], []
    for k, (Dl, Ul, b) in enumerate(_strand_systems(solver, scene.frame_dt / 4)):
        r64.append(_residual(Dl, Ul, b, _block_thomas(Dl, Ul, b, np.float64)))
        r32.append(_residual(Dl, Ul, b, _block_thomas(Dl, Ul, b, np.float32)))
        if k < 8:
            conds.append(np.linalg.cond(_dense(Dl, Ul)))
    print(f"condition median {np.median(conds):.2e}; residual FP64 median {np.median(r64):.2e}, "
          f"FP32 median {np.median(r32):.2e} max {np.max(r32):.2e}")
    assert len(r64) == 64
    assert np.median(conds) > 1e11
    assert np.max(r64) < 1e-8
    # six orders of magnitude lost: an FP32 direction is not a Newton direction
    assert np.median(r32) > 1e-5


def _block_factors(Dl, Ul, dtype):
    m = len(Dl)
    inv, C = [None] * m, [None] * m
    for i in range(m):
        S = Dl[i].astype(dtype)
        if i:
            S = S - Ul[i - 1].astype(dtype).T @ C[i - 1]
        shift = max(1e-12 * float(np.max(np.abs(np.diag(S)))), 1e-24)
        inv[i] = np.linalg.inv((S + shift * np.eye(6, dtype=dtype)).astype(dtype)).astype(dtype)
        C[i] = (inv[i] @ Ul[i].astype(dtype)).astype(dtype)
    return inv, C


def _resolve(inv, Ul, C, b, dtype):
    m = len(inv)
    y = [None] * m
    for i in range(m):
        r = b[i].astype(dtype)
        if i:
            r = r - Ul[i - 1].astype(dtype).T @ y[i - 1]
        y[i] = inv[i] @ r
    x = [None] * m
    for i in range(m - 1, -1, -1):
        x[i] = y[i] - (C[i] @ x[i + 1] if i + 1 < m else 0)
    return np.concatenate(x).astype(np.float64)


def test_fp64_refinement_of_fp32_factors_diverges(oracle_lib):
    """Symmetric Jacobi scaling, FP32 factors, FP64 residuals: refinement
    sweeps that reuse the factors make the residual grow."""
    scene = grooms.small_sphere_scene()
    solver = grooms.build_solver(oracle_lib, scene)
    history = []
    for Dl, Ul, b in _strand_systems(solver, scene.frame_dt / 4):
        m = len(Dl)
        sc = [1.0 / np.sqrt(np.diag(D)) for D in Dl]
        Ds = [D * np.outer(c, c) for D, c in zip(Dl, sc)]
        Us = [U * np.outer(sc[i], sc[i + 1]) if i + 1 < m else U * 0.0 for i, U in enumerate(Ul)]
        inv, C = _block_factors(Ds, Us, np.float32)
        x = _resolve(inv, Us, C, [bb * c for bb, c in zip(b, sc)], np.float32)
        res = []
        for sweep in range(3):
            d = x.reshape(m, 6) * np.array(sc)
            res.append(_residual(Dl, Ul, b, d.ravel()))
            r = [(b[i] - (Dl[i] @ d[i] + (Ul[i] @ d[i + 1] if i + 1 < m else 0) + (Ul[i - 1].T @ d[i - 1] if i else 0)))
                 * sc[i] for i in range(m)]
            x = x + _resolve(inv, Us, C, r, np.float32)
        history.append(res)
    h = np.array(history)
    print("median relative residual after 0, 1, 2 refinement sweeps:", np.median(h, axis=0))
    assert np.median(h[:, 0]) > 1e-4          # scaling does not rescue FP32
    assert np.median(h[:, 2]) > np.median(h[:, 0])  # refinement makes it worse


def test_fp32_newton_directions_stall_the_step(oracle_lib):
    lifted = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)
    scene = grooms.small_sphere_scene()
    runs = {}
    for mode in (0, 1):
        s = grooms.build_solver(oracle_lib, scene, **lifted)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_precision_mode(mode))
        s.set_threads(4)
        s.check(s.step(scene.frame_dt))
        st = s.step_stats()
        runs[mode] = (st.newton_iterations, st.line_search_evaluations, st.relative_residual_norm,
                      s.direction_counts(), s.internal_positions())
        s.close()
    print(f"FP64: newton {runs[0][0]}, line search {runs[0][1]}, residual {runs[0][2]:.2e}, directions {runs[0][3]}; "
          f"FP32 solve: newton {runs[1][0]}, line search {runs[1][1]}, residual {runs[1][2]:.2e}, "
          f"directions {runs[1][3]}; max position difference {np.abs(runs[0][4] - runs[1][4]).max():.2e} m")
    cap = 4 * 24
    assert runs[0][0] < cap and runs[0][2] <= 1e-6
    assert runs[1][0] == cap and runs[1][2] > 1e-3
    assert runs[1][1] > 3 * runs[0][1]
