"""Why the step has no mixed-precision form: the per-strand Newton systems
defeat an FP32 block solve.

On the small-sphere scene (64 strands of 32 points, E = 3 GPa, r = 40 um)
the oracle's Gauss-Newton blocks are taken at a perturbed state and each
strand's free-node system H d = -g is solved by block Thomas (block LU of the
tridiagonal matrix, the reference's algorithm) in FP64 and in FP32. The
systems have condition numbers near 3e12 (stretch EA/L ~ 1e3 against twist
GJ/L ~ 1e-7), so the FP32 solve leaves relative residuals of order 1e-3
where FP64 leaves 1e-10: an FP32 direction needs several FP64 refinement
passes (each another sequential sweep of the strand) to be usable, while on
MI355X non-packed FP32 and FP64 vector instructions issue at the same rate.
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
