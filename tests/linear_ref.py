"""Residual checks of the per-strand Newton systems, shared by the solve
tests: a direction d of the block-tridiagonal system H d = -g (diagonal
blocks D, element coupling blocks U between an element's two nodes) is
judged by its backward error in that system, which any solve of it must
make small whatever its rounding."""
import numpy as np


def newton_residual(D, U, g, d, solver):
    """||H d + g|| / ||g|| over the free nodes of `solver`'s model."""
    _, _, _, flags = solver.node_model()
    n = D.shape[0]
    dd = d.reshape(n, 6)
    r = np.einsum("nab,nb->na", D, dd)
    nodes2, _, _, _ = solver.element_model()
    i, j = nodes2[:, 0], nodes2[:, 1]
    r[i] += np.einsum("nab,nb->na", U[i], dd[j])
    r[j] += np.einsum("nba,nb->na", U[i], dd[i])
    r += g.reshape(n, 6)
    free = (flags & 1) == 0
    return float(np.linalg.norm(r[free]) / np.linalg.norm(g.reshape(n, 6)[free]))


def perturbed_state(solver, seed, dx=1e-4, du=0.05, dv=1e-3):
    """(q, q + dq, v): the solver's state moved by seeded uniform noise in
    positions (dx), rotation vectors (du) and velocities (dv)."""
    q, _ = solver.get_state()
    rng = np.random.default_rng(seed)
    n = q.size // 6
    dq = np.zeros((n, 6))
    dq[:, :3] = rng.uniform(-dx, dx, (n, 3))
    dq[:, 3:] = rng.uniform(-du, du, (n, 3))
    vel = rng.uniform(-dv, dv, (n, 6))
    return q, q + dq.ravel(), vel.ravel()


SHORT_COUNTS = [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 40, 5, 2, 7, 3]


def short_strands():
    """Points and offsets of 16 curled strands of 2..40 nodes, 5 mm apart
    along each strand: with one fixed root node their free counts are
    1..39 (even and odd), with two the 2-node strands have none."""
    pts, offs = [], [0]
    for s, c in enumerate(SHORT_COUNTS):
        t = np.arange(c) * 0.005
        pts.append(np.stack([0.01 * s + 0.002 * np.sin(7 * t + s), 0.002 * np.cos(5 * t + s), -t], 1))
        offs.append(offs[-1] + c)
    return np.concatenate(pts), offs


def build_short(lib, fixed_root_nodes, **desc_overrides):
    import khs_native as kn

    pts, offs = short_strands()
    desc = kn.default_desc(lib)
    desc.fixed_root_nodes = fixed_root_nodes
    desc.maximum_element_length = 0.02
    for k, v in desc_overrides.items():
        setattr(desc, k, v)
    s = kn.HairSolver(lib, desc)
    s.check(s.set_hair(pts, offs, kn.default_material(lib)))
    s.check(s.build())
    return s
