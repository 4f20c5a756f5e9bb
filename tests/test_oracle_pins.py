"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no golden vectors, and its CUDA/Eigen build cannot run
here, so each part of the oracle's step is checked against closed forms,
invariants or brute force, chosen so that a plausible slip (a dropped
factor, a wrong strain index, a transposed block, a sign) fails one of them:

  elastic energy   uniform stretch, pure twist and pure bend closed forms
                   (reference cuda_backend.cu:905-919 weights and strains)
  gradient         central differences of the objective
  Gauss-Newton     at the rest state the Gauss-Newton blocks equal the
                   exact Hessian: central differences of the gradient
  block solve      dense solve of the assembled per-strand matrix
  barrier          b(g) = -k (d - g)^2 ln(g / d) for a segment parallel to
                   a plane (cuda_backend.cu:255-265)
  closest point    segment above a triangle edge: distance in closed form
  broad phase      brute-force box test over every (element, triangle)
  root targets     a rigid rotation of the roots about z by phi gives the
                   rotation vector (0, 0, phi) on the fixed node
                   (solver.cpp:19-26, 416-447)
"""
import math

import numpy as np
import pytest

import grooms
import khs_native as kn


def _strand_solver(lib, points, *, collider=None, E=3.0e6, nu=0.3, radius=1.0e-4, damping=0.0, fixed=1,
                   barrier=1.0e-3, stiffness=10.0, h_max=0.05):
    desc = kn.default_desc(lib)
    desc.gravity = kn.KhsVec3(0.0, 0.0, 0.0)
    desc.fixed_root_nodes = fixed
    desc.maximum_element_length = h_max
    desc.minimum_dynamic_length = 0.0
    mat = kn.default_material(lib)
    mat.young_modulus = E
    mat.poisson_ratio = nu
    mat.radius = radius
    mat.mass_damping = damping
    mat.barrier_distance = barrier
    mat.contact_stiffness = stiffness
    s = kn.HairSolver(lib, desc)
    s.check(s.set_hair(np.asarray(points, float), [0, len(points)], mat))
    if collider is not None:
        s.check(s.set_collider(*collider))
    s.check(s.build())
    return s, desc, mat


def _straight(n, seg=0.01, z=0.1):
    return np.array([[seg * k, 0.0, z] for k in range(n)])


def _static_state(s, q):
    """q with zero velocity and q_old = q: inertia and damping vanish."""
    s.check(s.set_state(q, np.zeros_like(q), q))


def _weights(mat):
    r = mat.radius
    area, inertia = math.pi * r * r, math.pi * r ** 4 / 4.0
    G = mat.young_modulus / (2.0 * (1.0 + mat.poisson_ratio))
    return dict(EA=mat.young_modulus * area, EI=mat.young_modulus * inertia, GJ=2.0 * G * inertia)


def test_rest_state_has_zero_energy_and_gradient(oracle_lib):
    s, _, _ = _strand_solver(oracle_lib, _straight(6))
    q, _ = s.get_state()
    _static_state(s, q)
    ev, g, _ = s.evaluate(1e-3, direction=False)
    assert ev.objective == 0.0 and ev.elastic == 0.0
    assert np.all(g == 0.0)


def test_uniform_stretch_energy(oracle_lib):
    s, _, mat = _strand_solver(oracle_lib, _straight(6))
    q, _ = s.get_state()
    x = q.reshape(-1, 6)
    lam = 1.003
    x[:, :3] = x[0, :3] + lam * (x[:, :3] - x[0, :3])
    _static_state(s, q)
    ev, _, _ = s.evaluate(1e-3, direction=False)
    L0 = 0.01
    want = 5 * 0.5 * L0 * _weights(mat)["EA"] * (lam - 1.0) ** 2
    assert abs(ev.elastic - want) <= 1e-12 * want


def test_pure_twist_energy(oracle_lib):
    s, _, mat = _strand_solver(oracle_lib, _straight(6))
    q, _ = s.get_state()
    x = q.reshape(-1, 6)
    theta = 0.07
    for k in range(x.shape[0]):
        x[k, 3:] = [k * theta, 0.0, 0.0]  # rotation about the tangent (+x)
    _static_state(s, q)
    ev, _, _ = s.evaluate(1e-3, direction=False)
    L0 = 0.01
    kappa = 2.0 * math.tan(theta / 2.0) / L0
    want = 5 * 0.5 * L0 * _weights(mat)["GJ"] * kappa ** 2
    assert abs(ev.elastic - want) <= 1e-10 * want


def _axis_rotation(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * K @ K


def test_pure_bend_energy(oracle_lib):
    """Frames rotated by k theta about the first director and chords of
    length L0 / cos(theta/2) turning by theta: zero shear and stretch, and
    curvature 2 tan(theta/2) / L0 about the first director."""
    s, _, mat = _strand_solver(oracle_lib, _straight(6))
    frames, _, _, _ = s.node_model()
    d1 = frames[0].reshape(3, 3)[:, 0]
    t = np.array([1.0, 0.0, 0.0])
    q, _ = s.get_state()
    x = q.reshape(-1, 6)
    theta, L0 = 0.05, 0.01
    chord = L0 / math.cos(theta / 2.0)
    for k in range(1, x.shape[0]):
        x[k, :3] = x[k - 1, :3] + chord * (_axis_rotation(d1, (k - 0.5) * theta) @ t)
    for k in range(x.shape[0]):
        x[k, 3:] = k * theta * d1
    _static_state(s, q)
    ev, _, _ = s.evaluate(1e-3, direction=False)
    kappa = 2.0 * math.tan(theta / 2.0) / L0
    want = 5 * 0.5 * L0 * _weights(mat)["EI"] * kappa ** 2
    assert abs(ev.elastic - want) <= 1e-9 * want


def _perturbed(s, seed, dx=2e-4, du=0.05):
    q, _ = s.get_state()
    rng = np.random.default_rng(seed)
    n = q.size // 6
    dq = np.concatenate([rng.uniform(-dx, dx, (n, 3)), rng.uniform(-du, du, (n, 3))], axis=1)
    _, _, _, flags = s.node_model()
    dq[(flags & 1) != 0] = 0.0
    return q, q + dq.ravel(), rng.uniform(-1e-3, 1e-3, q.size)


def test_gradient_matches_central_differences(oracle_lib):
    s, _, _ = _strand_solver(oracle_lib, _straight(7), damping=3.0)
    q_old, q, v = _perturbed(s, 3)
    dt = 2e-3
    s.check(s.set_state(q, v, q_old))
    _, g, _ = s.evaluate(dt, direction=False)
    _, _, _, flags = s.node_model()
    free = np.repeat((flags & 1) == 0, 6)
    rng = np.random.default_rng(4)
    for trial in range(4):
        p = rng.normal(size=q.size) * free
        p[np.arange(q.size) % 6 < 3] *= 1e-4  # positions in metres, rotations in radians
        h = 1e-4
        e = []
        for sign in (1.0, -1.0):
            s.check(s.set_state(q + sign * h * p, v, q_old))
            e.append(s.evaluate(dt, direction=False)[0].objective)
        fd = (e[0] - e[1]) / (2 * h)
        assert abs(fd - g @ p) <= 1e-6 * abs(g @ p) + 1e-18, trial


def test_gauss_newton_blocks_are_the_hessian_at_rest(oracle_lib):
    s, _, _ = _strand_solver(oracle_lib, _straight(5), damping=2.0)
    q, _ = s.get_state()
    dt = 1e-2
    v = np.zeros_like(q)
    s.check(s.set_state(q, v, q))
    s.evaluate(dt, direction=False)
    D, U = s.blocks()
    n = q.size // 6
    H = np.zeros((q.size, q.size))
    for j in range(q.size):
        h = 1e-6 if j % 6 < 3 else 1e-4
        gs = []
        for sign in (1.0, -1.0):
            qq = q.copy()
            qq[j] += sign * h
            s.check(s.set_state(qq, v, q))
            gs.append(s.evaluate(dt, direction=False)[1])
        H[:, j] = (gs[0] - gs[1]) / (2 * h)
    _, _, _, flags = s.node_model()
    free = [k for k in range(n) if not flags[k] & 1]
    for k in free:
        blk = H[6 * k:6 * k + 6, 6 * k:6 * k + 6]
        assert np.allclose(D[k], blk, rtol=1e-5, atol=1e-6 * np.abs(blk).max()), k
        assert np.allclose(D[k], D[k].T, rtol=0, atol=1e-12 * np.abs(D[k]).max())
        if k + 1 < n and k + 1 in free:
            up = H[6 * k:6 * k + 6, 6 * (k + 1):6 * (k + 1) + 6]
            assert np.allclose(U[k], up, rtol=1e-5, atol=1e-6 * np.abs(up).max()), k


def test_block_solve_matches_dense_solve(oracle_lib):
    # A soft rod with large inertia keeps the per-strand system well conditioned.
    s, _, _ = _strand_solver(oracle_lib, _straight(8), E=1.0e4, radius=5e-4, damping=1.0)
    q_old, q, v = _perturbed(s, 5)
    dt = 1e-3
    s.check(s.set_state(q, v, q_old))
    ev, g, d = s.evaluate(dt)
    assert ev.solve_failed == 0
    D, U = s.blocks()
    n = q.size // 6
    H = np.zeros((q.size, q.size))
    for k in range(n):
        H[6 * k:6 * k + 6, 6 * k:6 * k + 6] = D[k]
        if k + 1 < n:
            H[6 * k:6 * k + 6, 6 * k + 6:6 * k + 12] = U[k]
            H[6 * k + 6:6 * k + 12, 6 * k:6 * k + 6] = U[k].T
    _, _, _, flags = s.node_model()
    free_nodes = [k for k in range(n) if not flags[k] & 1]
    # The solve regularizes each Schur complement S_k by the shift
    # max(1e-12 max|S_k,ii|, 1e-24) on its diagonal (cuda_backend.cu:855-858);
    # the dense system carries the same shifts, with S_k from numpy.
    S_prev = None
    for k in free_nodes:
        S = D[k].copy()
        if S_prev is not None:
            S -= U[k - 1].T @ np.linalg.solve(S_prev, U[k - 1])
        shift = max(1e-12 * np.abs(np.diag(S)).max(), 1e-24)
        S += shift * np.eye(6)
        H[6 * k:6 * k + 6, 6 * k:6 * k + 6] += shift * np.eye(6)
        S_prev = S
    free = np.repeat((flags & 1) == 0, 6)
    Hf = H[np.ix_(free, free)]
    want = np.linalg.solve(Hf, -g[free])
    assert np.linalg.cond(Hf) < 1e9
    assert np.linalg.norm(d[free] - want) <= 1e-8 * np.linalg.norm(want)
    assert np.all(d[~free] == 0.0)


def _plane(half=0.05, z=0.0):
    v = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]])
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def test_barrier_energy_parallel_to_plane(oracle_lib):
    h, r, barrier, k = 6.0e-4, 1.0e-4, 1.0e-3, 10.0
    pts = _straight(4, seg=0.01, z=h)
    pts[:, 0] -= 0.015
    s, _, mat = _strand_solver(oracle_lib, pts, collider=_plane(), radius=r, barrier=barrier, stiffness=k)
    q, _ = s.get_state()
    _static_state(s, q)
    ev, _, _ = s.evaluate(1e-3, direction=False)
    gap = h - r - mat.collider_offset
    b = -k * (barrier - gap) ** 2 * math.log(gap / barrier)
    # the root element's left node is fixed: its contact is not evaluated
    assert abs(ev.contact - 2 * b) <= 1e-12 * abs(2 * b)
    assert abs(ev.minimum_gap - gap) <= 1e-15
    assert ev.active_contacts == 2


def test_closest_point_to_a_triangle_edge(oracle_lib):
    """Element 1 runs parallel to the edge x = 0 of a small triangle in
    z = 0, 0.5 mm outside it and 0.3 mm above: distance sqrt(0.5^2 + 0.3^2)
    mm."""
    mm = 1e-3
    tri = (np.array([[0, 0, 0], [2 * mm, 0, 0], [0, 2 * mm, 0]]), np.array([[0, 1, 2]], np.uint32))
    pts = np.array([[-0.5 * mm, -8 * mm, 0.3 * mm], [-0.5 * mm, 0.4 * mm, 0.3 * mm], [-0.5 * mm, 1.6 * mm, 0.3 * mm]])
    s, _, mat = _strand_solver(oracle_lib, pts, collider=tri, radius=1e-5, barrier=1e-3, h_max=0.05)
    q, _ = s.get_state()
    _static_state(s, q)
    ev, _, _ = s.evaluate(1e-3, direction=False)
    want = math.hypot(0.5 * mm, 0.3 * mm) - mat.radius - mat.collider_offset
    assert abs(ev.minimum_gap - want) <= 1e-15


def test_broadphase_matches_brute_force(oracle_lib):
    scene = grooms.small_sphere_scene()
    s = grooms.build_solver(oracle_lib, scene)
    q, v = s.get_state()
    # free nodes pulled onto shells 0.2..1.6 mm above the collider's vertex
    # sphere, so that many element boxes reach triangle boxes
    x = q.reshape(-1, 6)
    _, _, _, flags0 = s.node_model()
    rng = np.random.default_rng(7)
    for k in np.nonzero((flags0 & 1) == 0)[0]:
        x[k, :3] *= (0.098 + rng.uniform(2e-4, 1.6e-3)) / np.linalg.norm(x[k, :3])
    s.check(s.set_state(q, v, q))
    got = s.broadphase_pairs()
    nodes2, strand, contact, _ = s.element_model()
    _, _, _, flags = s.node_model()
    x = q.reshape(-1, 6)[:, :3]
    mat = kn.default_material(oracle_lib)
    for k_, v_ in scene.material.items():
        setattr(mat, k_, v_)
    pad = mat.radius + mat.collider_offset + mat.barrier_distance
    i, j = nodes2[:, 0], nodes2[:, 1]
    use = (contact != 0) & ((flags[i] & 1) == 0)
    elo, ehi = np.minimum(x[i], x[j]), np.maximum(x[i], x[j])
    tv = scene.vertices[scene.triangles]
    tlo, thi = tv.min(axis=1), tv.max(axis=1)
    want = []
    for e in np.nonzero(use)[0]:
        ok = np.all((elo[e] <= thi + pad) & (tlo <= ehi[e] + pad), axis=1)
        want += [(e, t) for t in np.nonzero(ok)[0]]
    want = np.array(sorted(want), np.uint32).reshape(-1, 2)
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    assert len(want) > 100
    assert np.array_equal(got, want)


@pytest.mark.parametrize("degrees", [30.0, 80.0, 170.0])
def test_root_rotation_target(oracle_lib, degrees):
    """Roots rotated rigidly about z by phi. The new root frame takes the old
    first director, projected off the new tangent, as its first director
    (solver.cpp:440): below 90 degrees this tracks the rotation and the
    target is (0, 0, phi); beyond, the projected director flips and the
    target is a half turn that still carries the old tangent onto the new
    one (the near-pi branch of the rotation-vector extraction)."""
    pts = np.array([[0.05, 0.0, 0.0], [0.06, 0.0, 0.0], [0.07, 0.0, 0.0]])
    s, desc, _ = _strand_solver(oracle_lib, pts, h_max=0.05)
    phi = math.radians(degrees)
    R = _axis_rotation([0, 0, 1], phi)
    s.check(s.allocate_animation(2))
    s.check(s.set_root_frame(0, pts))
    s.check(s.set_root_frame(1, pts @ R.T))
    s.check(s.finalize_animation())
    s.check(s.step_frame(1, 1.0 / 24.0))
    q, _ = s.get_state()
    u = q.reshape(-1, 6)[0, 3:]
    if degrees < 90.0:
        assert np.allclose(u, [0.0, 0.0, phi], rtol=0, atol=1e-12)
    else:
        angle = np.linalg.norm(u)
        assert abs(angle - math.pi) <= 1e-12
        t_new = _axis_rotation(u, angle) @ np.array([1.0, 0.0, 0.0])
        assert np.allclose(t_new, R @ np.array([1.0, 0.0, 0.0]), rtol=0, atol=1e-12)
