"""Pins of the oracle's contact, inertia and fallback stages against what
the mathematics fixes, through the stage hooks (khsxApplyFriction,
khsxStepBound, khsxSweepCheck, khsxSetDirectionMode):

  friction     a node resting on a plane: slip branch |dv| = mu f_n dt / m,
               stick branch dv = -v_t, both times the smoothing factor
               v_t / sqrt(v_t^2 + (eps / dt)^2); the normal force f_n equals
               -dE_contact/dgap by central differences of the oracle's own
               energy; the normal velocity is untouched; a plane sliding
               under a resting node drags it along with the body velocity
               (x_cur - x_old) / dt (reference cuda_backend.cu:941-964)
  step bound   a segment driven straight at a plane: alpha is the time of
               impact at the target clearance (conservative advancement
               converges on it) and scales as 1 / |d|; random small
               segment/triangle motions: brute-force time sampling with
               textbook distances never finds the clearance violated before
               alpha (cuda_backend.cu:989-1000)
  sweep        a plane moving at a resting segment: feasible exactly when
               the final clearance stays above the target, with the failing
               pair's diagnostics in closed form; random motions: a sampled
               violation is always reported (cuda_backend.cu:966-987)
  inertia      a strand translated rigidly (no elastic energy): objective,
               gradient and kinetic energy of the implicit-Euler inertia
               and mass-proportional damping with gravity in closed form
               (cuda_backend.cu:883-897)
  fallback     -P g with the Jacobi diagonal P in closed form; its axial
               translation entry equals the exact Hessian diagonal by central
               differences of the gradient; forced steepest descent steps a
               frame and is counted (cuda_backend.cu:815-846, 632-638)
"""
import math

import numpy as np
import pytest

import geometry_ref as geo
import khs_native as kn

R, BARRIER, KAPPA = 1.0e-4, 1.0e-3, 10.0


def _solver(lib, points, *, collider=None, fixed=1, friction=0.3, smoothing=1e-7, damping=0.0, gravity=(0, 0, 0),
            E=3.0e6, nu=0.3, h_max=0.05, density=1300.0, **desc_over):
    desc = kn.default_desc(lib)
    for k, v in desc_over.items():
        setattr(desc, k, v)
    desc.gravity = kn.KhsVec3(*gravity)
    desc.fixed_root_nodes = fixed
    desc.maximum_element_length = h_max
    desc.minimum_dynamic_length = 0.0
    mat = kn.default_material(lib)
    mat.density = density
    mat.young_modulus = E
    mat.poisson_ratio = nu
    mat.radius = R
    mat.mass_damping = damping
    mat.barrier_distance = BARRIER
    mat.contact_stiffness = KAPPA
    mat.friction = friction
    mat.friction_smoothing = smoothing
    s = kn.HairSolver(lib, desc)
    s.check(s.set_hair(np.asarray(points, float), [0, len(points)], mat))
    if collider is not None:
        s.check(s.set_collider(*collider))
    s.check(s.build())
    return s, desc, mat


def _plane(half=1.0, z=0.0):
    v = np.array([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]])
    return v, np.array([[0, 1, 2], [0, 2, 3]], np.uint32)


def _resting(gap):
    """Fixed root high up, then a free node, then a free node resting `gap`
    above the plane z = 0: the element (1, 2) comes closest at node 2."""
    return np.array([[0.0, 0.0, 0.02], [0.01, 0.0, 0.01], [0.02, 0.0, R + gap]])


# ----------------------------------------------------------------- friction
def _friction_dv(lib, gap, v2, *, friction, smoothing=1e-7, dt=1e-3, plane_velocity=(0.0, 0.0, 0.0)):
    pts = _resting(gap)
    s, _, _ = _solver(lib, pts, collider=_plane(), friction=friction, smoothing=smoothing)
    q, _ = s.get_state()
    v = np.zeros_like(q)
    v[12:15] = v2
    s.check(s.set_state(q, v, q))
    cur, _ = _plane()
    s.check(s.set_collider_state(cur, cur - np.asarray(plane_velocity) * dt))
    s.check(s.apply_friction(dt))
    _, after = s.get_state()
    _, mass, _, _ = s.node_model()
    return (after - v).reshape(-1, 6), mass, s


def _normal_force(lib, gap, h=1e-7):
    """-dE_contact/dgap by central differences of the oracle's energy."""
    e = []
    for sign in (1.0, -1.0):
        s, _, _ = _solver(lib, _resting(gap), collider=_plane())
        q, _ = s.get_state()
        q[14] += sign * h
        s.check(s.set_state(q, np.zeros_like(q), q))
        e.append(s.evaluate(1e-3, direction=False)[0].contact)
        s.close()
    return -(e[0] - e[1]) / (2 * h)


def test_friction_slip_branch(oracle_lib):
    gap, dt, mu, eps = 4.0e-4, 1e-3, 0.002, 1e-7
    vt = np.array([0.3, 0.4, 0.0])  # 0.5 m/s along the plane
    v2 = vt + np.array([0.0, 0.0, -0.05])
    dv, mass, _ = _friction_dv(oracle_lib, gap, v2, friction=mu, smoothing=eps, dt=dt)
    fn = _normal_force(oracle_lib, gap)
    ts = np.linalg.norm(vt)
    assert mu * fn * dt < ts * mass[2]  # the slip branch
    want = -vt / ts * (mu * fn * dt / mass[2]) * ts / math.hypot(ts, eps / dt)
    assert np.allclose(dv[2, :3], want, rtol=1e-6, atol=1e-12 * np.linalg.norm(want))
    # the normal component is untouched (up to the rounding of the normal)
    assert abs(dv[2, 2]) <= 1e-12 * np.linalg.norm(want)
    assert np.all(dv[:2] == 0.0) and np.all(dv[:, 3:] == 0.0)


def test_friction_stick_branch(oracle_lib):
    gap, dt, mu, eps = 4.0e-4, 1e-3, 0.3, 1e-7
    vt = np.array([0.03, -0.04, 0.0])
    dv, mass, _ = _friction_dv(oracle_lib, gap, vt, friction=mu, smoothing=eps, dt=dt)
    fn = _normal_force(oracle_lib, gap)
    ts = np.linalg.norm(vt)
    assert mu * fn * dt > ts * mass[2]  # the stick branch
    # the impulse stops the tangential motion, up to the smoothing factor
    want = -vt * ts / math.hypot(ts, eps / dt)
    assert np.allclose(dv[2, :3], want, rtol=1e-12, atol=1e-12 * ts)
    assert np.allclose(dv[2, :3], -vt, rtol=1e-5, atol=1e-12 * ts)


def test_friction_drags_hair_with_a_moving_body(oracle_lib):
    u = np.array([0.02, 0.01, 0.0])
    dv, _, _ = _friction_dv(oracle_lib, 4.0e-4, np.zeros(3), friction=0.3, plane_velocity=u)
    assert np.allclose(dv[2, :3], u, rtol=1e-5, atol=1e-12 * np.linalg.norm(u))


def test_friction_outside_the_barrier_does_nothing(oracle_lib):
    dv, _, _ = _friction_dv(oracle_lib, 1.2 * BARRIER, np.array([0.3, 0.4, 0.0]), friction=0.3)
    assert np.all(dv == 0.0)


# --------------------------------------------------------------- step bound
def _bound_solver(lib, height, collider=None):
    pts = np.array([[0.0, 0.0, 0.02], [0.01, 0.0, height], [0.02, 0.0, height]])
    return _solver(lib, pts, collider=collider if collider is not None else _plane())


def test_step_bound_is_time_of_impact_on_a_plane(oracle_lib):
    h = 3.0e-3
    s, desc, mat = _bound_solver(oracle_lib, h)
    target = mat.radius + mat.collider_offset + desc.minimum_gap
    q, _ = s.get_state()
    for speed in (5.0e-3, 1.0e-2, 4.0e-2):
        d = np.zeros_like(q)
        d[8] = d[14] = -speed  # nodes 1 and 2 straight down
        alpha = s.step_bound(d)
        toi = (h - target) / speed
        if toi >= 1.0:
            assert alpha == 1.0
        else:
            assert toi - 1e-9 <= alpha <= toi, (speed, alpha, toi)
    # alpha scales as 1 / |d| (advancement stops once a step falls below
    # 1e-12, which bounds the difference)
    d = np.zeros_like(q)
    d[8] = d[14] = -0.01
    d[6] = d[12] = 0.004
    a1, a2 = s.step_bound(d), s.step_bound(2 * d)
    assert a1 < 1.0 and abs(a2 - a1 / 2) <= 1e-11
    # moving away from the plane is never bounded
    assert s.step_bound(-d) == 1.0


def test_step_bound_never_passes_a_sampled_impact(oracle_lib):
    rng = np.random.default_rng(11)
    mm = 1e-3
    checked = bounded = 0
    for trial in range(40):
        tri = rng.uniform(-2 * mm, 2 * mm, (3, 3))
        tri[:, 2] = rng.uniform(-0.3 * mm, 0.3 * mm, 3)
        p1 = np.array([rng.uniform(-2, 2) * mm, rng.uniform(-2, 2) * mm, rng.uniform(1.0, 2.5) * mm])
        p2 = p1 + rng.normal(size=3) * mm
        p2[2] = max(p2[2], 1.0 * mm)
        pts = np.array([p1 + [0.0, 0.0, 0.02], p1, p2])
        s, desc, mat = _solver(oracle_lib, pts, collider=(tri, np.array([[0, 1, 2]], np.uint32)))
        target = mat.radius + mat.collider_offset + desc.minimum_gap
        q, _ = s.get_state()
        d = np.zeros_like(q)
        d[6:9] = rng.normal(size=3) * 2 * mm - [0, 0, 2 * mm]
        d[12:15] = rng.normal(size=3) * 2 * mm - [0, 0, 2 * mm]
        a = s.step_bound(d)
        x1, x2 = q[6:9], q[12:15]
        ts = np.linspace(0.0, a, 401)
        if any(geo.pierces(x1 + t * d[6:9], x2 + t * d[12:15], *tri) for t in ts):
            continue
        clear = [geo.segment_triangle(x1 + t * d[6:9], x2 + t * d[12:15], *tri) - target for t in ts]
        assert min(clear) > -1e-15, (trial, a, min(clear))
        checked += 1
        bounded += a < 1.0
        s.close()
    assert checked >= 30 and bounded >= 5


# -------------------------------------------------------------------- sweep
def test_sweep_plane_closed_form(oracle_lib):
    h = 2.0e-3
    s, desc, mat = _bound_solver(oracle_lib, h)
    target = mat.radius + mat.collider_offset + desc.minimum_gap
    old, _ = _plane()
    for lift in (0.5e-3, 1.5e-3, 1.85e-3, 1.95e-3, 2.5e-3):
        cur = old + [0.0, 0.0, lift]
        s.check(s.set_collider_state(cur, old))
        ok, diag = s.sweep_check()
        want_ok = h - lift - target > 1e-6
        assert ok == want_ok, lift
        if not ok:
            assert diag.kind == kn.FAILURE_SWEEP
            assert diag.element_index == 1 and diag.strand_index == 0
            assert abs(diag.required_distance - target) <= 1e-18
            assert abs(diag.collider_substep_displacement - lift) <= 1e-15
            assert abs(diag.clearance - (diag.distance - target)) <= 1e-18
            assert diag.clearance <= 1e-9
    # a plane moving away, or sideways within its own plane, never fails
    for motion in ([0.0, 0.0, -1.0e-3], [5.0e-3, 0.0, 0.0]):
        s.check(s.set_collider_state(old + motion, old))
        assert s.sweep_check()[0]


def test_sweep_reports_every_sampled_violation(oracle_lib):
    rng = np.random.default_rng(12)
    mm = 1e-3
    seen = {True: 0, False: 0}
    for trial in range(40):
        tri_old = rng.uniform(-2 * mm, 2 * mm, (3, 3))
        tri_old[:, 2] = rng.uniform(-0.3 * mm, 0.3 * mm, 3)
        motion = rng.normal(size=3) * mm + [0, 0, 2.0 * mm]
        tri_new = tri_old + motion + rng.normal(size=(3, 3)) * 0.2 * mm
        p1 = np.array([rng.uniform(-1, 1) * mm, rng.uniform(-1, 1) * mm, rng.uniform(1.5, 3.0) * mm])
        p2 = p1 + rng.normal(size=3) * mm
        p2[2] = max(p2[2], 1.5 * mm)
        pts = np.array([p1 + [0.0, 0.0, 0.02], p1, p2])
        s, desc, mat = _solver(oracle_lib, pts, collider=(tri_old, np.array([[0, 1, 2]], np.uint32)))
        target = mat.radius + mat.collider_offset + desc.minimum_gap
        s.check(s.set_collider_state(tri_new, tri_old))
        ok, _ = s.sweep_check()
        samples = [tri_old + t * (tri_new - tri_old) for t in np.linspace(0.0, 1.0, 401)]
        if any(geo.pierces(p1, p2, *tri) for tri in samples):
            violated = True
        else:
            violated = min(geo.segment_triangle(p1, p2, *tri) for tri in samples) - target <= 0.0
        if violated:
            assert not ok, trial
        seen[ok] += 1
        s.close()
    assert seen[True] >= 5 and seen[False] >= 5


# ------------------------------------------------------------------ inertia
def test_inertia_with_gravity_and_damping_closed_form(oracle_lib):
    pts = np.array([[0.01 * k, 0.0, 0.0] for k in range(4)])
    gravity = np.array([0.3, -1.2, -9.81])
    damping, dt = 4.0, 2e-3
    s, _, mat = _solver(oracle_lib, pts, damping=damping, gravity=gravity)
    q_rest, _ = s.get_state()
    rng = np.random.default_rng(13)
    shift = rng.uniform(-1e-3, 1e-3, 3)
    q = q_rest.copy().reshape(-1, 6)
    q[:, :3] += shift  # a rigid translation: no elastic energy
    q = q.ravel()
    v = rng.uniform(-0.05, 0.05, q.size)
    s.check(s.set_state(q, v, q_rest))
    ev, g, _ = s.evaluate(dt, direction=False)
    frames, mass, rmass, flags = s.node_model()
    # lumped masses: half of each adjacent element, rho pi r^2 L0 and
    # rho (4/3)(pi r^4 / 4) L0 for the rotational mass
    L0 = 0.01
    m_elem = mat.density * math.pi * R * R * L0
    rm_elem = mat.density * (4.0 / 3.0) * (math.pi * R ** 4 / 4.0) * L0
    ends = np.array([1, 2, 2, 1]) / 2.0
    assert np.allclose(mass, m_elem * ends, rtol=1e-14, atol=0)
    assert np.allclose(rmass, rm_elem * ends, rtol=1e-14, atol=0)
    want_e = want_k = 0.0
    want_g = np.zeros_like(q)
    for n in range(4):
        if flags[n] & 1:
            continue
        for d in range(6):
            i = 6 * n + d
            m = mass[n] if d < 3 else rmass[n]
            pred = q_rest[i] + dt * v[i] + (gravity[d] * dt * dt if d < 3 else 0.0)
            k, c = m / dt ** 2, m * damping / dt
            want_e += 0.5 * k * (q[i] - pred) ** 2 + 0.5 * c * (q[i] - q_rest[i]) ** 2
            want_g[i] = k * (q[i] - pred) + c * (q[i] - q_rest[i])
            if d < 3:
                want_k += 0.5 * mass[n] * (q[i] - q_rest[i]) ** 2 / dt ** 2
    assert abs(ev.elastic) <= 1e-12 * want_e
    assert abs(ev.objective - want_e) <= 1e-10 * want_e
    assert abs(ev.kinetic - want_k) <= 1e-12 * want_k
    assert np.allclose(g, want_g, rtol=1e-9, atol=1e-12 * np.abs(want_g).max())


# ----------------------------------------------------------------- fallback
def test_steepest_descent_is_jacobi_preconditioned(oracle_lib):
    pts = np.array([[0.01 * k, 0.0, 0.0] for k in range(5)])
    damping, dt = 2.0, 1e-3
    s, _, mat = _solver(oracle_lib, pts, damping=damping)
    q_rest, _ = s.get_state()
    rng = np.random.default_rng(14)
    q = q_rest + np.concatenate([rng.uniform(-1e-4, 1e-4, (5, 3)), rng.uniform(-0.02, 0.02, (5, 3))], 1).ravel()
    q[:6] = q_rest[:6]
    v = rng.uniform(-0.01, 0.01, q.size)
    s.check(s.set_state(q, v, q_rest))
    s.check(s.set_direction_mode(1))
    _, g, d = s.evaluate(dt)
    _, mass, rmass, flags = s.node_model()
    area, inertia = math.pi * R * R, math.pi * R ** 4 / 4.0
    G = mat.young_modulus / (2.0 * (1.0 + mat.poisson_ratio))
    EA, kGA = mat.young_modulus * area, mat.shear_correction * G * area
    EI, GJ = mat.young_modulus * inertia, 2.0 * G * inertia
    L0 = 0.01
    P = np.zeros(q.size)
    for n in range(5):
        if flags[n] & 1:
            continue
        elems = (n > 0) + (n < 4)
        t = mass[n] * (1 / dt ** 2 + damping / dt) + elems * max(EA, kGA) / L0
        r = rmass[n] * (1 / dt ** 2 + damping / dt) + elems * (kGA * L0 + max(EI, GJ) / L0)
        P[6 * n:6 * n + 3] = 1.0 / t
        P[6 * n + 3:6 * n + 6] = 1.0 / r
    assert np.allclose(d, -P * g, rtol=1e-13, atol=0)
    assert np.all(d[:6] == 0.0) and g @ d < 0.0
    # The axial translation entry of the diagonal is exact: at rest the
    # Hessian's (x, x) entry of a node is m (1/dt^2 + c/dt) + EA/L0 per
    # element, measured here by central differences of the gradient.
    s.check(s.set_state(q_rest, np.zeros_like(q), q_rest))
    for n in (2, 4):
        i, h = 6 * n, 1e-7
        gs = []
        for sign in (1.0, -1.0):
            qq = q_rest.copy()
            qq[i] += sign * h
            s.check(s.set_state(qq, np.zeros_like(q), q_rest))
            gs.append(s.evaluate(dt, direction=False)[1][i])
        hxx = (gs[0] - gs[1]) / (2 * h)
        assert abs(hxx * P[i] - 1.0) <= 1e-6, n


def test_forced_steepest_descent_steps_and_counts(oracle_lib):
    pts = np.array([[0.01 * k, 0.0, 0.0] for k in range(6)])
    runs = []
    for mode in (0, 1):
        s, desc, _ = _solver(oracle_lib, pts, damping=2.0, gravity=(0, 0, -9.81), E=1.0e7, substeps=2,
                             newton_iterations=40, relative_tolerance=1e-6, absolute_tolerance=1e-30,
                             increment_tolerance=1e-14)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_direction_mode(mode))
        s.check(s.step(1.0 / 60.0))
        runs.append((s.internal_positions(), s.direction_counts(), s.step_stats()))
    (x0, c0, st0), (x1, c1, st1) = runs
    assert c0[0] > 0 and c0[1] == 0  # the block direction serves a plain frame
    assert c1[0] == 0 and c1[1] == st1.linear_solves > 0
    assert st1.phase == kn.PHASE_FINISHED
    # both fall under gravity from the same start; steepest descent is slower
    # to converge, so it is compared only in direction of travel
    assert np.all(x1[1:, 2] < 0.0) and np.all(x0[1:, 2] < 0.0)
