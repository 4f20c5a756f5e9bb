"""Stage-by-stage and moving-collider parity of the HIP step with the oracle.

Each stage hook (friction, step bound, sweep, steepest descent) runs on both
libraries from the same state; the oracle's side of each is pinned by
test_oracle_stage_pins.py. Bars: 1e-10 relative for energies, gradients,
blocks, velocity changes and directions; step bounds within 1e-12; sweep
verdicts and reported pairs equal; positions after each frame within 1e-7 m
with equal Newton, line-search, candidate and active counts.
"""
import math

import numpy as np
import pytest

import grooms
import khs_native as kn

pytestmark = pytest.mark.gpu

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _pulled_in(solver, seed=7, lo=5e-5, hi=2e-4):
    """Free nodes of the sphere scene pulled onto shells lo..hi above the
    collider's vertex sphere: by default inside the 0.2 mm barrier, so that
    many barriers are active (about 150); 0.2-1.6 mm for the continuous
    collision stages, which need room to move."""
    q, v = solver.get_state()
    x = q.reshape(-1, 6)
    _, _, _, flags = solver.node_model()
    rng = np.random.default_rng(seed)
    for k in np.nonzero((flags & 1) == 0)[0]:
        x[k, :3] *= (0.098 + rng.uniform(lo, hi)) / np.linalg.norm(x[k, :3])
    return q, rng


def _sphere_pair(oracle_lib, gpu_lib, **desc):
    scene = grooms.small_sphere_scene()
    out = []
    for lib in (oracle_lib, gpu_lib):
        s = grooms.build_solver(lib, scene, **desc)
        out.append(s)
    return scene, out[0], out[1]


def test_active_barrier_evaluation_matches_oracle(oracle_lib, gpu_lib):
    print("[gpu] evaluation with active barriers, sphere scene pulled in", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    q, rng = _pulled_in(o)
    v = rng.uniform(-1e-3, 1e-3, q.size)
    dt = scene.frame_dt / 4
    res = []
    for s in (o, g):
        s.check(s.set_state(q, v, q - 1e-4 * v))
        ev, grad, d = s.evaluate(dt)
        D, U = s.blocks()
        res.append((ev, grad, d, D, U))
    (eo, go, do, Do, Uo), (eg, gg, dg, Dg, Ug) = res
    print(f"active oracle {eo.active_contacts} gpu {eg.active_contacts}; candidates {eo.candidate_pairs} "
          f"{eg.candidate_pairs}; contact energy {eo.contact:.6e} {eg.contact:.6e}")
    assert eo.active_contacts > 100 and eg.active_contacts == eo.active_contacts
    assert eg.candidate_pairs == eo.candidate_pairs
    assert eg.feasible == eo.feasible == 1
    for f in ("objective", "elastic", "contact", "kinetic"):
        assert abs(getattr(eg, f) - getattr(eo, f)) <= 1e-10 * abs(getattr(eo, f)), f
    assert eg.minimum_gap == eo.minimum_gap
    assert _rel(gg, go) <= 1e-10
    assert max(_rel(Dg, Do), _rel(Ug, Uo)) <= 1e-10
    # the direction by its backward error in the oracle's system (the
    # per-strand systems are too ill-conditioned for a forward bar)
    from test_gpu_parity import _newton_residual
    res_o = _newton_residual(Do, Uo, go, do, o)
    res_g = _newton_residual(Do, Uo, go, dg, o)
    print(f"direction backward error oracle {res_o:.3e} gpu {res_g:.3e}; forward {_rel(dg, do):.3e}")
    assert res_g <= max(10 * res_o, 1e-10)


def test_friction_stage_matches_oracle(oracle_lib, gpu_lib):
    """Friction against a collider that moved (body velocity nonzero) at a
    state with many elements inside the barrier."""
    print("[gpu] friction stage, moving sphere collider", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    q, rng = _pulled_in(o, seed=8)
    v = rng.uniform(-0.05, 0.05, q.size)
    phi = 0.002
    c, s_ = math.cos(phi), math.sin(phi)
    prev = scene.vertices @ np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]).T
    dv = []
    for s in (o, g):
        s.check(s.set_state(q, v, q))
        s.check(s.set_collider_state(scene.vertices, prev))
        s.check(s.apply_friction(scene.frame_dt / 4))
        dv.append(s.get_state()[1] - v)
    moved_o = np.count_nonzero(np.abs(dv[0]) > 0)
    moved_g = np.count_nonzero(np.abs(dv[1]) > 0)
    print(f"dofs changed oracle {moved_o} gpu {moved_g}; rel {_rel(dv[1], dv[0]):.3e}")
    assert moved_o > 50 and moved_g == moved_o
    assert _rel(dv[1], dv[0]) <= 1e-10


def test_friction_closed_form_cases_match_oracle(oracle_lib, gpu_lib):
    """The single-node slip, stick and moving-plane cases the oracle pins."""
    print("[gpu] friction closed-form cases", flush=True)
    import test_oracle_stage_pins as pins

    cases = [dict(gap=4e-4, v2=np.array([0.3, 0.4, -0.05]), friction=0.002),
             dict(gap=4e-4, v2=np.array([0.03, -0.04, 0.0]), friction=0.3),
             dict(gap=4e-4, v2=np.zeros(3), friction=0.3, plane_velocity=np.array([0.02, 0.01, 0.0]))]
    for case in cases:
        got = [pins._friction_dv(lib, **case)[0] for lib in (oracle_lib, gpu_lib)]
        assert np.abs(got[0]).max() > 0
        assert _rel(got[1], got[0]) <= 1e-12


def test_step_bound_matches_oracle(oracle_lib, gpu_lib):
    print("[gpu] step bound, sphere scene pulled in", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    q, rng = _pulled_in(o, seed=9, lo=2e-4, hi=1.6e-3)
    _, _, _, flags = o.node_model()
    free = np.repeat((flags & 1) == 0, 6)
    for trial in range(3):
        d = rng.normal(size=q.size) * free
        d.reshape(-1, 6)[:, :3] *= 2e-3  # up to a few mm of travel
        d.reshape(-1, 6)[:, 3:] *= 0.05
        alphas = []
        for s in (o, g):
            s.check(s.set_state(q, np.zeros_like(q), q))
            alphas.append(s.step_bound(d))
        print(f"trial {trial}: alpha oracle {alphas[0]!r} gpu {alphas[1]!r}")
        assert alphas[0] < 1.0
        assert abs(alphas[1] - alphas[0]) <= 1e-12


def test_step_bound_plane_matches_oracle(oracle_lib, gpu_lib):
    import test_oracle_stage_pins as pins

    print("[gpu] step bound, segment driven at a plane", flush=True)
    out = []
    for lib in (oracle_lib, gpu_lib):
        s, _, _ = pins._bound_solver(lib, 3.0e-3)
        q, _ = s.get_state()
        d = np.zeros_like(q)
        d[8] = d[14] = -0.01
        d[6] = d[12] = 0.004
        out.append((s.step_bound(d), s.step_bound(2 * d), s.step_bound(-d)))
    assert out[1] == out[0]


def test_step_bound_with_overflowed_lists_matches_oracle(oracle_lib, gpu_lib):
    """A strand 1 mm above a plane of 0.4 mm quads, where an element's thin
    box holds more triangles than a list can: the elements whose lists
    overflow at the state, and those whose swept box leaves their list and
    overflow when regathered, are bounded by walks of the swept box."""
    import test_gpu_parity as par

    print("[gpu] step bound with overflowed lists, strand over a 0.4 mm grid", flush=True)
    verts, tris = par._grid_plane(200, 4.0e-4)
    pts = np.array([[-0.03 + 0.01 * k, 0.0013 * k, 0.006] for k in range(7)])
    solvers = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.substeps = 4
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.01
        mat = kn.default_material(lib)
        mat.radius = 1.0e-4
        mat.barrier_distance = 2.0e-3
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, [0, len(pts)], mat))
        s.check(s.set_collider(verts, tris))
        s.check(s.build())
        solvers.append(s)
    q, _ = solvers[0].get_state()
    x = q.reshape(-1, 6)
    _, _, _, flags = solvers[0].node_model()
    free = (flags & 1) == 0
    x[free, 2] = 1.0e-3
    rng = np.random.default_rng(12)
    for trial in range(3):
        d = np.zeros_like(q)
        dv = d.reshape(-1, 6)
        dv[free, 2] = -rng.uniform(1.5e-3, 3.0e-3, int(free.sum()))
        dv[free, :2] = rng.normal(scale=2.0e-3, size=(int(free.sum()), 2))
        alphas = []
        for s in solvers:
            s.check(s.set_state(q, np.zeros_like(q), q))
            alphas.append(s.step_bound(d))
        lists = solvers[1].contact_list_stats()
        print(f"trial {trial}: alpha oracle {alphas[0]!r} gpu {alphas[1]!r}; overflowed lists {lists.overflowed}")
        assert lists.overflowed > 0
        assert alphas[0] < 1.0
        assert abs(alphas[1] - alphas[0]) <= 1e-12


def test_sweep_matches_oracle(oracle_lib, gpu_lib):
    """The collider swelling outward at the pulled-in state: small swells
    pass, larger ones fail on the same (element, triangle)."""
    print("[gpu] collider sweep, sphere scene pulled in", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    q, _ = _pulled_in(o, seed=10, lo=2e-4, hi=1.6e-3)
    verdicts = []
    for swell in (1.0005, 1.002, 1.01, 1.03):
        res = []
        for s in (o, g):
            s.check(s.set_state(q, np.zeros_like(q), q))
            s.check(s.set_collider_state(scene.vertices * swell, scene.vertices))
            res.append(s.sweep_check())
        (ok_o, fo), (ok_g, fg) = res
        print(f"swell {swell}: oracle {ok_o} ({fo.element_index}, {fo.collider_triangle_index}) "
              f"gpu {ok_g} ({fg.element_index}, {fg.collider_triangle_index})")
        assert ok_g == ok_o
        if not ok_o:
            assert (fg.strand_index, fg.element_index, fg.collider_triangle_index) == \
                   (fo.strand_index, fo.element_index, fo.collider_triangle_index)
            assert abs(fg.distance - fo.distance) <= 1e-15
            assert abs(fg.collider_substep_displacement - fo.collider_substep_displacement) <= 1e-18
        verdicts.append(ok_o)
    assert True in verdicts and False in verdicts


def test_steepest_direction_matches_oracle(oracle_lib, gpu_lib):
    print("[gpu] forced steepest-descent direction", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    q, rng = _pulled_in(o, seed=11)
    v = rng.uniform(-1e-3, 1e-3, q.size)
    out = []
    for s in (o, g):
        s.check(s.set_state(q, v, q))
        s.check(s.set_direction_mode(1))
        out.append(s.evaluate(scene.frame_dt / 4))
    (_, go, do), (_, gg, dg) = out
    assert np.dot(go, do) < 0
    assert _rel(dg, do) <= 1e-10


def _frames_match(results, label):
    (xo, so, co), (xg, sg, cg) = results
    err = float(np.max(np.abs(xo - xg)))
    print(f"{label}: max position error {err:.3e} m; newton {so[0]} {sg[0]}; line search {so[1]} {sg[1]}; "
          f"directions oracle {co} gpu {cg}")
    assert err <= 1e-7
    assert sg == so
    assert cg == co


def _trust_strand(lib, min_step):
    pts = np.array([[0.01 * k, 0.0, 0.0] for k in range(8)])
    desc = kn.default_desc(lib)
    desc.substeps = 1
    desc.maximum_substeps = 1
    desc.fixed_root_nodes = 1
    desc.maximum_element_length = 0.02
    desc.newton_iterations = 12
    desc.minimum_line_search_step = min_step
    mat = kn.default_material(lib)
    mat.young_modulus = 1.0e7
    mat.radius = 1.0e-4
    s = kn.HairSolver(lib, desc)
    s.check(s.set_hair(pts, [0, len(pts)], mat))
    s.check(s.build())
    s.check(s.set_convergence_mode(1))
    return s


def test_natural_fallback_frames_match_oracle(oracle_lib, gpu_lib):
    """Trust-limited frames with a minimum line-search step of 0.6: a block
    direction whose Armijo test needs a shorter step is rejected and the
    steepest-descent retry takes over (reference cuda_backend.cu:668-676)."""
    print("[gpu] natural steepest-descent fallback", flush=True)
    results = []
    for lib in (oracle_lib, gpu_lib):
        s = _trust_strand(lib, 0.6)
        stats, counts = [], []
        for _ in range(2):
            s.check(s.step(0.25))
            st = s.step_stats()
            stats.append((st.newton_iterations, st.line_search_evaluations, st.linear_solves))
            counts.append(s.direction_counts())
        results.append((s.internal_positions(), tuple(zip(*stats)), counts))
    _frames_match(results, "fallback")
    assert all(c[1] > 0 for c in results[1][2])  # the GPU took the fallback in each frame


def test_forced_steepest_frames_match_oracle(oracle_lib, gpu_lib):
    print("[gpu] forced steepest descent, two frames", flush=True)
    results = []
    for lib in (oracle_lib, gpu_lib):
        s = _trust_strand(lib, 1e-9)
        s.check(s.set_direction_mode(1))
        stats, counts = [], []
        for _ in range(2):
            s.check(s.step(0.25))
            st = s.step_stats()
            stats.append((st.newton_iterations, st.line_search_evaluations, st.linear_solves))
            counts.append(s.direction_counts())
        results.append((s.internal_positions(), tuple(zip(*stats)), counts))
    _frames_match(results, "forced steepest")
    assert all(c[0] == 0 and c[1] > 0 for c in results[1][2])


def _grid_plane(n, h):
    xs = (np.arange(n + 1) - n / 2) * h
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    verts = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            tris += [[a, b, c], [a, c, d]]
    return verts, np.array(tris, np.uint32)


def test_turntable_contact_frames_match_oracle(oracle_lib, gpu_lib):
    """Three strands draped on a tessellated plane that turns under them at
    60 deg/s and rises 0.3 mm per frame: the swept tree with current-box
    leaves, the sweep's success path on a moving mesh, and friction with the
    body velocity, over four animation frames."""
    print("[gpu] turntable contact frames", flush=True)
    verts, tris = _grid_plane(20, 5e-3)
    strands, offsets = [], [0]
    for k in range(3):
        y = -0.012 + 0.012 * k
        strands.append(np.array([[-0.035 + 0.01 * i, y + 0.0013 * i, 0.006] for i in range(7)]))
        offsets.append(offsets[-1] + 7)
    pts = np.concatenate(strands)
    spin, rise, frames = math.radians(60.0) / 30.0, 3.0e-4, 4

    def collider(f):
        c, s_ = math.cos(spin * f), math.sin(spin * f)
        return verts @ np.array([[c, -s_, 0], [s_, c, 0], [0, 0, 1]]).T + [0.0, 0.0, rise * f]

    results = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.substeps = 4
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.01
        desc.newton_iterations = 60
        for k, v in LIFTED.items():
            setattr(desc, k, v)
        mat = kn.default_material(lib)
        mat.radius = 1.0e-4
        mat.young_modulus = 1.0e7
        mat.contact_stiffness = 100.0
        mat.barrier_distance = 2.0e-3
        mat.mass_damping = 2.0
        mat.friction = 0.3
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, offsets, mat))
        s.check(s.set_collider(verts, tris))
        s.check(s.build())
        s.check(s.set_convergence_mode(1))
        s.check(s.allocate_animation(frames + 1))
        for f in range(frames + 1):
            s.check(s.set_root_frame(f, pts))
            s.check(s.set_collider_frame(f, collider(f)))
        s.check(s.finalize_animation())
        stats, counts = [], []
        for f in range(1, frames + 1):
            s.check(s.step_frame(f, 1.0 / 30.0))
            st = s.step_stats()
            stats.append((st.newton_iterations, st.line_search_evaluations, st.contact_candidate_count,
                          st.active_contact_count, st.substeps))
            counts.append(s.direction_counts())
            assert st.minimum_gap > desc.minimum_gap
        results.append((s.internal_positions(), tuple(zip(*stats)), counts))
        s.close()
    _frames_match(results, "turntable")
    assert sum(results[1][1][3]) > 0  # barriers were active


def _groom_solver(lib, scene, frames, **desc):
    s = grooms.build_solver(lib, scene, **{**LIFTED, **desc})
    s.check(s.set_convergence_mode(1))
    s.set_threads(16)
    s.check(s.allocate_animation(frames))
    for f in range(frames):
        s.check(s.set_root_frame(f, scene.frame_points(f)))
        s.check(s.set_collider_frame(f, scene.frame_vertices(f)))
    s.check(s.finalize_animation())
    return s


GROOM_STRANDS, GROOM_WARM = 600, 12


@pytest.fixture(scope="module")
def groom_subset():
    return grooms.head_scene(strands=GROOM_STRANDS)


@pytest.fixture(scope="module")
def groom_contact_checkpoint(gpu_lib, groom_subset):
    """The groom subset's checkpoint after frames 1-12 stepped on the GPU,
    when its hair lies on the body."""
    scene = groom_subset
    print(f"[gpu] groom subset: frames 1-{GROOM_WARM} on the GPU", flush=True)
    s = _groom_solver(gpu_lib, scene, GROOM_WARM + 1)
    for f in range(1, GROOM_WARM + 1):
        s.check(s.step_frame(f, scene.frame_dt))
    state = s.save_checkpoint()
    s.close()
    return state


def test_groom_subset_first_frames_match_oracle(oracle_lib, gpu_lib, groom_subset):
    """The first 600 strands of the 6,757-strand groom on the full rotating
    225,184-vertex body, frames 1-2 at 8 substeps to the 1e-6 residual with
    the floors lifted: positions and every count."""
    scene = groom_subset
    print(f"[gpu] groom subset: {scene.strand_count} strands, frames 1-2 on both", flush=True)
    results = []
    for lib in (oracle_lib, gpu_lib):
        s = _groom_solver(lib, scene, 3)
        stats, counts = [], []
        for f in (1, 2):
            s.check(s.step_frame(f, scene.frame_dt))
            st = s.step_stats()
            stats.append((st.newton_iterations, st.line_search_evaluations, st.contact_candidate_count,
                          st.active_contact_count, st.substeps))
            counts.append(s.direction_counts())
            assert st.minimum_gap > 1e-7
        results.append((s.internal_positions(), tuple(zip(*stats)), counts))
        s.close()
    _frames_match(results, "groom frames 1-2")


def test_groom_contact_state_stages_match_oracle(oracle_lib, gpu_lib, groom_subset, groom_contact_checkpoint):
    """The same groom once its hair lies on the body (the GPU steps frames
    1-12 alone, and both libraries restore its checkpoint): at the first
    substep of frame 13, with the body between its frame-12 and substep
    positions, the sweep, the evaluation (energies, gradient, Gauss-Newton
    blocks, candidate and active counts, minimum gap), the direction, its
    step bound and the friction impulse are compared stage by stage. Whole
    contact frames are not compared: in them the reference's Newton loop
    stops at its iteration cap in every substep (24), short of the residual,
    so the iterate depends on the last bits of every operation; the host and
    device libm (log, sin, cos) and the oracle's dual-number Jacobian against
    the device's analytic one differ there."""
    scene = groom_subset
    state = groom_contact_checkpoint
    substeps = scene.desc["substeps"]
    dt = scene.frame_dt / substeps
    prev = scene.frame_vertices(GROOM_WARM)
    cur = prev * (1.0 - 1.0 / substeps) + scene.frame_vertices(GROOM_WARM + 1) * (1.0 / substeps)
    out = []
    for lib in (oracle_lib, gpu_lib):
        print(f"[gpu] groom contact stages on {'oracle' if lib is oracle_lib else 'gpu'}", flush=True)
        t = _groom_solver(lib, scene, GROOM_WARM + 2)
        t.check(t.restore_checkpoint(state))
        q, v = t.get_state()
        t.check(t.set_state(q, v, q))
        t.check(t.set_collider_state(cur, prev))
        ok, diag = t.sweep_check()
        ev, grad, d = t.evaluate(dt)
        D, U = t.blocks()
        out.append(dict(solver=t, q=q, v=v, ok=ok, diag=diag, ev=ev, g=grad, d=d, D=D, U=U))
    o, g = out
    eo, eg = o["ev"], g["ev"]
    print(f"sweep {o['ok']} {g['ok']}; active {eo.active_contacts} {eg.active_contacts}; candidates "
          f"{eo.candidate_pairs} {eg.candidate_pairs}; min gap {eo.minimum_gap!r} {eg.minimum_gap!r}")
    assert g["ok"] == o["ok"]
    assert eo.active_contacts > 20 and eg.active_contacts == eo.active_contacts
    assert eg.candidate_pairs == eo.candidate_pairs
    assert eg.feasible == eo.feasible == 1
    assert eg.minimum_gap == eo.minimum_gap
    for f in ("objective", "elastic", "contact", "kinetic"):
        assert abs(getattr(eg, f) - getattr(eo, f)) <= 1e-10 * abs(getattr(eo, f)), f
    assert _rel(g["g"], o["g"]) <= 1e-10
    assert max(_rel(g["D"], o["D"]), _rel(g["U"], o["U"])) <= 1e-10
    from test_gpu_parity import _newton_residual
    res_o = _newton_residual(o["D"], o["U"], o["g"], o["d"], o["solver"])
    res_g = _newton_residual(o["D"], o["U"], o["g"], g["d"], o["solver"])
    print(f"direction backward error oracle {res_o:.3e} gpu {res_g:.3e}")
    assert res_g <= max(10 * res_o, 1e-10)
    alphas = [x["solver"].step_bound(o["d"]) for x in out]
    print(f"step bound oracle {alphas[0]!r} gpu {alphas[1]!r}")
    assert abs(alphas[1] - alphas[0]) <= 1e-12
    dv = []
    for x in out:
        x["solver"].check(x["solver"].set_state(x["q"], x["v"], x["q"]))
        x["solver"].check(x["solver"].set_collider_state(cur, prev))
        x["solver"].check(x["solver"].apply_friction(dt))
        dv.append(x["solver"].get_state()[1] - x["v"])
        x["solver"].close()
    print(f"friction: dofs changed {np.count_nonzero(dv[0])} {np.count_nonzero(dv[1])}; rel {_rel(dv[1], dv[0]):.3e}")
    assert np.count_nonzero(dv[0]) > 20
    assert np.count_nonzero(dv[1]) == np.count_nonzero(dv[0])
    assert _rel(dv[1], dv[0]) <= 1e-10


def test_contact_reuse_is_exact(gpu_lib, groom_subset, groom_contact_checkpoint):
    """The device's reuse of contact-list work (the gradient at an accepted
    trial takes the trial's closest triangles; a trial inside the boxes the
    swept pass covered runs no check) changes nothing: from the groom
    subset's frame-12 contact state, frames 13-14 stepped with the reuse on
    and off give the same positions bit for bit and every count equal."""
    scene = groom_subset
    results = []
    for mode in (1, 0):
        print(f"[gpu] groom contact frames 13-14, contact reuse {mode}", flush=True)
        s = _groom_solver(gpu_lib, scene, GROOM_WARM + 3)
        s.check(s.set_contact_reuse(mode))
        s.check(s.restore_checkpoint(groom_contact_checkpoint))
        stats = []
        for f in (GROOM_WARM + 1, GROOM_WARM + 2):
            s.check(s.step_frame(f, scene.frame_dt))
            st = s.step_stats()
            stats.append((st.newton_iterations, st.line_search_evaluations, st.contact_candidate_count,
                          st.active_contact_count, st.minimum_gap))
        results.append((s.internal_positions(), stats))
        s.close()
    (x_on, st_on), (x_off, st_off) = results
    print(f"reuse on {st_on}\nreuse off {st_off}")
    assert st_on[0][3] > 20  # active barriers: the frames are in contact
    assert st_on == st_off
    assert np.array_equal(x_on, x_off)


def test_fp32_block_solve_matches_oracle(oracle_lib, gpu_lib):
    """Precision mode 1: the FP32 block solve from the same FP64 blocks and
    gradient gives the oracle's FP32 direction (the same single-precision
    operations in the same order), and a frame in that mode stalls at the
    Newton cap on the GPU as on the oracle."""
    print("[gpu] FP32 block solve against the oracle's", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    q, rng = _pulled_in(o, seed=12, lo=2e-4, hi=1.6e-3)
    v = rng.uniform(-1e-3, 1e-3, q.size)
    out = []
    for s in (o, g):
        s.check(s.set_state(q, v, q - 1e-4 * v))
        s.check(s.set_precision_mode(1))
        out.append(s.evaluate(scene.frame_dt / 4))
    (eo, go, do), (eg, gg, dg) = out
    same = float(np.mean(dg == do))
    print(f"FP32 direction: relative difference {_rel(dg, do):.3e}, bit-equal entries {same:.4f}")
    assert eo.solve_failed == eg.solve_failed == 0
    assert _rel(gg, go) <= 1e-10
    assert _rel(dg, do) <= 1e-5
    stats = []
    for lib in (oracle_lib, gpu_lib):
        s = grooms.build_solver(lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.check(s.set_precision_mode(1))
        s.set_threads(16)
        s.check(s.step(scene.frame_dt))
        st = s.step_stats()
        stats.append((st.newton_iterations, st.relative_residual_norm, st.phase))
        s.close()
    print(f"FP32 frame: oracle {stats[0]} gpu {stats[1]}")
    assert stats[1][2] == kn.PHASE_FINISHED
    assert stats[1][0] == stats[0][0] == 4 * 24


def _config3_materials(strands, seed=3):
    """Per-strand Young's modulus U(1, 5) GPa and radius U(30, 80) um."""
    rng = np.random.default_rng(seed)
    return rng.uniform(1.0e9, 5.0e9, strands), rng.uniform(3.0e-5, 8.0e-5, strands)


def test_strand_materials_evaluation_matches_oracle(oracle_lib, gpu_lib):
    """Per-strand E and radius at a state with many active barriers: the
    radius enters the contact gap, the query padding and the masses, E the
    stiffness; energies, gradient, blocks and counts against the oracle."""
    print("[gpu] per-strand materials, evaluation with active barriers", flush=True)
    scene, o, g = _sphere_pair(oracle_lib, gpu_lib)
    E, r = _config3_materials(scene.strand_count)
    for s in (o, g):
        s.check(s.set_strand_materials(E, r))
    _, mo, ro, _ = o.node_model()
    _, mg, rg, _ = g.node_model()
    assert np.max(np.abs(mo - mg) / mo) <= 1e-15 and np.max(np.abs(ro - rg) / ro) <= 1e-15
    q, rng = _pulled_in(o, seed=13)
    v = rng.uniform(-1e-3, 1e-3, q.size)
    res = []
    for s in (o, g):
        s.check(s.set_state(q, v, q - 1e-4 * v))
        ev, grad, d = s.evaluate(scene.frame_dt / 4)
        D, U = s.blocks()
        res.append((ev, grad, D, U))
    (eo, go, Do, Uo), (eg, gg, Dg, Ug) = res
    print(f"active {eo.active_contacts} {eg.active_contacts}; candidates {eo.candidate_pairs} {eg.candidate_pairs}")
    assert eo.active_contacts > 50 and eg.active_contacts == eo.active_contacts
    assert eg.candidate_pairs == eo.candidate_pairs and eg.minimum_gap == eo.minimum_gap
    for f in ("objective", "elastic", "contact", "kinetic"):
        assert abs(getattr(eg, f) - getattr(eo, f)) <= 1e-10 * abs(getattr(eo, f)), f
    assert _rel(gg, go) <= 1e-10
    assert max(_rel(Dg, Do), _rel(Ug, Uo)) <= 1e-10


def test_strand_materials_frame_matches_oracle(oracle_lib, gpu_lib):
    """One frame of the sphere scene with the uneven-material groom's per-strand
    E and radius; and the same values for every strand reproduce the
    solver-wide material bit for bit on the GPU."""
    print("[gpu] per-strand materials, one frame", flush=True)
    scene = grooms.small_sphere_scene()
    E, r = _config3_materials(scene.strand_count)
    results = []
    for lib in (oracle_lib, gpu_lib):
        s = grooms.build_solver(lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.set_threads(16)
        s.check(s.set_strand_materials(E, r))
        s.check(s.step(scene.frame_dt))
        st = s.step_stats()
        results.append((s.internal_positions(), ((st.newton_iterations,), (st.line_search_evaluations,),
                                                 (st.contact_candidate_count,), (st.active_contact_count,)),
                        [s.direction_counts()]))
        s.close()
    _frames_match(results, "per-strand materials")
    plain = []
    for uniform in (False, True):
        s = grooms.build_solver(gpu_lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        if uniform:
            n = scene.strand_count
            s.check(s.set_strand_materials(np.full(n, scene.material["young_modulus"]),
                                           np.full(n, scene.material["radius"])))
        s.check(s.step(scene.frame_dt))
        plain.append(s.internal_positions())
        s.close()
    assert np.array_equal(plain[0], plain[1])
