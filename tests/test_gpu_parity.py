"""Parity of the HIP step with the CPU oracle on the same seeded inputs.

Bars: rest model bit-for-bit (same FP64 operations, contraction off);
energy, gradient and Gauss-Newton direction within 1e-10 relative
(norm-wise); broad-phase candidate pairs bit-exact; positions after one frame
within 1e-7 m; minimum gap above the feasibility threshold.
"""
import math
import os

import numpy as np
import pytest

import grooms
import khs_native as kn

pytestmark = pytest.mark.gpu

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)


def _pair(oracle_lib, gpu_lib, scene, **desc):
    o = grooms.build_solver(oracle_lib, scene, **desc)
    g = grooms.build_solver(gpu_lib, scene, **desc)
    return o, g


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _perturbed_state(solver, seed, dx=1e-4, du=0.05, dv=1e-3):
    q, v = solver.get_state()
    rng = np.random.default_rng(seed)
    n = q.size // 6
    dq = np.zeros((n, 6))
    dq[:, :3] = rng.uniform(-dx, dx, (n, 3))
    dq[:, 3:] = rng.uniform(-du, du, (n, 3))
    vel = rng.uniform(-dv, dv, (n, 6))
    return q, q + dq.ravel(), vel.ravel()


def test_rest_model_matches_oracle(oracle_lib, gpu_lib):
    print("[gpu] rest model, small sphere scene", flush=True)
    scene = grooms.small_sphere_scene()
    o, g = _pair(oracle_lib, gpu_lib, scene)
    # the same FP64 operations in the same order, contraction off: bit for bit
    fo, mo, ro, flo = o.node_model()
    fg, mg, rg, flg = g.node_model()
    assert np.array_equal(flo, flg)
    assert np.array_equal(fo, fg)
    assert np.array_equal(mo, mg)
    assert np.array_equal(ro, rg)
    no, so, co, resto = o.element_model()
    ng, sg, cg, restg = g.element_model()
    assert np.array_equal(no, ng) and np.array_equal(so, sg) and np.array_equal(co, cg)
    assert np.array_equal(resto, restg)
    bo, bg = o.build_stats(), g.build_stats()
    for f in ("internal_node_count", "element_count", "fixed_node_count", "degree_of_freedom_count",
              "collider_triangle_count", "collider_boundary_edge_count", "collider_inverted_closed_component_count"):
        assert getattr(bo, f) == getattr(bg, f), f


@pytest.mark.parametrize("seed", [0, 1])
def test_energy_gradient_direction_match_oracle(oracle_lib, gpu_lib, seed):
    print(f"[gpu] evaluate parity seed {seed}", flush=True)
    scene = grooms.small_sphere_scene()
    o, g = _pair(oracle_lib, gpu_lib, scene)
    q_old, q, v = _perturbed_state(o, seed)
    for s in (o, g):
        s.check(s.set_state(q, v, q_old))
    dt = scene.frame_dt / 4
    eo, go, do = o.evaluate(dt)
    eg, gg, dg = g.evaluate(dt)
    energy_rel = {f: abs(getattr(eg, f) - getattr(eo, f)) / abs(getattr(eo, f)) for f in ("objective", "elastic", "kinetic")}
    rel_g = _rel(gg, go)
    rel_d = _rel(dg, do)
    print(f"energy rel {energy_rel} gradient rel {rel_g:.3e} direction rel {rel_d:.3e}")
    for f, e in energy_rel.items():
        assert e <= 1e-10, (f, e)
    assert eg.feasible == eo.feasible == 1
    assert eg.solve_failed == eo.solve_failed == 0
    assert rel_g <= 1e-10
    # The per-strand Newton systems are ill-conditioned (stretch stiffness
    # EA/L ~ 1e3 against twist GJ/L ~ 1e-7), so the direction is compared by
    # its backward error in the oracle's own system: the GPU direction must
    # solve H_oracle d = -g as well as the oracle's direction does.
    Do, Uo = o.blocks()
    Dg, Ug = g.blocks()
    rel_h = max(_rel(Dg, Do), _rel(Ug, Uo))
    res_o = _newton_residual(Do, Uo, go, do, o)
    res_g = _newton_residual(Do, Uo, go, dg, o)
    print(f"blocks rel {rel_h:.3e} backward error oracle {res_o:.3e} gpu {res_g:.3e}")
    assert rel_h <= 1e-10
    assert res_g <= max(10 * res_o, 1e-10)


def _newton_residual(D, U, g, d, solver):
    """||H d + g|| / ||g|| over free nodes, H block tridiagonal per strand."""
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


def _draped_scene():
    """Small-sphere strands pulled radially onto the collider so that many
    elements sit inside the barrier padding."""
    scene = grooms.small_sphere_scene()
    return scene


def test_broadphase_pairs_bit_exact(oracle_lib, gpu_lib):
    print("[gpu] broad phase pairs", flush=True)
    scene = grooms.small_sphere_scene()
    o, g = _pair(oracle_lib, gpu_lib, scene)
    q, v = o.get_state()
    pts = q.reshape(-1, 6)[:, :3]
    r = np.linalg.norm(pts, axis=1, keepdims=True)
    # shrink the radial distance of free nodes so segments come within the padding
    pts2 = pts * np.minimum(1.0, (0.098 + 0.00025 + (r - 0.1) * 0.002) / r)
    q2 = q.reshape(-1, 6).copy()
    q2[:, :3] = pts2
    q2 = q2.ravel()
    for s in (o, g):
        s.check(s.set_state(q2, v, q2))
    check = g.check_broadphase()
    print(f"broad-phase self-check {check}")
    assert check[:5] == [0, 0, 0, 0, 0]
    po = o.broadphase_pairs()
    pg = g.broadphase_pairs()
    print(f"pairs oracle {len(po)} gpu {len(pg)}")
    assert len(po) > 100
    assert np.array_equal(po, pg)


def test_one_frame_positions_match_oracle(oracle_lib, gpu_lib):
    print("[gpu] one frame, small sphere scene, floors lifted", flush=True)
    scene = grooms.small_sphere_scene()
    o, g = _pair(oracle_lib, gpu_lib, scene, **LIFTED)
    for s in (o, g):
        s.check(s.set_convergence_mode(1))
        s.check(s.step(scene.frame_dt))
    xo, xg = o.internal_positions(), g.internal_positions()
    err = float(np.max(np.abs(xo - xg)))
    so, sg = o.step_stats(), g.step_stats()
    print(f"max position error {err:.3e} m; newton oracle {so.newton_iterations} gpu {sg.newton_iterations}; "
          f"candidates oracle {so.contact_candidate_count} gpu {sg.contact_candidate_count}; "
          f"active oracle {so.active_contact_count} gpu {sg.active_contact_count}")
    assert err <= 1e-7
    assert sg.phase == kn.PHASE_FINISHED
    # The candidate count sums the triangles every evaluation tested, so it
    # also checks that per-element candidate lists report the full set.
    assert sg.contact_candidate_count == so.contact_candidate_count
    assert sg.active_contact_count == so.active_contact_count
    assert sg.minimum_gap > scene.desc.get("minimum_gap", 1e-7) or np.isinf(sg.minimum_gap)


def test_contact_frames_match_oracle(oracle_lib, gpu_lib):
    """The reference's barrier scenario (tests/test_solver.cpp:218-251), three
    frames, compared node by node."""
    print("[gpu] barrier contact frames", flush=True)
    pts = np.array([[0, 0, 0.03], [0, 0, 0.02], [0, 0, 0.01], [0, 0, 0.001]])
    col = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]])
    tris = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    out = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.substeps = 4
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.02
        desc.newton_iterations = 40
        mat = kn.default_material(lib)
        mat.radius = 2.0e-4
        mat.young_modulus = 1.0e6
        mat.contact_stiffness = 100.0
        mat.barrier_distance = 2.0e-3
        mat.mass_damping = 2.0
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, [0, 4], mat))
        s.check(s.set_collider(col, tris))
        s.check(s.build())
        for _ in range(3):
            s.check(s.step(1.0 / 60.0))
        out.append((s.internal_positions(), s.step_stats()))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    print(f"max position error {err:.3e} m; active oracle {out[0][1].active_contact_count} gpu {out[1][1].active_contact_count}")
    assert err <= 1e-7
    assert out[1][1].active_contact_count > 0
    assert out[1][1].minimum_gap > 1e-7


def _grid_plane(n, h):
    """n x n quads of side h centred on the origin in z = 0, two triangles each."""
    xs = (np.arange(n + 1) - n / 2) * h
    gx, gy = np.meshgrid(xs, xs, indexing="ij")
    verts = np.stack([gx.ravel(), gy.ravel(), np.zeros(gx.size)], axis=1)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            tris += [[a, b, c], [a, c, d]]
    return verts, np.array(tris, np.uint32)


@pytest.mark.parametrize("quad", [5.0e-3, 1.0e-3, 4.0e-4])
def test_draped_strand_on_grid_matches_oracle(oracle_lib, gpu_lib, quad):
    """A strand falls onto a tessellated plane. With 5 mm quads the lists are
    short; with 1 mm quads an element's skin holds several hundred triangles
    (long lists, thin-skin regathers); with 0.4 mm quads even the thin box
    holds more than a list can, so those elements take the hierarchy path."""
    print(f"[gpu] draped strand on grid, quad {quad}", flush=True)
    verts, tris = _grid_plane(int(round(0.08 / quad)), quad)
    pts = np.array([[-0.03 + 0.01 * k, 0.0013 * k, 0.006] for k in range(7)])
    out = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.substeps = 4
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.01
        desc.newton_iterations = 60
        desc.relative_tolerance = 1e-6
        desc.absolute_tolerance = 1e-30
        desc.increment_tolerance = 1e-14
        mat = kn.default_material(lib)
        mat.radius = 1.0e-4
        mat.young_modulus = 1.0e7
        mat.contact_stiffness = 100.0
        mat.barrier_distance = 2.0e-3
        mat.mass_damping = 2.0
        mat.friction = 0.3
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, [0, len(pts)], mat))
        s.check(s.set_collider(verts, tris))
        s.check(s.build())
        s.check(s.set_convergence_mode(1))
        cand, active = [], []
        for _ in range(4 if quad > 5.0e-4 else 2):  # the hierarchy path is slow: two frames
            s.check(s.step(1.0 / 30.0))
            st = s.step_stats()
            cand.append(st.contact_candidate_count)
            active.append(st.active_contact_count)
        lists = s.contact_list_stats() if lib is gpu_lib else None
        out.append((s.internal_positions(), cand, active, s.step_stats(), lists))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    lists = out[1][4]
    print(f"max position error {err:.3e} m; candidates oracle {out[0][1]} gpu {out[1][1]}; "
          f"active oracle {out[0][2]} gpu {out[1][2]}; longest list {lists.longest}, overflowed {lists.overflowed}")
    assert err <= 1e-7
    assert out[1][1] == out[0][1]
    assert out[1][2] == out[0][2]
    assert sum(out[1][2]) > 0
    assert out[1][3].minimum_gap > 1e-7
    if quad < 5.0e-4:
        assert lists.overflowed > 0


def test_trust_radius_frames_match_oracle(oracle_lib, gpu_lib):
    """A long frame with one substep: the free fall g dt^2 / 2 is far above
    the trust radius, so every Newton direction is scaled down to it (and
    the steepest-descent fallback may be taken). Positions and iteration
    counts must follow the oracle."""
    print("[gpu] trust-radius limited frames", flush=True)
    pts = np.array([[0.01 * k, 0.0, 0.0] for k in range(8)])
    out = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.substeps = 1
        desc.maximum_substeps = 1
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.02
        desc.newton_iterations = 12
        mat = kn.default_material(lib)
        mat.young_modulus = 1.0e7
        mat.radius = 1.0e-4
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, [0, len(pts)], mat))
        s.check(s.build())
        s.check(s.set_convergence_mode(1))
        stats = []
        for _ in range(2):
            s.check(s.step(0.25))
            st = s.step_stats()
            stats.append((st.newton_iterations, st.line_search_evaluations, st.linear_solves))
        out.append((s.internal_positions(), stats))
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    print(f"max position error {err:.3e} m; oracle {out[0][1]} gpu {out[1][1]}")
    assert out[1][1] == out[0][1]
    assert err <= 1e-7
    moved = float(np.max(np.abs(out[1][0] - pts)))
    assert moved > 0.01  # the steps really were trust-limited


def test_concurrent_solvers_match_alone(gpu_lib):
    """Independent solvers stepped at once from several host threads (each
    on its own stream) give bit-for-bit the positions of a solver run
    alone: the batched form shares nothing between jobs."""
    import threading
    print("[gpu] concurrent solvers against one alone", flush=True)
    scene = grooms.small_sphere_scene()

    def run(s):
        s.check(s.set_convergence_mode(1))
        for _ in range(2):
            s.check(s.step(scene.frame_dt))
        return s.internal_positions()

    alone = run(grooms.build_solver(gpu_lib, scene, **LIFTED))
    solvers = [grooms.build_solver(gpu_lib, scene, **LIFTED) for _ in range(3)]
    results = [None] * len(solvers)

    def work(k):
        results[k] = run(solvers[k])

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(solvers))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for r in results:
        assert r is not None and np.array_equal(r, alone)


def _axis_rotation(axis, angle):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


@pytest.mark.parametrize("degrees", [35.0, 130.0])
def test_device_root_targets_match_oracle(oracle_lib, gpu_lib, degrees):
    """Root targets of uploaded frames are computed on the device (binding
    lerp, transported frames, rotation vectors of the fixed nodes). After one
    frame the fixed nodes sit exactly on the last frame's targets, so their
    state compares the device targets with the oracle's host ones (whose
    rigid-rotation closed forms test_oracle_pins checks); two fixed nodes per
    strand exercise the transport between them, and 130 degrees about a
    tilted axis the half-turn branch of the rotation vector."""
    print(f"[gpu] device root targets, rotation {degrees} deg", flush=True)
    scene = grooms.small_sphere_scene()
    phi = np.radians(degrees)
    R = _axis_rotation([0.0, 0.0, 1.0], phi) if degrees < 90 else _axis_rotation([0.3, -0.2, 1.0], phi)
    states = []
    for lib in (oracle_lib, gpu_lib):
        s = grooms.build_solver(lib, scene, collider=False, fixed_root_nodes=2)
        s.check(s.allocate_animation(2))
        s.check(s.set_root_frame(0, scene.points))
        s.check(s.set_root_frame(1, scene.points @ R.T))
        s.check(s.finalize_animation())
        s.check(s.step_frame(1, scene.frame_dt))
        q, _ = s.get_state()
        _, _, _, flags = s.node_model()
        states.append((q.reshape(-1, 6), flags))
    (qo, fo), (qg, fg) = states
    assert np.array_equal(fo, fg)
    fixed = (fo & 1) == 1
    assert fixed.sum() == 2 * 64
    err = np.max(np.abs(qg[fixed] - qo[fixed]))
    print(f"fixed-node target difference {err:.3e}")
    assert err <= 1e-13


def _curly_strands(rng, strands):
    """Curly strands of 3..40 points with repeated points (zero-length
    segments) scattered in, and some short strands for the extension."""
    points, offsets = [], [0]
    for s in range(strands):
        count = int(rng.integers(3, 41))
        root = rng.uniform(-0.05, 0.05, 3)
        step = rng.uniform(0.002, 0.006)
        angle = np.cumsum(rng.uniform(0.0, 0.4, count))
        p = root + np.stack([0.004 * np.cos(angle), -step * np.arange(count), 0.004 * np.sin(angle)], axis=1)
        if count > 4 and s % 3 == 0:
            k = int(rng.integers(1, count - 1))
            p = np.insert(p, k, p[k], axis=0)  # a zero-length segment
        points.append(p)
        offsets.append(offsets[-1] + len(p))
    return np.concatenate(points), offsets


def test_device_model_build_matches_oracle(oracle_lib, gpu_lib):
    """The rod model built on the device (discretization, merges, virtual
    extension, frames, rest strains, masses, mapping) against the oracle's
    host build, and the masses again after a material change."""
    print("[gpu] device model build against the oracle", flush=True)
    rng = np.random.default_rng(7)
    points, offsets = _curly_strands(rng, 300)
    solvers = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.maximum_element_length = 0.003
        desc.minimum_dynamic_length = 0.08
        desc.fixed_root_nodes = 2
        s = kn.HairSolver(lib, desc)
        mat = kn.default_material(lib)
        s.check(s.set_hair(points, offsets, mat))
        s.check(s.build())
        solvers.append((s, desc, mat))
    (o, _, _), (g, _, _) = solvers
    bo, bg = o.build_stats(), g.build_stats()
    for f in ("internal_node_count", "element_count", "fixed_node_count", "degree_of_freedom_count",
              "merged_zero_length_segment_count", "virtual_extension_strand_count", "virtual_extension_node_count"):
        assert getattr(bo, f) == getattr(bg, f), f
    assert bo.merged_zero_length_segment_count > 0 and bo.virtual_extension_strand_count > 0
    assert abs(bo.virtual_extension_rest_length - bg.virtual_extension_rest_length) <= 1e-15 * bo.virtual_extension_rest_length
    assert np.array_equal(o.mapping(), g.mapping())
    # positions, frames, rest strains and masses bit for bit
    assert np.array_equal(o.internal_positions(), g.internal_positions())
    fo, mo, ro, flo = o.node_model()
    fg, mg, rg, flg = g.node_model()
    assert np.array_equal(flo, flg)
    assert np.array_equal(fo, fg)
    assert np.array_equal(mo, mg)
    assert np.array_equal(ro, rg)
    no, so, co, resto = o.element_model()
    ng, sg, cg, restg = g.element_model()
    assert np.array_equal(no, ng) and np.array_equal(so, sg) and np.array_equal(co, cg)
    assert np.array_equal(resto, restg)
    for s, desc, mat in solvers:
        mat.density *= 1.5
        mat.radius *= 0.75
        s.check(s.update_runtime(desc, mat))
    _, mo, ro, _ = o.node_model()
    _, mg, rg, _ = g.node_model()
    assert np.array_equal(mo, mg)
    assert np.array_equal(ro, rg)
    # a strand with every point at one place is refused, naming the strand
    bad = np.concatenate([points[:offsets[2]], np.repeat(points[:1], 3, axis=0)])
    for lib in (oracle_lib, gpu_lib):
        s = kn.HairSolver(lib, kn.default_desc(lib))
        s.check(s.set_hair(bad, [0, offsets[1], offsets[2], offsets[2] + 3], kn.default_material(lib)))
        assert s.build() == kn.KHS_ERROR_INVALID_HAIR
        assert "Strand 2" in s.last_error()
        s.close()
    for s, _, _ in solvers:
        s.close()


def test_dense_pole_contact_matches_oracle(oracle_lib, gpu_lib):
    """Hair lying across the pole of a UV sphere, where 160 sliver triangles
    meet: candidate lists of more than 128 triangles, gathered again with
    the thin skin where the skin holds more than a list can. Two frames
    compared node by node with the oracle's exhaustive search."""
    print("[gpu] dense pole contact frames", flush=True)
    R = 0.05
    verts, tris = grooms.uv_ellipsoid(24, 160, (R, R, R), (0.0, 0.0, 0.0))
    strands, offsets = [], [0]
    for k in range(8):
        phi = math.pi * k / 8
        u = np.array([math.cos(phi), math.sin(phi), 0.0])
        s = np.linspace(-0.012, 0.012, 9)[:, None] * u + np.array([0.0, 0.0, R + 0.0004])
        strands.append(s)
        offsets.append(offsets[-1] + len(s))
    pts = np.concatenate(strands)
    out = []
    for lib in (oracle_lib, gpu_lib):
        desc = kn.default_desc(lib)
        desc.substeps = 4
        desc.fixed_root_nodes = 1
        desc.maximum_element_length = 0.01
        desc.newton_iterations = 40
        mat = kn.default_material(lib)
        mat.radius = 2.0e-4
        mat.young_modulus = 1.0e6
        mat.contact_stiffness = 100.0
        mat.barrier_distance = 2.0e-3
        mat.mass_damping = 2.0
        s = kn.HairSolver(lib, desc)
        s.check(s.set_hair(pts, offsets, mat))
        s.check(s.set_collider(verts, tris))
        s.check(s.build())
        s.check(s.step(1.0 / 60.0))
        lists = s.contact_list_stats() if lib is gpu_lib else None
        first = s.step_stats()
        s.check(s.step(1.0 / 60.0))
        out.append((s.internal_positions(), first, lists))
        s.close()
    err = float(np.max(np.abs(out[0][0] - out[1][0])))
    lists = out[1][2]
    print(f"max position error {err:.3e} m; active oracle {out[0][1].active_contact_count} "
          f"gpu {out[1][1].active_contact_count}; longest list {lists.longest}, overflowed {lists.overflowed}")
    assert err <= 1e-7
    assert out[1][1].active_contact_count > 0
    assert out[1][1].minimum_gap > 1e-7
    assert lists.longest > 128


def test_sphere_scene_matches_golden(gpu_lib):
    """The GPU against the oracle's committed results on the 64-strand sphere
    scene (tests/golden/config1_frame.npz), without running the oracle:
    positions after one frame within 1e-7 m with the same Newton count, and
    objective and gradient at the perturbed state within 1e-10 relative."""
    import golden.make_golden as make_golden

    print("[gpu] sphere scene against the committed oracle results", flush=True)
    want = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "config1_frame.npz"))
    got = make_golden.compute(gpu_lib)
    err = float(np.max(np.abs(got["positions"] - want["positions"])))
    print(f"max position error {err:.3e} m; newton {got['newton'][0]} (golden {want['newton'][0]})")
    assert err <= 1e-7
    assert got["newton"][0] == want["newton"][0]
    assert abs(got["objective"][0] - want["objective"][0]) <= 1e-10 * abs(want["objective"][0])
    assert _rel(got["gradient"], want["gradient"]) <= 1e-10


def test_gpu_stats_phase_times(gpu_lib):
    """With kernel timing on, khsGetGpuStats reports the frame's assembly,
    collision and optimization (block solve) milliseconds from the
    per-launch HIP events; they are the frame's share of the kernel-class
    totals and fit inside the frame."""
    print("[gpu] phase times in the GPU stats", flush=True)
    scene = grooms.small_sphere_scene()
    s = grooms.build_solver(gpu_lib, scene, **LIFTED)
    s.check(s.set_convergence_mode(1))
    s.check(s.set_kernel_timing(True))
    s.check(s.step(scene.frame_dt))
    st, t = s.gpu_stats(), s.kernel_timing()
    print(f"frame {st.last_frame_milliseconds:.3f} ms: assembly {st.last_assembly_milliseconds:.3f}, collision "
          f"{st.last_collision_milliseconds:.3f}, optimization {st.last_optimization_milliseconds:.3f}")
    assert st.last_assembly_milliseconds > 0 and st.last_collision_milliseconds > 0
    assert st.last_optimization_milliseconds > 0
    assert abs(st.last_assembly_milliseconds - (t.assembly_ms + t.energy_ms)) <= 1e-9
    assert abs(st.last_optimization_milliseconds - t.solve_ms) <= 1e-9
    assert (st.last_assembly_milliseconds + st.last_collision_milliseconds + st.last_optimization_milliseconds
            <= st.last_frame_milliseconds * 1.001)
    s.check(s.set_kernel_timing(False))
    s.check(s.step(scene.frame_dt))
    assert s.gpu_stats().last_optimization_milliseconds == 0.0
    s.close()
