"""Per-strand Young's modulus and radius (khsxSetStrandMaterials) on the
oracle: the same values for every strand reproduce the solver-wide material
bit for bit; different values enter masses, stiffness and the contact
radius as the closed forms say; bad input is refused."""
import math

import numpy as np
import pytest

import grooms
import khs_native as kn

LIFTED = dict(relative_tolerance=1e-6, absolute_tolerance=1e-30, increment_tolerance=1e-14)


def _two_strands(lib, z=0.1, collider=None, E=3.0e6, radius=1.0e-4):
    pts = np.array([[0.01 * k, 0.0, z] for k in range(6)] + [[0.01 * k, 0.05, z] for k in range(6)])
    desc = kn.default_desc(lib)
    desc.gravity = kn.KhsVec3(0.0, 0.0, 0.0)
    desc.fixed_root_nodes = 1
    desc.maximum_element_length = 0.05
    mat = kn.default_material(lib)
    mat.young_modulus = E
    mat.radius = radius
    mat.mass_damping = 0.0
    mat.barrier_distance = 1.0e-3
    mat.contact_stiffness = 10.0
    s = kn.HairSolver(lib, desc)
    s.check(s.set_hair(pts, [0, 6, 12], mat))
    if collider is not None:
        s.check(s.set_collider(*collider))
    s.check(s.build())
    return s, mat


def test_uniform_values_reproduce_the_material(oracle_lib):
    scene = grooms.small_sphere_scene()
    out = []
    for per_strand in (False, True):
        s = grooms.build_solver(oracle_lib, scene, **LIFTED)
        s.check(s.set_convergence_mode(1))
        s.set_threads(4)
        if per_strand:
            n = scene.strand_count
            s.check(s.set_strand_materials(np.full(n, scene.material["young_modulus"]),
                                           np.full(n, scene.material["radius"])))
        s.check(s.step(scene.frame_dt))
        out.append((s.internal_positions(), s.step_stats().newton_iterations))
        s.close()
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_stiffness_and_mass_follow_each_strand(oracle_lib):
    s, mat = _two_strands(oracle_lib)
    E = np.array([2.0e6, 5.0e6])
    r = np.array([8.0e-5, 1.3e-4])
    s.check(s.set_strand_materials(E, r))
    _, mass, rmass, _ = s.node_model()
    L0 = 0.01
    for k in range(2):
        m_elem = mat.density * math.pi * r[k] ** 2 * L0
        want = m_elem * np.array([0.5, 1, 1, 1, 1, 0.5])
        assert np.allclose(mass[6 * k:6 * k + 6], want, rtol=1e-14, atol=0)
    q, _ = s.get_state()
    x = q.reshape(-1, 6)
    lam = 1.002
    for k in range(2):
        root = x[6 * k, :3].copy()
        x[6 * k:6 * k + 6, :3] = root + lam * (x[6 * k:6 * k + 6, :3] - root)
    s.check(s.set_state(q, np.zeros_like(q), q))
    ev, _, _ = s.evaluate(1e-3, direction=False)
    want = sum(5 * 0.5 * L0 * E[k] * math.pi * r[k] ** 2 * (lam - 1.0) ** 2 for k in range(2))
    assert abs(ev.elastic - want) <= 1e-12 * want


def test_contact_radius_follows_each_strand(oracle_lib):
    h = 6.0e-4
    plane = (np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [1.0, 1.0, 0.0], [-1.0, 1.0, 0.0]]),
             np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    s, mat = _two_strands(oracle_lib, z=h, collider=plane)
    r = np.array([5.0e-5, 2.0e-4])
    s.check(s.set_strand_materials(np.full(2, mat.young_modulus), r))
    q, _ = s.get_state()
    s.check(s.set_state(q, np.zeros_like(q), q))
    ev, _, _ = s.evaluate(1e-3, direction=False)
    b = lambda g: -mat.contact_stiffness * (mat.barrier_distance - g) ** 2 * math.log(g / mat.barrier_distance)
    # 4 free-left elements per strand (the root element takes no barrier)
    want = sum(4 * b(h - r[k] - mat.collider_offset) for k in range(2))
    assert abs(ev.contact - want) <= 1e-12 * want
    assert abs(ev.minimum_gap - (h - r[1])) <= 1e-18
    assert ev.active_contacts == 8


def test_bad_strand_materials_are_refused(oracle_lib):
    s, _ = _two_strands(oracle_lib)
    assert s.set_strand_materials(np.ones(3), np.ones(3) * 1e-4) == kn.KHS_ERROR_INVALID_ARGUMENT
    assert s.set_strand_materials(np.array([1e6, -1.0]), np.ones(2) * 1e-4) == kn.KHS_ERROR_INVALID_ARGUMENT
    assert s.set_strand_materials(np.array([1e6, 1e6]), np.array([1e-4, np.nan])) == kn.KHS_ERROR_INVALID_ARGUMENT
    assert s.set_strand_materials() == kn.KHS_OK
