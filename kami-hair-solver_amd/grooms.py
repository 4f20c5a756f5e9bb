"""Synthetic grooms and colliders for tests and benchmarks.

Every random draw goes through SplitMix64, so this container and the GPU
machine build bit-identical scenes from the same seed.

small_sphere_scene: 64 strands x 32 points on a rigid sphere of radius 0.10 m
    with an icosphere collider (subdivision 3, 1,280 triangles) 2 mm inside.
head_scene: 6,757 strands, lognormal node counts (median 36, clipped 8..80),
    on the upper part of a deformed ellipsoid head, with a head + torso body
    of 225,184 vertices rotating 30 deg/s about the vertical axis.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; uniform doubles take the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def normal(self) -> float:
        # Box-Muller; u1 is kept away from zero
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def unit_vector(self) -> np.ndarray:
        while True:
            v = np.array([self.normal(), self.normal(), self.normal()])
            n = np.linalg.norm(v)
            if n > 1e-9:
                return v / n


@dataclass
class Scene:
    points: np.ndarray            # (P, 3)
    offsets: np.ndarray           # (S + 1,)
    vertices: np.ndarray          # (V, 3) collider at frame 0
    triangles: np.ndarray         # (T, 3)
    desc: dict = field(default_factory=dict)
    material: dict = field(default_factory=dict)
    frame_dt: float = 1.0 / 24.0
    # optional per-frame rigid motion: rotation angle (rad) about z per frame
    spin_per_frame: float = 0.0
    redraws: int = 0

    @property
    def strand_count(self) -> int:
        return self.offsets.size - 1

    def frame_points(self, frame: int) -> np.ndarray:
        return _rotate_z(self.points, self.spin_per_frame * frame)

    def frame_vertices(self, frame: int) -> np.ndarray:
        return _rotate_z(self.vertices, self.spin_per_frame * frame)


def _rotate_z(p: np.ndarray, angle: float) -> np.ndarray:
    if angle == 0.0:
        return p.copy()
    c, s = math.cos(angle), math.sin(angle)
    r = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    return p @ r.T


def icosphere(subdivisions: int, radius: float):
    """Icosahedron subdivided `subdivisions` times, vertices on the sphere,
    faces oriented outward."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, float) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}
        new_faces = []

        def midpoint(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            new_faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = new_faces
    return np.array(verts) * radius, np.array(faces, dtype=np.uint32)


def _perpendicular_rotation(rng: SplitMix64, t: np.ndarray, angle: float) -> np.ndarray:
    """Rotates unit tangent t by `angle` about a random axis perpendicular to it."""
    while True:
        a = rng.unit_vector()
        a = a - t * np.dot(a, t)
        n = np.linalg.norm(a)
        if n > 1e-6:
            a /= n
            break
    return t * math.cos(angle) + np.cross(a, t) * math.sin(angle)


def _segment_origin_distance(p0: np.ndarray, p1: np.ndarray) -> float:
    d = p1 - p0
    s = np.clip(-np.dot(p0, d) / max(np.dot(d, d), 1e-300), 0.0, 1.0)
    return float(np.linalg.norm(p0 + s * d))


def small_sphere_scene(strands: int = 64, points_per_strand: int = 32, seed: int = 1) -> Scene:
    """64 strands x 32 points, roots uniform on a sphere r = 0.10 m, length
    0.15 m, curl U(0, 0.3) rad per segment, icosphere collider (subdivision
    3) with vertex radius 0.098 m; E = 3 GPa with Poisson ratio 0.4999
    (G = 1.00007 GPa) since the ABI requires a ratio below 0.5."""
    rng = SplitMix64(seed)
    seg = 0.15 / (points_per_strand - 1)
    radius, body = 0.10, 0.098
    hair_radius = 4.0e-5
    pts = []
    redraws = 0
    for _ in range(strands):
        root_dir = rng.unit_vector()
        while True:
            p = root_dir * radius
            t = root_dir.copy()
            strand = [p.copy()]
            for _k in range(points_per_strand - 1):
                t = _perpendicular_rotation(rng, t, rng.uniform(0.0, 0.3))
                p = p + seg * t
                strand.append(p.copy())
            ok = True
            for k in range(2, points_per_strand - 1):
                if _segment_origin_distance(strand[k], strand[k + 1]) - body - hair_radius <= 1.0e-3:
                    ok = False
                    break
            if ok:
                break
            redraws += 1
        pts.extend(strand)
    verts, tris = icosphere(3, body)
    offsets = np.arange(0, strands * points_per_strand + 1, points_per_strand, dtype=np.uint32)
    return Scene(
        points=np.array(pts),
        offsets=offsets,
        vertices=verts,
        triangles=tris,
        desc=dict(substeps=4, maximum_substeps=32),
        material=dict(density=1.3e3, radius=hair_radius, young_modulus=3.0e9, poisson_ratio=0.4999),
        frame_dt=1.0 / 24.0,
        redraws=redraws,
    )


def configure(lib, scene: Scene, **desc_overrides):
    """Returns (desc, material) ABI structs for a scene."""
    import khs_native as kn

    desc = kn.default_desc(lib)
    for k, v in {**scene.desc, **desc_overrides}.items():
        if k == "gravity":
            desc.gravity = kn.KhsVec3(*v)
        else:
            setattr(desc, k, v)
    mat = kn.default_material(lib)
    for k, v in scene.material.items():
        setattr(mat, k, v)
    return desc, mat


def build_solver(lib, scene: Scene, collider: bool = True, **desc_overrides):
    """Creates, fills and builds a solver for `scene`."""
    import khs_native as kn

    desc, mat = configure(lib, scene, **desc_overrides)
    s = kn.HairSolver(lib, desc)
    s.check(s.set_hair(scene.points, scene.offsets, mat))
    if collider:
        s.check(s.set_collider(scene.vertices, scene.triangles))
    s.check(s.build())
    return s


def uv_ellipsoid(rings: int, segments: int, axes, center):
    """Closed UV ellipsoid: `rings` latitude rings of `segments` vertices plus
    two poles, faces oriented outward."""
    a, b, c = axes
    verts = [np.array([0.0, 0.0, c])]
    for i in range(1, rings + 1):
        theta = math.pi * i / (rings + 1)
        st, ct = math.sin(theta), math.cos(theta)
        phi = 2.0 * math.pi * np.arange(segments) / segments
        ring = np.stack([a * st * np.cos(phi), b * st * np.sin(phi), np.full(segments, c * ct)], axis=1)
        verts.extend(ring)
    verts.append(np.array([0.0, 0.0, -c]))
    verts = np.array(verts) + np.asarray(center, float)
    tris = []
    top, bottom = 0, 1 + rings * segments
    ring0 = lambda i: 1 + i * segments
    for j in range(segments):
        tris.append((top, ring0(0) + j, ring0(0) + (j + 1) % segments))
    for i in range(rings - 1):
        for j in range(segments):
            a0, a1 = ring0(i) + j, ring0(i) + (j + 1) % segments
            b0, b1 = ring0(i + 1) + j, ring0(i + 1) + (j + 1) % segments
            tris.append((a0, b0, b1))
            tris.append((a0, b1, a1))
    for j in range(segments):
        tris.append((bottom, ring0(rings - 1) + (j + 1) % segments, ring0(rings - 1) + j))
    return verts, np.array(tris, dtype=np.uint32)


HEAD_AXES = (0.078, 0.095, 0.115)


def _head_bulge(d: np.ndarray) -> np.ndarray:
    """Radial scale of the deformed head along unit directions d (rows)."""
    phi = np.arctan2(d[..., 1], d[..., 0])
    return 1.0 + 0.04 * np.sin(3.0 * phi) * (1.0 - d[..., 2] ** 2) + 0.03 * d[..., 2] ** 3


def _head_point(d: np.ndarray) -> np.ndarray:
    return _head_bulge(d)[..., None] * d * np.asarray(HEAD_AXES)


def head_body():
    """The head + torso body of the groom scenes: the deformed ellipsoid head
    (icosphere subdivision 6, 40,962 vertices) and a UV-ellipsoid torso
    (305 x 604 + 2 poles, 184,222 vertices); 225,184 vertices and 450,360
    triangles."""
    hv, ht = icosphere(6, 1.0)
    head_v = _head_point(hv)
    tv, tt = uv_ellipsoid(305, 604, TORSO_AXES, TORSO_CENTER)
    vertices = np.concatenate([head_v, tv])
    triangles = np.concatenate([ht, tt + head_v.shape[0]]).astype(np.uint32)
    return vertices, triangles


TORSO_AXES, TORSO_CENTER = (0.19, 0.12, 0.28), (0.0, 0.0, -0.43)


def uneven_scene(strands: int = 50000, seed: int = 3, lengths=(4, 256), spin_deg_per_s: float = 30.0,
                 fps: float = 24.0):
    """The uneven-length groom: `strands` strands with node counts uniform
    over `lengths` (inclusive; about 6.5M nodes at 50,000 strands) rooted
    on the upper part of the head body, 8 mm segments leaving along the
    surface normal and turning sideways by at most 1 rad over the strand (so
    no strand turns down into the body); per-strand Young's modulus U(1, 5)
    GPa and radius U(30, 80) um. Returns the scene and the two per-strand
    arrays. Generated with numpy's seeded generator (vectorized per strand)."""
    rng = np.random.default_rng(seed)
    vertices, triangles = head_body()
    d = rng.normal(size=(3 * strands, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d[d[:, 2] > 0.2][:strands]
    surface = _head_point(d)
    axes = np.asarray(HEAD_AXES)
    normal = surface / axes ** 2
    normal /= np.linalg.norm(normal, axis=1, keepdims=True)
    side = np.cross(normal, [0.0, 0.0, 1.0])
    side /= np.maximum(np.linalg.norm(side, axis=1, keepdims=True), 1e-12)
    counts = rng.integers(lengths[0], lengths[1] + 1, strands)
    turn = rng.uniform(0.0, 1.0, strands) / np.maximum(counts - 1, 1)
    pts, offsets = [], [0]
    for s in range(strands):
        ang = turn[s] * np.arange(counts[s] - 1)
        t = np.cos(ang)[:, None] * normal[s] + np.sin(ang)[:, None] * side[s]
        # 3 mm off the head surface so the deformed surface's facets stay clear
        p = surface[s] + 0.003 * normal[s] + 0.008 * np.concatenate([[np.zeros(3)], np.cumsum(t, axis=0)])
        pts.append(p)
        offsets.append(offsets[-1] + counts[s])
    young = rng.uniform(1.0e9, 5.0e9, strands)
    radius = rng.uniform(3.0e-5, 8.0e-5, strands)
    scene = Scene(points=np.concatenate(pts), offsets=np.array(offsets, dtype=np.uint32), vertices=vertices,
                  triangles=triangles, desc=dict(substeps=8, maximum_substeps=32),
                  material=dict(friction=0.3, barrier_distance=1.0e-3), frame_dt=1.0 / fps,
                  spin_per_frame=math.radians(spin_deg_per_s) / fps)
    return scene, young, radius


def head_scene(strands: int = 6757, seed: int = 2, spin_deg_per_s: float = 30.0, fps: float = 24.0) -> Scene:
    """Groom-scale scene: 6,757 strands with lognormal node counts
    (median 36, sigma 0.41, clipped 8..80) 8 mm apart on the upper part of a
    deformed ellipsoid head; body = head (icosphere subdivision 6, 40,962
    vertices) + torso (UV ellipsoid 305 x 604 + 2 poles, 184,222 vertices),
    225,184 vertices and 450,360 triangles; mu = 0.3, barrier 1 mm."""
    rng = SplitMix64(seed)
    vertices, triangles = head_body()
    torso_axes, torso_center = TORSO_AXES, TORSO_CENTER
    seg = 0.008
    pts, offsets = [], [0]
    redraws = 0
    axes = np.asarray(HEAD_AXES)

    def outside(p):
        # scaled-radius test against the deformed head and the torso
        r = np.linalg.norm(p / axes)
        d = p / max(np.linalg.norm(p), 1e-12)
        head_ok = r > float(_head_bulge(d / np.linalg.norm(d))) * 1.04 + 0.0
        q = (p - np.asarray(torso_center)) / np.asarray(torso_axes)
        torso_ok = np.linalg.norm(q) > 1.08
        return head_ok and torso_ok

    for _ in range(strands):
        n = int(min(80, max(8, round(36.0 * math.exp(0.41 * rng.normal())))))
        while True:
            d = rng.unit_vector()
            if d[2] > 0.2:
                break
        surface = _head_point(d)
        grad = surface / axes ** 2
        normal = grad / np.linalg.norm(grad)
        while True:
            p = surface + 0.0005 * normal
            t = normal.copy()
            strand = [p.copy()]
            for _k in range(n - 1):
                t = _perpendicular_rotation(rng, t, rng.uniform(0.0, 0.1))
                p = p + seg * t
                strand.append(p.copy())
            if all(outside(s) for s in strand[2:]):
                break
            redraws += 1
        pts.extend(strand)
        offsets.append(len(pts))
    spin = math.radians(spin_deg_per_s) / fps
    return Scene(
        points=np.array(pts),
        offsets=np.array(offsets, dtype=np.uint32),
        vertices=vertices,
        triangles=triangles,
        desc=dict(substeps=8, maximum_substeps=32),
        material=dict(friction=0.3, barrier_distance=1.0e-3),
        frame_dt=1.0 / fps,
        spin_per_frame=spin,
        redraws=redraws,
    )


def strand_subset(scene: Scene, count: int) -> Scene:
    """The first `count` strands of a scene with the same collider."""
    end = int(scene.offsets[count])
    return Scene(points=scene.points[:end].copy(), offsets=scene.offsets[: count + 1].copy(), vertices=scene.vertices,
                 triangles=scene.triangles, desc=dict(scene.desc), material=dict(scene.material),
                 frame_dt=scene.frame_dt, spin_per_frame=scene.spin_per_frame, redraws=scene.redraws)
