"""Textbook closest-point distances in numpy, used to check the oracle's
continuous-collision stages by brute-force sampling.

closest_point_triangle follows Ericson, Real-Time Collision Detection,
section 5.1.5 (Voronoi regions of the triangle); segment_segment follows
section 5.1.9. segment_triangle is the minimum over the segment's endpoints
against the triangle and the segment against the triangle's three edges,
which is the distance whenever the segment does not pierce the triangle
(the cases sampled here are kept clear of that).
"""
import numpy as np


def closest_point_triangle(p, a, b, c):
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def segment_segment(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    eps = 1e-300
    if a <= eps and e <= eps:
        return float(np.linalg.norm(p1 - p2))
    if a <= eps:
        s, t = 0.0, np.clip(f / e, 0.0, 1.0)
    else:
        c = d1 @ r
        if e <= eps:
            t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
        else:
            b = d1 @ d2
            denom = a * e - b * b
            s = np.clip((b * f - c * e) / denom, 0.0, 1.0) if denom > 0 else 0.0
            t = (b * s + f) / e
            if t < 0:
                t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
            elif t > 1:
                t, s = 1.0, np.clip((b - c) / a, 0.0, 1.0)
    return float(np.linalg.norm((p1 + d1 * s) - (p2 + d2 * t)))


def segment_triangle(p0, p1, a, b, c):
    d = min(np.linalg.norm(p0 - closest_point_triangle(p0, a, b, c)),
            np.linalg.norm(p1 - closest_point_triangle(p1, a, b, c)))
    for u, v in ((a, b), (b, c), (c, a)):
        d = min(d, segment_segment(p0, p1, u, v))
    return float(d)


def pierces(p0, p1, a, b, c):
    """Whether the segment crosses the triangle's interior (Moller-Trumbore)."""
    d = p1 - p0
    e1, e2 = b - a, c - a
    h = np.cross(d, e2)
    det = e1 @ h
    if abs(det) < 1e-300:
        return False
    s = p0 - a
    u = (s @ h) / det
    qv = np.cross(s, e1)
    v = (d @ qv) / det
    t = (e2 @ qv) / det
    return 0 <= u <= 1 and 0 <= v and u + v <= 1 and 0 <= t <= 1
