"""Oracle intersection check pinned by exact rational arithmetic.

The oracle's orient3d sign is compared with the determinant evaluated in
Python Fractions (every float converts exactly), and its edge-triangle rule
with an independent exact formulation: the plane crossing parameter
s = dp / (dp - dq) of the segment, then the three barycentric signs of
the crossing point (closed triangle; a segment lying in the plane of its
triangle, or a zero-area triangle, counts as no intersection). Cases:
random points, exactly coplanar and collinear configurations with dyadic
coordinates, segments through a triangle's interior, edges and vertices
and through the lines of its edges outside it, near-coplanar points made
in floating point, and coordinates spread over sixty orders of magnitude.
The scene-level check (every intersecting edge-face pair of close bodies)
is compared with a Fraction brute force over all pairs.
"""
import os
import sys
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rigid-ipc_amd"))

import oracle_lib as O  # noqa: E402
from rigidipc_amd import scenes as S  # noqa: E402


def frac(v):
    return [Fraction(float(x)) for x in v]


def sub(a, b):
    return [x - y for x, y in zip(a, b)]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def sign(x):
    return (x > 0) - (x < 0)


def orient_exact(a, b, c, d):
    a, b, c, d = frac(a), frac(b), frac(c), frac(d)
    return sign(dot(sub(a, d), cross(sub(b, d), sub(c, d))))


def edge_triangle_exact(p, q, t0, t1, t2):
    p, q, t0, t1, t2 = frac(p), frac(q), frac(t0), frac(t1), frac(t2)
    n = cross(sub(t1, t0), sub(t2, t0))
    if dot(n, n) == 0:
        return False
    dp, dq = dot(n, sub(p, t0)), dot(n, sub(q, t0))
    if dp == 0 and dq == 0:
        return False
    if sign(dp) * sign(dq) > 0:
        return False
    s = dp / (dp - dq)
    x = [pi + s * (qi - pi) for pi, qi in zip(p, q)]
    lam = [dot(n, cross(sub(t1, x), sub(t2, x))), dot(n, cross(sub(t2, x), sub(t0, x))),
           dot(n, cross(sub(t0, x), sub(t1, x)))]
    return all(v >= 0 for v in lam)


def test_orient3d_random_and_antisymmetric():
    rng = np.random.default_rng(1)
    for _ in range(400):
        a, b, c, d = rng.uniform(-1, 1, (4, 3))
        s = O.orient3d(a, b, c, d)
        assert s == orient_exact(a, b, c, d)
        assert O.orient3d(b, a, c, d) == -s
        assert O.orient3d(b, c, a, d) == s


def test_orient3d_coplanar_and_near_coplanar():
    rng = np.random.default_rng(2)
    zeros = nonzero = 0
    for _ in range(400):
        # dyadic points on one plane: exact zero
        a, b, c = rng.integers(-64, 64, (3, 3)) / 8.0
        u, v = rng.integers(-8, 9, 2) / 4.0
        d = a + u * (b - a) + v * (c - a)
        assert O.orient3d(a, b, c, d) == orient_exact(a, b, c, d)
        # the same in floating point with random coordinates: residuals of
        # a few ulps, exact sign either way
        a, b, c = rng.uniform(-1, 1, (3, 3))
        u, v = rng.uniform(-1, 2, 2)
        d = a + u * (b - a) + v * (c - a)
        s = orient_exact(a, b, c, d)
        zeros += s == 0
        nonzero += s != 0
        assert O.orient3d(a, b, c, d) == s
    assert nonzero > 100


def test_orient3d_wide_exponent_range():
    rng = np.random.default_rng(3)
    for _ in range(200):
        pts = rng.uniform(-1, 1, (4, 3)) * 10.0 ** rng.integers(-30, 31, (4, 3))
        assert O.orient3d(*pts) == orient_exact(*pts)
    # differences that need every bit: large coordinates with tiny offsets
    for _ in range(200):
        base = rng.uniform(-1, 1, 3) * 1e8
        pts = [base + rng.uniform(-1, 1, 3) * 1e-8 for _ in range(4)]
        assert O.orient3d(*pts) == orient_exact(*pts)


def triangle_point(t, w):
    return w[0] * t[0] + w[1] * t[1] + w[2] * t[2]


def test_edge_triangle_cases():
    rng = np.random.default_rng(4)
    seen = {True: 0, False: 0}
    weights = [(0.25, 0.25, 0.5), (0.5, 0.5, 0.0), (1.0, 0.0, 0.0), (1.5, -0.5, 0.0), (0.75, 0.5, -0.25),
               (0.125, 0.375, 0.5)]
    for it in range(600):
        t = rng.integers(-32, 32, (3, 3)) / 8.0
        kind = it % 6
        if kind == 0:  # random segment
            p, q = rng.uniform(-4, 4, (2, 3))
        elif kind == 1:  # through a point of the plane (interior, edge, vertex, outside)
            x = triangle_point(t, weights[rng.integers(len(weights))])
            n = np.cross(t[1] - t[0], t[2] - t[0])
            p, q = x + n * rng.integers(1, 4) / 4.0, x - n * rng.integers(1, 4) / 4.0
        elif kind == 2:  # one end point on the triangle's plane
            p = triangle_point(t, weights[rng.integers(len(weights))])
            q = rng.integers(-32, 32, 3) / 8.0
        elif kind == 3:  # segment in the plane
            p = triangle_point(t, weights[rng.integers(len(weights))])
            q = triangle_point(t, weights[rng.integers(len(weights))])
        elif kind == 4:  # crossing point a few ulps from an edge (floating point)
            tt = rng.uniform(-1, 1, (3, 3))
            x = tt[0] + rng.uniform(0, 1) * (tt[1] - tt[0])
            n = np.cross(tt[1] - tt[0], tt[2] - tt[0])
            t, p, q = tt, x + 0.3 * n, x - 0.2 * n
        else:  # degenerate triangle or segment
            if rng.integers(2):
                t[2] = t[0] + (t[1] - t[0]) * 0.5
                p, q = rng.uniform(-4, 4, (2, 3))
            else:
                p = triangle_point(t, (0.25, 0.25, 0.5))
                q = p.copy()
        want = edge_triangle_exact(p, q, *t)
        seen[want] += 1
        assert O.edge_triangle(p, q, *t) == want, (kind, p, q, t)
    assert seen[True] > 100 and seen[False] > 100


def static_body(V, F, p):
    return S.Body(np.asarray(V, float), np.asarray(F), np.eye(3), np.asarray(p, float), np.zeros(3), np.zeros(3),
                  static=True)


def brute_force(sc, poses):
    W = O.world_vertices(poses, len(sc.vertices))
    out = []
    nb = len(sc.group_ids)
    for i in range(nb):
        for j in range(i + 1, nb):
            if sc.group_ids[i] == sc.group_ids[j]:
                continue
            for X, Y in ((i, j), (j, i)):
                for e in range(sc.body_edge_start[X], sc.body_edge_start[X + 1]):
                    p, q = W[sc.edges[e]]
                    for f in range(sc.body_face_start[Y], sc.body_face_start[Y + 1]):
                        if edge_triangle_exact(p, q, *W[sc.faces[f]]):
                            out.append((e, f))
    return np.array(sorted(out), np.int32).reshape(-1, 2)


def test_scene_matches_fraction_brute_force():
    V, F = S.cube(1.0)
    Vp, Fp = S.plane(4.0)
    cases = {
        # interpenetrating cubes, dyadic offsets
        "overlap": S.assemble([static_body(V, F, (0, 0.5, 0)), static_body(V, F, (0.5, 0.75, 0.25))]),
        # a cube resting exactly on the plane (touching: zero orientations)
        "touching": S.assemble([static_body(Vp, Fp, (0, 0, 0)), static_body(V, F, (0.25, 0.5, 0.125))]),
        # separated
        "apart": S.assemble([static_body(V, F, (0, 0.5, 0)), static_body(V, F, (2, 0.5, 0))]),
        # same group: never checked
        "grouped": S.assemble([S.Body(V, F, np.eye(3), np.zeros(3), np.zeros(3), np.zeros(3), static=True, group=7),
                               S.Body(V, F, np.eye(3), np.full(3, 0.25), np.zeros(3), np.zeros(3), static=True,
                                      group=7)]),
    }
    counts = {}
    for name, sc in cases.items():
        O.set_scene(sc)
        got = O.detect_intersections(sc.poses)
        want = brute_force(sc, sc.poses)
        np.testing.assert_array_equal(got, want, err_msg=name)
        counts[name] = len(got)
    assert counts["overlap"] > 0 and counts["touching"] > 0
    assert counts["apart"] == 0 and counts["grouped"] == 0


def test_rotated_bodies_match_fraction_brute_force():
    rng = np.random.default_rng(5)
    V, F = S.icosphere(1, 0.5)
    Vc, Fc = S.cube(0.8)
    bodies = [S.Body(V, F, S.random_rotation(rng), rng.uniform(0, 0.6, 3), np.zeros(3), np.zeros(3)),
              S.Body(Vc, Fc, S.random_rotation(rng), rng.uniform(0, 0.6, 3), np.zeros(3), np.zeros(3)),
              S.Body(V, F, S.random_rotation(rng), rng.uniform(0, 0.6, 3), np.zeros(3), np.zeros(3))]
    sc = S.assemble(bodies)
    O.set_scene(sc)
    got = O.detect_intersections(sc.poses)
    np.testing.assert_array_equal(got, brute_force(sc, sc.poses))
    assert len(got) > 0
