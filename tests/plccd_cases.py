"""Piecewise-linear CCD test cases shared by the CPU and GPU tests.

* The reference's closed-form translation scenes
  (tests/ccd/test_rigid_body_time_of_impact.cpp:170-303): an edge sweeping
  across another edge, toi = 1 / (1 - y_t1); a vertex falling onto a
  triangle, toi = 1 / (1 - z_t1). Bodies are given in body-local
  coordinates with zero rotation, which places every vertex where the
  reference's test puts it.
* A crossing detector for rotating queries: the first time the exact rigid
  trajectories make the two elements cross (sign change of the signed
  volume with the crossing point inside both elements), found by dense
  sampling and bisection. A conservative CCD must report an impact no later
  than that time.
"""
import itertools
import math

import numpy as np


def ee_translation_cases():
    out = []
    for z, sign, y_t1 in itertools.product((1000.0, 10.0, 1.0 + 1e-8, 1.0, 1.0 - 1e-8, 0.5, 0.0), (-1, 1),
                                           (2.0, 1.0, 0.1, 1e-6, 0.0, -1e-6, -1, -2, -10, -1000)):
        z_t0 = sign * z
        verts = [[-1, 0, 0], [1, 0, 0], [0, 0, -1], [0, 0, 1]]
        pa0 = [0, 1.0, z_t0, 0, 0, 0]
        pa1 = [0, y_t1, z_t0, 0, 0, 0]
        zero = [0.0] * 6
        q = np.array(sum(verts, []) + pa0 + pa1 + zero + zero, np.float64)
        expected = 1 / (1 - y_t1) if y_t1 != 1 else math.inf
        hit = y_t1 <= 0.0 and -1 <= z_t0 <= 1
        out.append((q, hit, expected))
    return out


def fv_translation_cases():
    out = []
    for y, z_t1 in itertools.product((1.0 + 1e-8, 1.0, 1.0 - 1e-8, 0.5, 1e-8, 0.0), (1, 1e-8, 0.0, -1e-8, -2.0, -10)):
        verts = [[0, y, 1], [-1, 0, 0], [1, 0, 0], [0, 1, 0]]
        zero = [0.0] * 6
        pa1 = [0, 0, z_t1 - 1, 0, 0, 0]
        q = np.array(sum(verts, []) + zero + pa1 + zero + zero, np.float64)
        expected = 1 / (1 - z_t1) if z_t1 != 1 else math.inf
        hit = z_t1 <= 0.0 and 0 <= y <= 1.0
        out.append((q, hit, expected))
    return out


def _rot(r):
    th = np.linalg.norm(r, axis=-1)[..., None, None]
    K = np.zeros(r.shape[:-1] + (3, 3))
    K[..., 0, 1], K[..., 0, 2], K[..., 1, 2] = -r[..., 2], r[..., 1], -r[..., 0]
    K[..., 1, 0], K[..., 2, 0], K[..., 2, 1] = r[..., 2], -r[..., 1], r[..., 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.where(th > 1e-12, np.sin(th) / th, 1.0)
        b = np.where(th > 1e-12, (1 - np.cos(th)) / th ** 2, 0.5)
    return np.eye(3) + a * K + b * (K @ K)


def _points(q, t):
    t = np.atleast_1d(np.asarray(t, np.float64))
    verts = q[:12].reshape(4, 3)
    pa0, pa1, pb0, pb1 = q[12:18], q[18:24], q[24:30], q[30:36]
    out = []
    for p0, p1, idx in ((pa0, pa1, None), (pb0, pb1, None)):
        pose = p0[None, :] + (p1 - p0)[None, :] * t[:, None]
        out.append((_rot(pose[:, 3:]), pose[:, :3]))
    return verts, out


def world(q, kind, t):
    """World positions [len(t), 4, 3] of the query's four vertices."""
    verts, ((RA, PA), (RB, PB)) = _points(q, t)
    nA = 2 if kind == 0 else 1
    X = np.empty((len(np.atleast_1d(t)), 4, 3))
    for i in range(4):
        R, P = (RA, PA) if i < nA else (RB, PB)
        X[:, i] = np.einsum("nij,j->ni", R, verts[i]) + P
    return X


def _volume(X):
    return np.einsum("ni,ni->n", X[:, 1] - X[:, 0], np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]))


def _inside(kind, X, margin=1e-7):
    x = X[0]
    if kind == 1:  # vertex x[0] against triangle x[1..3]
        e0, e1, p = x[2] - x[1], x[3] - x[1], x[0] - x[1]
        M = np.array([[e0 @ e0, e0 @ e1], [e0 @ e1, e1 @ e1]])
        u, v = np.linalg.solve(M, [p @ e0, p @ e1])
        return u >= margin and v >= margin and u + v <= 1 - margin
    a, b = x[1] - x[0], x[3] - x[2]
    w = x[0] - x[2]
    M = np.array([[a @ a, -(a @ b)], [a @ b, -(b @ b)]])
    s, r = np.linalg.solve(M, [-(w @ a), -(w @ b)])
    return margin <= s <= 1 - margin and margin <= r <= 1 - margin


def first_crossing(q, kind, n=4000):
    """First t in [0,1] where the elements cross along the exact rigid
    trajectories (inf if none is found)."""
    ts = np.linspace(0.0, 1.0, n + 1)
    vol = _volume(world(q, kind, ts))
    for k in np.nonzero(np.sign(vol[:-1]) * np.sign(vol[1:]) < 0)[0]:
        lo, hi = ts[k], ts[k + 1]
        vlo = vol[k]
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            vm = _volume(world(q, kind, [mid]))[0]
            if np.sign(vm) == np.sign(vlo):
                lo, vlo = mid, vm
            else:
                hi = mid
        if _inside(kind, world(q, kind, [hi])):
            return hi
    return math.inf


def resting_cases(msd=1e-3):
    """Stationary pairs at the separation distance's edge: a vertex above a
    triangle's interior and two crossing edges, msd * f apart. For f just
    above 1 the gap beyond msd is below the linear CCD's error filter, so
    it reports an impact at the piece's start for every piece; the
    reference then halves the first piece without end. Returns
    (kind, query, f)."""
    out = []
    zero = [0.0] * 6
    for f in (1.5, 1.000000000001, 1.0000000000001):
        h = f * msd
        fv = [[0.2, 0.2, h], [-1, 0, 0], [1, 0, 0], [0, 1, 0]]
        ee = [[-1, 0.1, h], [1, 0.1, h], [0.2, -1, 0], [0.2, 1, 0]]
        for kind, verts in ((1, fv), (0, ee)):
            out.append((kind, np.array(sum(verts, []) + zero * 4, np.float64), f))
    return out
