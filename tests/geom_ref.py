"""Closest points between mesh elements, for pinning the oracle's
restatement of ipc-toolkit's distances and constraint set (CPU tests only).

Nothing here classifies distance regions the way the oracle or ipc-toolkit
do. A squared distance is the minimum over an explicit list of candidate
closest-point configurations, each evaluated directly:

* the interior stationary point of the quadratic |x(s) - y(t)|^2 when it
  lies strictly inside the parameter domain (for two segments a convex
  quadratic, so such a point is the global minimum);
* every vertex against every opposite segment or vertex, with the segment
  parameter clamped to [0, 1].

The winning candidate names the closest features: interior, a vertex and
the inside of an edge, or two vertices. `margin` is the gap to the best
candidate with different features, so callers can tell a clear winner from
a tie. Slots follow ipc-toolkit's vertex order: edge-edge (ea0, ea1, eb0,
eb1), point-triangle (p, t0, t1, t2).
"""
import numpy as np


def point_segment(p, a, b):
    """(squared distance, clamped parameter t) of point p to segment ab."""
    ab = b - a
    l2 = float(ab @ ab)
    t = 0.0 if l2 == 0 else min(1.0, max(0.0, float((p - a) @ ab) / l2))
    c = a + t * ab - p
    return float(c @ c), t


def _pick(cands):
    cands = sorted(cands, key=lambda c: c[0])
    best = cands[0]
    other = [c[0] for c in cands if c[1] != best[1]]
    margin = (min(other) - best[0]) if other else np.inf
    return best[0], best[1], margin


def edge_edge(a0, a1, b0, b1):
    """Closest features of segments a = (a0, a1), b = (b0, b1).
    Returns (d2, feature, margin): feature ("EE",), ("EV", vertex slot,
    edge "a" | "b"), or ("VV", slot, slot) with the smaller slot first."""
    X = [np.asarray(x, float) for x in (a0, a1, b0, b1)]
    u, v, w = X[1] - X[0], X[3] - X[2], X[0] - X[2]
    a, b, c, d, e = u @ u, u @ v, v @ v, u @ w, v @ w
    D = a * c - b * b
    cands = []
    if D > 1e-12 * a * c:
        s = (b * e - c * d) / D
        t = (a * e - b * d) / D
        if 0 < s < 1 and 0 < t < 1:
            r = w + s * u - t * v
            cands.append((float(r @ r), ("EE",)))
    for slot, (q0, q1, side, ends) in enumerate(((2, 3, "b", (2, 3)), (2, 3, "b", (2, 3)),
                                                (0, 1, "a", (0, 1)), (0, 1, "a", (0, 1)))):
        d2, t = point_segment(X[slot], X[q0], X[q1])
        if 0 < t < 1:
            cands.append((d2, ("EV", slot, side)))
        else:
            other = ends[0] if t == 0 else ends[1]
            cands.append((d2, ("VV",) + tuple(sorted((slot, other)))))
    return _pick(cands)


def point_triangle(p, t0, t1, t2):
    """Closest features of point p and triangle (t0, t1, t2). Returns (d2,
    feature, margin): feature ("FV",), ("EV", (i, j)) with the triangle
    slots of the edge (1, 2), (2, 3) or (3, 1), or ("VV", slot)."""
    X = [np.asarray(x, float) for x in (p, t0, t1, t2)]
    e0, e1, r = X[2] - X[1], X[3] - X[1], X[0] - X[1]
    M = np.array([[e0 @ e0, e0 @ e1], [e0 @ e1, e1 @ e1]])
    cands = []
    if np.linalg.det(M) > 1e-12 * M[0, 0] * M[1, 1]:
        u, v = np.linalg.solve(M, [e0 @ r, e1 @ r])
        if u > 0 and v > 0 and u + v < 1:
            q = X[1] + u * e0 + v * e1 - X[0]
            cands.append((float(q @ q), ("FV",)))
    for i, j in ((1, 2), (2, 3), (3, 1)):
        d2, t = point_segment(X[0], X[i], X[j])
        if 0 < t < 1:
            cands.append((d2, ("EV", (i, j))))
        else:
            cands.append((d2, ("VV", i if t == 0 else j)))
    return _pick(cands)


def grid_min_edge_edge(a0, a1, b0, b1, n=201):
    """Brute-force minimum over an n x n grid of segment parameters."""
    s = np.linspace(0, 1, n)
    A = a0[None, :] + s[:, None] * (a1 - a0)[None, :]
    B = b0[None, :] + s[:, None] * (b1 - b0)[None, :]
    d = A[:, None, :] - B[None, :, :]
    return float((d * d).sum(-1).min())


def grid_min_point_triangle(p, t0, t1, t2, n=201):
    s = np.linspace(0, 1, n)
    u, v = np.meshgrid(s, s, indexing="ij")
    keep = u + v <= 1
    Q = t0 + u[keep][:, None] * (t1 - t0) + v[keep][:, None] * (t2 - t0)
    d = Q - p
    return float((d * d).sum(-1).min())


def rodrigues(r):
    th = float(np.linalg.norm(r))
    if th == 0:
        return np.eye(3)
    K = np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
    return np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K


def world_vertices(sc, poses):
    out = np.zeros_like(sc.vertices, dtype=float)
    for b in range(len(sc.group_ids)):
        a, z = sc.body_vertex_start[b], sc.body_vertex_start[b + 1]
        out[a:z] = sc.vertices[a:z] @ rodrigues(poses[b, 3:]).T + poses[b, :3]
    return out


def brute_force_constraints(sc, poses, dhat, tie_tol=1e-9):
    """ipc-toolkit's constraint set (ipc::construct_constraint_set as the
    reference calls it, distance_barrier_constraint.cpp:296-300) from every
    element pair of different groups whose squared distance is below dhat^2,
    with the closest features found by `edge_edge` / `point_triangle`:

    * an edge-edge pair whose cross product |ea x eb|^2 is below
      eps_x = 1e-3 |ea|^2 |eb|^2 stays an edge-edge constraint (mollified)
      whatever its closest features;
    * otherwise the closest features decide: edge-edge interior -> EE;
      vertex and the inside of an edge -> EV (edge, vertex); two vertices
      -> VV; point-triangle interior -> FV;
    * EV and VV constraints reached from several pairs are merged, with the
      number of pairs as multiplicity.

    Returns ({(type, a, b): multiplicity}, stats) with stats["ties"] the
    pairs whose features tie within tie_tol, stats["edge_cases"] the pairs
    within tie_tol of dhat^2, and stats["parallel_overrides"] the nearly
    parallel edge pairs kept as EE whose closest features are not interior
    to both edges. Keys:
    VV (0, min v, max v); EV (1, edge, vertex); EE (2, e0, e1) with e0 < e1;
    FV (3, face, vertex)."""
    W = world_vertices(sc, poses)
    nb = len(sc.group_ids)
    vb = np.repeat(np.arange(nb), np.diff(sc.body_vertex_start))
    E, F = np.asarray(sc.edges), np.asarray(sc.faces)
    eb = vb[E[:, 0]]
    fb = vb[F[:, 0]]
    edge_id = {}
    for i, (x, y) in enumerate(E):
        edge_id[(min(x, y), max(x, y))] = i
    g = np.asarray(sc.group_ids)
    d2max = dhat * dhat
    scale = max(1.0, float(np.abs(W).max())) ** 2
    out = {}
    stats = {"ties": 0, "edge_cases": 0, "parallel_overrides": 0}

    def add(key):
        out[key] = out.get(key, 0) + 1

    def boxes(idx_lists):
        P = W[idx_lists]
        return P.min(1), P.max(1)

    emn, emx = boxes(E)
    fmn, fmx = boxes(F)
    vmn = vmx = W
    for A in range(nb):
        for B in range(nb):
            if g[A] == g[B]:
                continue
            # edge-edge: each unordered pair once (A < B)
            if A < B:
                ia = np.nonzero(eb == A)[0]
                ib = np.nonzero(eb == B)[0]
                if len(ia) and len(ib):
                    gap = np.maximum(emn[ia][:, None, :] - emx[ib][None, :, :],
                                     emn[ib][None, :, :] - emx[ia][:, None, :]).max(-1)
                    for i, j in zip(*np.nonzero(gap < dhat)):
                        ea, ebb = ia[i], ib[j]
                        x = [W[E[ea, 0]], W[E[ea, 1]], W[E[ebb, 0]], W[E[ebb, 1]]]
                        d2, feat, margin = edge_edge(*x)
                        if abs(d2 - d2max) <= tie_tol * scale:
                            stats["edge_cases"] += 1
                        if not d2 < d2max:
                            continue
                        ua, ub = x[1] - x[0], x[3] - x[2]
                        n = np.cross(ua, ub)
                        if n @ n < 1e-3 * (ua @ ua) * (ub @ ub):
                            add((2, min(ea, ebb), max(ea, ebb)))
                            stats["parallel_overrides"] += feat[0] != "EE"
                            continue
                        if margin <= tie_tol * scale:
                            stats["ties"] += 1
                        vid = [E[ea, 0], E[ea, 1], E[ebb, 0], E[ebb, 1]]
                        if feat[0] == "EE":
                            add((2, min(ea, ebb), max(ea, ebb)))
                        elif feat[0] == "EV":
                            add((1, ebb if feat[2] == "b" else ea, vid[feat[1]]))
                        else:
                            v0, v1 = vid[feat[1]], vid[feat[2]]
                            add((0, min(v0, v1), max(v0, v1)))
            # face of A against vertex of B
            fa = np.nonzero(fb == A)[0]
            vbi = np.nonzero(vb == B)[0]
            if len(fa) and len(vbi):
                gap = np.maximum(fmn[fa][:, None, :] - vmx[vbi][None, :, :],
                                 vmn[vbi][None, :, :] - fmx[fa][:, None, :]).max(-1)
                for i, j in zip(*np.nonzero(gap < dhat)):
                    f, v = fa[i], vbi[j]
                    tri = [F[f, 0], F[f, 1], F[f, 2]]
                    d2, feat, margin = point_triangle(W[v], W[tri[0]], W[tri[1]], W[tri[2]])
                    if abs(d2 - d2max) <= tie_tol * scale:
                        stats["edge_cases"] += 1
                    if not d2 < d2max:
                        continue
                    if margin <= tie_tol * scale:
                        stats["ties"] += 1
                    if feat[0] == "FV":
                        add((3, f, v))
                    elif feat[0] == "EV":
                        x, y = tri[feat[1][0] - 1], tri[feat[1][1] - 1]
                        add((1, edge_id[(min(x, y), max(x, y))], v))
                    else:
                        w = tri[feat[1] - 1]
                        add((0, min(v, w), max(v, w)))
    return out, stats
