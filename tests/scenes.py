"""Scene builders shared by the tests: packing recorded CCD queries into one
multi-body scene, random queries, and procedural meshes."""
import types

import numpy as np


def queries_to_scene(kinds, qs):
    """Pack per-query records (36 doubles: 4 body-local vertices, poses
    A_t0, A_t1, B_t0, B_t1) into one scene with two bodies per query.

    EE query k -> body A = edge (a0,a1), body B = edge (b0,b1); candidate
    (edgeA, edgeB). FV query -> body A = the vertex, body B = the triangle
    (f0,f1,f2) with edges (f0,f1),(f1,f2),(f2,f0); candidate (face, vertex).
    Returns (scene, poses_t0, poses_t1, ee, fv, order) where order maps the
    narrow-phase output (EE first, then FV) back to the input index.
    """
    V, E, F, FE = [], [], [], []
    bvs, bes, bfs = [0], [0], [0]
    P0, P1 = [], []
    ee, fv, ee_idx, fv_idx = [], [], [], []
    for k, (kind, q) in enumerate(zip(kinds, qs)):
        q = np.asarray(q, np.float64)
        verts = q[:12].reshape(4, 3)
        pa0, pa1, pb0, pb1 = q[12:18], q[18:24], q[24:30], q[30:36]
        nv = len(V)
        if kind == 0:
            V += [verts[0], verts[1]]
            E.append([nv, nv + 1])
            ea = len(E) - 1
            bvs.append(len(V)); bes.append(len(E)); bfs.append(len(F))
            V += [verts[2], verts[3]]
            E.append([nv + 2, nv + 3])
            eb = len(E) - 1
            bvs.append(len(V)); bes.append(len(E)); bfs.append(len(F))
            ee.append([ea, eb]); ee_idx.append(k)
        else:
            V.append(verts[0])
            vi = len(V) - 1
            bvs.append(len(V)); bes.append(len(E)); bfs.append(len(F))
            b = len(V)
            V += [verts[1], verts[2], verts[3]]
            e0 = len(E)
            E += [[b, b + 1], [b + 1, b + 2], [b + 2, b]]
            F.append([b, b + 1, b + 2])
            FE.append([e0, e0 + 1, e0 + 2])
            bvs.append(len(V)); bes.append(len(E)); bfs.append(len(F))
            fv.append([len(F) - 1, vi]); fv_idx.append(k)
        P0 += [pa0, pb0]
        P1 += [pa1, pb1]
    nb = len(P0)
    scene = types.SimpleNamespace(
        vertices=np.array(V, np.float64).reshape(-1, 3),
        edges=np.array(E, np.int32).reshape(-1, 2),
        faces=np.array(F, np.int32).reshape(-1, 3),
        face_edges=np.array(FE, np.int32).reshape(-1, 3),
        body_vertex_start=np.array(bvs, np.int32),
        body_edge_start=np.array(bes, np.int32),
        body_face_start=np.array(bfs, np.int32),
        group_ids=np.arange(nb, dtype=np.int32),
        is_static=np.zeros(nb, np.uint8),
    )
    order = np.array(ee_idx + fv_idx, np.int64)
    return (scene, np.array(P0), np.array(P1), np.array(ee, np.int32).reshape(-1, 2),
            np.array(fv, np.int32).reshape(-1, 2), order)


def random_queries(n, seed, rot_scale=1.0, approach=True):
    """Random EE/FV queries between unit-size features that move and rotate
    towards each other over the step (so a fair share of them collide)."""
    rng = np.random.default_rng(seed)
    kinds, qs = [], []
    for i in range(n):
        kind = int(rng.integers(0, 2))
        verts = rng.normal(size=(4, 3)) * 0.5
        pa0 = np.concatenate([rng.normal(size=3) * 0.3, rng.normal(size=3) * rot_scale])
        pb0 = np.concatenate([rng.normal(size=3) * 0.3 + [0, 0.6, 0], rng.normal(size=3) * rot_scale])
        pa1 = pa0 + np.concatenate([rng.normal(size=3) * 0.4 + ([0, 0.5, 0] if approach else 0),
                                    rng.normal(size=3) * 0.5 * rot_scale])
        pb1 = pb0 + np.concatenate([rng.normal(size=3) * 0.4 - ([0, 0.5, 0] if approach else 0),
                                    rng.normal(size=3) * 0.5 * rot_scale])
        kinds.append(kind)
        qs.append(np.concatenate([verts.ravel(), pa0, pa1, pb0, pb1]))
    return np.array(kinds, np.int8), np.array(qs)
