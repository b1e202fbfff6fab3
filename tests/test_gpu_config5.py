"""Config 5 at full size: 10,000 bodies (9,000 ico3 spheres and 1,000 tori),
12.10M triangles, in a walled box above the plane.

The column starts with every body at least 1 cm from its neighbours and
falling with them, so the first time steps see only the bottom layer meet
the plane. The path is therefore also run at stressed poses: every body
moved by a random step (N(0, 2 cm) in position, N(0, 0.05) in rotation), so
that swept volumes overlap all over the column, and the constraint set with
a 2 cm activation distance.

* Properties at full size: candidate pairs of distinct, non-grouped bodies;
  earliest TOI in [0, 1] or +inf; constraint multiplicities >= 1; finite
  barrier energy and gradient; time and device memory are printed.
* Against the oracle on a sub-region (the bodies whose centres lie in the
  bottom corner of the box, with the plane and walls): what depends only on
  the two bodies of a pair must match the oracle run on the sub-scene alone,
  bit for bit: the discrete broad phase, the per-query rigid CCD of the
  sub-scene's continuous candidates, and the constraint set (members,
  multiplicities, mollifier thresholds).
* A few time steps of the real scene on the device stepper, with every
  step's end poses checked for intersections.
"""
import math
import time

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import PKG  # noqa: F401  (puts the package on sys.path)
from rigidipc_amd import scenes as S
from rigidipc_amd import sim as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    t = time.perf_counter()
    sc = S.config5(0)
    print(f"config5: {len(sc.group_ids)} bodies, {len(sc.vertices)} vertices, {len(sc.edges)} edges, "
          f"{len(sc.faces)} faces ({time.perf_counter() - t:.1f} s to generate)", flush=True)
    assert len(sc.faces) == 12096050
    return sc


@pytest.fixture(scope="module")
def ctx(scene):
    import gpu_util
    import torch
    free0, _ = torch.cuda.mem_get_info()
    c = gpu_util.context()
    t = time.perf_counter()
    c.set_scene(scene)
    free1, _ = torch.cuda.mem_get_info()
    print(f"set_scene {time.perf_counter() - t:.2f} s, device memory {(free0 - free1) / 2**30:.2f} GiB", flush=True)
    yield c
    c.close()


def stressed(sc, seed=1):
    rng = np.random.default_rng(seed)
    p0 = sc.poses.copy()
    p1 = p0.copy()
    dyn = ~sc.is_static.astype(bool)
    p1[dyn, :3] += rng.normal(0, 0.02, (dyn.sum(), 3))
    p1[dyn, 3:] += rng.normal(0, 0.05, (dyn.sum(), 3))
    return p0, p1


def region(sc, p):
    """Dynamic bodies whose centres lie in the box's bottom corner."""
    dyn = np.nonzero(~sc.is_static.astype(bool))[0]
    c = p[dyn, :3]
    return dyn[(c[:, 0] < -3.0) & (c[:, 2] < -3.0) & (c[:, 1] < 4.0)]


def sub_scene(sc, bodies):
    """The sub-scene of `bodies` plus every static body, and maps from its
    body / vertex / edge / face ids to the full scene's."""
    statics = np.nonzero(sc.is_static.astype(bool))[0]
    keep = np.concatenate([bodies, statics])
    sub = S.assemble([sc.bodies[i] for i in keep])

    def ranges(start):
        return np.concatenate([np.arange(start[i], start[i + 1]) for i in keep])
    return sub, keep, ranges(sc.body_vertex_start), ranges(sc.body_edge_start), ranges(sc.body_face_start)


def query_records(sc, p0, p1, kind, pairs):
    """The 36-double CCD records of candidate pairs (body-local vertices,
    then the two bodies' poses at t0 and t1), the layout of the reference's
    edge_edge_ccd / face_vertex_ccd (src/ccd/ccd.cpp:221-357)."""
    vb = np.repeat(np.arange(len(sc.group_ids)), np.diff(sc.body_vertex_start))
    V = sc.vertices
    a, b = pairs[:, 0], pairs[:, 1]
    if kind == 0:
        ea, eb = sc.edges[a], sc.edges[b]
        A, B = vb[ea[:, 0]], vb[eb[:, 0]]
        verts = [V[ea[:, 0]], V[ea[:, 1]], V[eb[:, 0]], V[eb[:, 1]]]
    else:
        f = sc.faces[a]
        A, B = vb[b], vb[f[:, 0]]
        verts = [V[b], V[f[:, 0]], V[f[:, 1]], V[f[:, 2]]]
    return np.ascontiguousarray(np.concatenate(verts + [p0[A], p1[A], p0[B], p1[B]], axis=1))


def test_full_size_properties(scene, ctx):
    sc = scene
    nb = len(sc.group_ids)
    p0, p1 = stressed(sc)
    t = time.perf_counter()
    ee, fv = ctx.broad_phase(p0, p1, 0.0)
    st = ctx.stats()
    print(f"continuous broad phase: {len(ee)} ee + {len(fv)} fv candidates, device {st['broad_phase_ms']:.2f} ms, "
          f"wall {1e3 * (time.perf_counter() - t):.1f} ms", flush=True)
    assert len(ee) + len(fv) > 10000
    vb = np.repeat(np.arange(nb), np.diff(sc.body_vertex_start))
    eb, fb = vb[sc.edges[:, 0]], vb[sc.faces[:, 0]]
    g = sc.group_ids
    assert np.all(g[eb[ee[:, 0]]] != g[eb[ee[:, 1]]]) and np.all(g[fb[fv[:, 0]]] != g[vb[fv[:, 1]]])
    t = time.perf_counter()
    toi = ctx.compute_earliest_toi(p0, p1)
    st = ctx.stats()
    print(f"earliest toi {toi}: {st['num_candidates']} candidates, broad {st['broad_phase_ms']:.2f} ms, narrow "
          f"{st['narrow_phase_ms']:.2f} ms, {st['box_evaluations']} box evaluations, "
          f"wall {1e3 * (time.perf_counter() - t):.1f} ms", flush=True)
    assert 0.0 <= toi <= 1.0
    t = time.perf_counter()
    counts = ctx.construct_constraint_set(p0, 0.02)
    e, grad, _ = ctx.barrier_term(p0, 0.02, grad=True)
    print(f"constraint set (dhat 2 cm) {counts.tolist()}, barrier {e:.6g}, "
          f"{1e3 * (time.perf_counter() - t):.1f} ms", flush=True)
    rec, _ = ctx.constraint_set()
    assert counts.sum() > 10000 and np.all(rec[:, 3] >= 1)
    assert math.isfinite(e) and e > 0 and np.all(np.isfinite(grad))
    # the static bodies share one group: no constraint between them
    st_mask = sc.is_static.astype(bool)
    assert not np.any(st_mask[rec[:, 4]] & st_mask[rec[:, 5]])


def test_sub_region_matches_oracle(scene, ctx):
    sc = scene
    p0, p1 = stressed(sc)
    bodies = region(sc, p0)
    sub, keep, vmap, emap, fmap = sub_scene(sc, bodies)
    print(f"sub-region: {len(bodies)} bodies + {len(keep) - len(bodies)} static", flush=True)
    assert len(bodies) >= 20
    inr = np.zeros(len(sc.group_ids), bool)
    inr[keep] = True
    vb = np.repeat(np.arange(len(sc.group_ids)), np.diff(sc.body_vertex_start))
    eb, fb = vb[sc.edges[:, 0]], vb[sc.faces[:, 0]]
    q0, q1 = p0[keep], p1[keep]
    O.set_scene(sub)
    # discrete broad phase at p1 (inflation 1 cm): no dependence on other bodies
    ee, fv = ctx.broad_phase(p1, None, 0.01)
    ee = ee[inr[eb[ee[:, 0]]] & inr[eb[ee[:, 1]]]]
    fv = fv[inr[fb[fv[:, 0]]] & inr[vb[fv[:, 1]]]]
    oee, ofv = O.broadphase(q1, None, 0.01)
    want_ee = np.sort(np.stack([emap[oee[:, 0]], emap[oee[:, 1]]], 1), axis=1) if len(oee) else oee
    want_fv = np.stack([fmap[ofv[:, 0]], vmap[ofv[:, 1]]], 1) if len(ofv) else ofv
    np.testing.assert_array_equal(ee[np.lexsort(ee.T[::-1])], want_ee[np.lexsort(want_ee.T[::-1])])
    np.testing.assert_array_equal(fv[np.lexsort(fv.T[::-1])], want_fv[np.lexsort(want_fv.T[::-1])])
    print(f"discrete broad phase: {len(ee)} ee + {len(fv)} fv in the region", flush=True)
    # per-query rigid CCD of the sub-scene's continuous candidates
    cee, cfv = O.broadphase(q0, q1, 0.0)
    want_toi = []
    for k, recs in ((0, query_records(sub, q0, q1, 0, cee)), (1, query_records(sub, q0, q1, 1, cfv))):
        for q in recs:
            hit, toi, _ = (O.ee_toi if k == 0 else O.fv_toi)(q)
            want_toi.append(toi if hit else math.inf)
    full_ee = np.stack([emap[cee[:, 0]], emap[cee[:, 1]]], 1).astype(np.int32)
    full_fv = np.stack([fmap[cfv[:, 0]], vmap[cfv[:, 1]]], 1).astype(np.int32)
    _, toi, hit = ctx.narrow_phase(p0, p1, full_ee, full_fv, per_query=True)
    got_toi = np.where(hit == 1, toi, math.inf)
    print(f"narrow phase: {len(want_toi)} queries, {int(np.isfinite(want_toi).sum())} hits", flush=True)
    assert len(want_toi) > 100
    np.testing.assert_array_equal(got_toi, np.array(want_toi))
    # constraint set at p0 with a 2 cm activation distance
    ctx.construct_constraint_set(p0, 0.02)
    rec, eps = ctx.constraint_set()
    sel = inr[rec[:, 4]] & inr[rec[:, 5]]
    orec, oeps = O.constraint_set(q0, 0.02)

    def full_key(r):
        t, a, b = int(r[0]), int(r[1]), int(r[2])
        if t == 0:
            a, b = sorted((int(vmap[a]), int(vmap[b])))
        elif t == 1:
            a, b = int(emap[a]), int(vmap[b])
        elif t == 2:
            a, b = int(emap[a]), int(emap[b])
        else:
            a, b = int(fmap[a]), int(vmap[b])
        return (t, a, b)
    want = {full_key(r): (int(r[3]), float(e)) for r, e in zip(orec, oeps)}
    got = {}
    for r, e in zip(rec[sel], eps[sel]):
        t, a, b = int(r[0]), int(r[1]), int(r[2])
        if t == 0:
            a, b = sorted((a, b))
        got[(t, a, b)] = (int(r[3]), float(e))
    print(f"constraint set: {len(want)} constraints in the region", flush=True)
    assert len(want) > 50
    assert got == want


def test_time_steps(scene):
    sc = scene
    p = M.default_params(sc)
    p.do_intersection_check = 1
    sim = M.Simulation(sc, params=p)
    for k in range(6):
        t = time.perf_counter()
        sim.step(1)
        st = sim.stats()
        print(f"step {k + 1}: {time.perf_counter() - t:.2f} s, Newton {st['newton_iterations']}, "
              f"constraints {st['num_constraints']}, failed {st['failed_solves']}", flush=True)
    poses, _ = sim.state()
    assert st["failed_solves"] == 0 and st["intersecting_steps"] == 0 and st["initial_intersections"] == 0
    assert np.all(np.isfinite(poses))
    sim.close()
