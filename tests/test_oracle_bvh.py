"""Pin the oracle's BVH broad phase (DetectionMethod::BVH, the reference's
default; oracle/oracle_bvh.cpp) on the CPU.

* An independent float64 restatement of the reference's rules
  (src/ccd/rigid/rigid_body_bvh.hpp:64-94, rigid_body_bvh.cpp:5-190): close
  body pairs from inflated body boxes; the body with fewer faces is A; A's
  vertices in B's frame, ((V_A R_A^T) + (p_A - p_B)) R_B, boxes grown by r;
  B's faces found by the BVH query (A's face box grown by r again against
  B's face boxes in B's frame, inclusive); FV candidates through the face
  that owns the vertex (lowest face index), EE candidates through the faces
  that own both edges, oriented (edge of A, edge of B). The oracle's
  interval boxes are wider by a few ulps only, so on generic scenes the two
  sets are equal.
* Coverage: every EE / FV pair of different groups closer than 2r is a
  candidate, so the candidates within the constraint set's activation
  distance are those of the hash grid (whose coverage is pinned by brute
  force in test_oracle_broadphase.py), and the constraint set built from
  either is the same.
* Continuous: every pair that comes closer than 2r at a sampled time of the
  interpolated trajectories is a candidate; the earliest TOI over the BVH's
  candidates equals the hash grid's on the test scenes.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rigid-ipc_amd"))

import geom_ref as G  # noqa: E402
import oracle_lib as O  # noqa: E402
from rigidipc_amd import scenes as S  # noqa: E402


@pytest.fixture(autouse=True)
def hash_grid_after():
    yield
    O.set_detection_method("hash_grid")


def small_cluster(seed, n=8):
    rng = np.random.default_rng(seed)
    meshes = {0: S.icosphere(1, 0.5), 1: S.torus(0.5, 0.15, 12, 6), 2: S.cube(1.0)}
    bodies = S.lattice_bodies(rng, n, (3.0, 3.0, 3.0), 1.2, 0.0, 1.0, 2.0, meshes=meshes)
    bodies.append(S._ground())
    return S.assemble(bodies)


def perturbed(sc, seed, dpos=0.08, drot=0.3):
    rng = np.random.default_rng(seed)
    p = sc.poses.copy()
    dyn = ~sc.is_static.astype(bool)
    p[dyn] += rng.normal(size=(dyn.sum(), 6)) * [dpos, dpos, dpos, drot, drot, drot]
    return p


def ownership(sc):
    nv, ne = len(sc.vertices), len(sc.edges)
    v2f = np.full(nv, np.iinfo(np.int64).max)
    e2f = np.full(ne, np.iinfo(np.int64).max)
    for f in range(len(sc.faces)):
        for j in range(3):
            v2f[sc.faces[f, j]] = min(v2f[sc.faces[f, j]], f)
            e2f[sc.face_edges[f, j]] = min(e2f[sc.face_edges[f, j]], f)
    return v2f, e2f


def boxes(lo, hi):
    h = (hi - lo) / 2
    return lo, hi, lo + h, h


def overlap_center(a, b):
    return bool(np.all(np.abs(a[2] - b[2]) <= a[3] + b[3]))


def merge(*bs):
    lo = np.min([b[0] for b in bs], axis=0)
    hi = np.max([b[1] for b in bs], axis=0)
    return boxes(lo, hi)


def float_bvh(sc, poses, r):
    """The reference's BVH candidate rules in float64 (discrete)."""
    nb = len(sc.group_ids)
    W = G.world_vertices(sc, poses)
    bv = sc.body_vertex_start
    bf = sc.body_face_start
    bb = [(W[bv[b]:bv[b + 1]].min(axis=0) - r, W[bv[b]:bv[b + 1]].max(axis=0) + r) for b in range(nb)]
    v2f, e2f = ownership(sc)
    ee, fv = set(), set()
    for i in range(nb):
        for j in range(i + 1, nb):
            if sc.group_ids[i] == sc.group_ids[j]:
                continue
            if not (np.all(bb[i][0] <= bb[j][1]) and np.all(bb[j][0] <= bb[i][1])):
                continue
            A, B = (j, i) if bf[i + 1] - bf[i] > bf[j + 1] - bf[j] else (i, j)
            RA, RB = G.rodrigues(poses[A, 3:]), G.rodrigues(poses[B, 3:])
            VA = ((sc.vertices[bv[A]:bv[A + 1]] @ RA.T) + (poses[A, :3] - poses[B, :3])) @ RB
            VB = sc.vertices[bv[B]:bv[B + 1]]
            abox = [boxes(x - r, x + r) for x in VA]
            bbox = [boxes(x - r, x + r) for x in VB]
            for fa in range(bf[A], bf[A + 1]):
                FA = sc.faces[fa]
                fa_box = merge(*[abox[k - bv[A]] for k in FA])
                q_lo, q_hi = fa_box[0] - r, fa_box[1] + r
                for fb in range(bf[B], bf[B + 1]):
                    FB = sc.faces[fb]
                    P = sc.vertices[FB]
                    if not (np.all(q_lo <= P.max(axis=0)) and np.all(P.min(axis=0) <= q_hi)):
                        continue
                    fb_box = merge(*[bbox[k - bv[B]] for k in FB])
                    for k in range(3):
                        va, vb = FA[k], FB[k]
                        if v2f[va] == fa and overlap_center(abox[va - bv[A]], fb_box):
                            fv.add((fb, va))
                        if v2f[vb] == fb and overlap_center(fa_box, bbox[vb - bv[B]]):
                            fv.add((fa, vb))
                    for ea in sc.face_edges[fa]:
                        if e2f[ea] != fa:
                            continue
                        ea_box = merge(*[abox[k - bv[A]] for k in sc.edges[ea]])
                        for eb in sc.face_edges[fb]:
                            if e2f[eb] != fb:
                                continue
                            eb_box = merge(*[bbox[k - bv[B]] for k in sc.edges[eb]])
                            if overlap_center(ea_box, eb_box):
                                ee.add((int(ea), int(eb)))
    return ee, fv


def as_set(a):
    return {tuple(int(x) for x in row) for row in a}


@pytest.mark.parametrize("seed,r", [(0, 0.02), (1, 0.05), (2, 0.01)])
def test_bvh_matches_an_independent_restatement(seed, r):
    sc = small_cluster(seed)
    poses = perturbed(sc, 100 + seed)
    O.set_scene(sc)
    assert O.bvh_supported()
    O.set_detection_method("bvh")
    ee, fv = O.broadphase(poses, None, r)
    want_ee, want_fv = float_bvh(sc, poses, r)
    print(f"seed {seed}: ee {len(ee)} fv {len(fv)}", flush=True)
    assert len(want_ee) > 0 and len(want_fv) > 0
    assert as_set(ee) == want_ee
    assert as_set(fv) == want_fv
    # sorted, no duplicates
    assert len(as_set(ee)) == len(ee) and len(as_set(fv)) == len(fv)


def within(sc, poses, ee, fv, d):
    """The candidates whose elements are closer than d (squared distance
    from the oracle's distance routine, pinned in test_oracle_ipc_pins.py),
    EE pairs normalised to (lower, higher) edge id."""
    W = G.world_vertices(sc, poses)
    out_ee, out_fv = set(), set()
    for a, b in ee:
        X = np.stack([W[sc.edges[a, 0]], W[sc.edges[a, 1]], W[sc.edges[b, 0]], W[sc.edges[b, 1]]])
        if O.element_distance_sq(0, X) < d * d:
            out_ee.add((min(a, b), max(a, b)))
    for f, v in fv:
        X = np.stack([W[v], W[sc.faces[f, 0]], W[sc.faces[f, 1]], W[sc.faces[f, 2]]])
        if O.element_distance_sq(1, X) < d * d:
            out_fv.add((int(f), int(v)))
    return out_ee, out_fv


@pytest.mark.parametrize("seed,dhat", [(3, 0.02), (4, 0.05)])
def test_candidates_within_dhat_are_the_hash_grids(seed, dhat):
    """Both broad phases cover every pair closer than twice the inflation
    radius, so what the constraint set keeps (pairs within dhat, inflation
    dhat / 2) is the same from either."""
    sc = small_cluster(seed)
    poses = perturbed(sc, 200 + seed, dpos=0.1)
    O.set_scene(sc)
    O.set_detection_method("bvh")
    ee_b, fv_b = O.broadphase(poses, None, dhat / 2)
    O.set_detection_method("hash_grid")
    ee_h, fv_h = O.broadphase(poses, None, dhat / 2)
    kb = within(sc, poses, ee_b, fv_b, dhat)
    kh = within(sc, poses, ee_h, fv_h, dhat)
    print(f"seed {seed}: bvh {len(ee_b)}/{len(fv_b)}, grid {len(ee_h)}/{len(fv_h)}, "
          f"within dhat {len(kb[0])}/{len(kb[1])}", flush=True)
    assert len(kb[0]) + len(kb[1]) > 0
    assert kb == kh


def test_continuous_bvh_covers_sampled_contacts():
    sc = small_cluster(5)
    p0 = perturbed(sc, 300, dpos=0.02, drot=0.05)
    p1 = perturbed(sc, 301, dpos=0.15, drot=0.6)
    r = 0.01
    O.set_scene(sc)
    O.set_detection_method("bvh")
    ee, fv = O.broadphase(p0, p1, r)
    cand_ee = {(min(a, b), max(a, b)) for a, b in as_set(ee)}
    cand_fv = as_set(fv)
    # pairs closer than 2r at some sampled time must be candidates
    need_ee, need_fv = set(), set()
    O.set_detection_method("hash_grid")
    ee_h, fv_h = O.broadphase(p0, p1, r)
    for t in np.linspace(0.0, 1.0, 17):
        pt = p0 + t * (p1 - p0)
        a, b = within(sc, pt, ee_h, fv_h, 2 * r * (1 - 1e-9))
        need_ee |= a
        need_fv |= b
    print(f"candidates {len(cand_ee)}/{len(cand_fv)}, sampled contacts {len(need_ee)}/{len(need_fv)}", flush=True)
    assert len(need_ee) + len(need_fv) > 0
    assert need_ee <= cand_ee
    assert need_fv <= cand_fv


@pytest.mark.parametrize("seed", [6, 7, 8])
def test_earliest_toi_over_bvh_candidates_equals_hash_grid(seed):
    sc = small_cluster(seed)
    p0 = perturbed(sc, 400 + seed, dpos=0.02, drot=0.05)
    p1 = perturbed(sc, 500 + seed, dpos=0.2, drot=0.5)
    O.set_scene(sc)
    O.set_eval_budget(0)
    O.set_detection_method("bvh")
    toi_b, n_b = O.earliest_toi(p0, p1, 0.0)
    O.set_detection_method("hash_grid")
    toi_h, n_h = O.earliest_toi(p0, p1, 0.0)
    print(f"seed {seed}: toi {toi_b} ({n_b} candidates) vs {toi_h} ({n_h})", flush=True)
    assert np.isfinite(toi_h)
    assert toi_b == toi_h
