"""GPU parity of the rigid CCD narrow phase against the CPU oracle.

Bar: integer results (hit flags) and TOIs bit-exact, query by query, on
  * the reference's recorded CCD queries (tests/golden/ccd_queries.npz,
    from tests/data/{wrecking-ball,kinematic}/ and tests/data/ccd-test-*),
  * the reference's closed-form TOI scenes (EE / FV translation),
  * random rotating queries (seeded), checked against the oracle live;
and the earliest-TOI mode (cross-query pruning) equal to the minimum of
the per-query TOIs.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import oracle_lib as O  # noqa: E402
import scenes  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import gpu_util
    c = gpu_util.context()
    yield c
    c.close()


def run_gpu(ctx, kinds, qs, per_query=True, budget=0, allow_unresolved=False, toi_tol=1e-4):
    scene, p0, p1, ee, fv, order = scenes.queries_to_scene(kinds, qs)
    ctx.set_scene(scene)
    best, toi, hit = ctx.narrow_phase(p0, p1, ee, fv, toi_tolerance=toi_tol, eval_budget=budget,
                                      per_query=per_query, allow_unresolved=allow_unresolved)
    out_toi = np.empty(len(kinds))
    out_hit = np.empty(len(kinds), np.int8)
    out_toi[order] = toi
    out_hit[order] = hit
    return best, out_toi, out_hit


def test_recorded_queries_bit_exact(ctx):
    g = np.load(os.path.join(HERE, "golden", "ccd_queries.npz"))
    keep = g["hit"] >= 0  # the 6 queries beyond the oracle's budget are checked below
    kinds, qs = g["kind"][keep], g["q"][keep]
    print(f"recorded queries: {keep.sum()} (EE {np.sum(kinds == 0)}, FV {np.sum(kinds == 1)})", flush=True)
    _, toi, hit = run_gpu(ctx, kinds, qs)
    np.testing.assert_array_equal(hit, g["hit"][keep])
    np.testing.assert_array_equal(toi.view(np.int64), g["toi"][keep].view(np.int64))


def test_recorded_queries_budget_exceeded_agrees(ctx):
    g = np.load(os.path.join(HERE, "golden", "ccd_queries.npz"))
    idx = np.nonzero(g["hit"] < 0)[0]
    budget = 20000
    O.set_eval_budget(budget)
    try:
        ref = [(O.ee_toi if g["kind"][i] == 0 else O.fv_toi)(g["q"][i]) for i in idx]
    finally:
        O.set_eval_budget(0)
    _, toi, hit = run_gpu(ctx, g["kind"][idx], g["q"][idx], budget=budget, allow_unresolved=True)
    assert [int(r[0]) for r in ref] == [int(h) for h in hit]


def _oracle_all(kinds, qs):
    hits, tois = [], []
    for k, q in zip(kinds, qs):
        h, t, _ = (O.ee_toi if k == 0 else O.fv_toi)(q)
        hits.append(int(h)); tois.append(t)
    return np.array(hits, np.int8), np.array(tois)


@pytest.mark.parametrize("seed,rot", [(0, 0.2), (1, 1.0), (2, 3.0)])
def test_random_rotating_queries_bit_exact(ctx, seed, rot):
    kinds, qs = scenes.random_queries(300, seed, rot_scale=rot)
    # The oracle walks each tree depth-first with a visit budget; the GPU
    # walks the same tree in parallel (a different visit order), so it gets
    # a larger budget and must resolve, bit for bit, every query the oracle
    # resolved.
    budget = 50000
    O.set_eval_budget(budget)
    try:
        h_ref, t_ref = _oracle_all(kinds, qs)
    finally:
        O.set_eval_budget(0)
    _, toi, hit = run_gpu(ctx, kinds, qs, budget=20 * budget, allow_unresolved=True)
    print(f"seed {seed}: hits {np.sum(h_ref == 1)} unresolved {np.sum(h_ref < 0)}", flush=True)
    ok = h_ref >= 0
    np.testing.assert_array_equal(hit[ok], h_ref[ok])
    np.testing.assert_array_equal(toi[ok].view(np.int64), t_ref[ok].view(np.int64))


def test_closed_form_translation_scenes(ctx):
    # tests/ccd/test_rigid_body_time_of_impact.cpp:164-303 with toi_tol 1e-6
    kinds, qs, expect = [], [], []
    Z6 = np.zeros(6)
    for z in [1000.0, 10.0, 1.0 + 1e-8, 1.0, 1.0 - 1e-8, 0.5, 0.0]:
        for sign in (-1, 1):
            for y_t1 in [2.0, 1.0, 0.1, 1e-6, 0.0, -1e-6, -1, -2, -10, -1000]:
                zt = sign * z
                q = np.concatenate([[-1, 0, 0], [1, 0, 0], [0, 0, -1], [0, 0, 1],
                                    [0, 1.0, zt, 0, 0, 0], [0, y_t1, zt, 0, 0, 0], Z6, Z6])
                kinds.append(0); qs.append(q)
                expect.append((y_t1 <= 0 and -1 <= zt <= 1, 1 / (1 - y_t1) if y_t1 != 1 else math.inf))
    for y in [1.0 + 1e-8, 1.0, 1.0 - 1e-8, 0.5, 1e-8, 0.0]:
        for z_t1 in [1, 1e-8, 0.0, -1e-8, -2.0, -10]:
            q = np.concatenate([[0, y, 1], [-1, 0, 0], [1, 0, 0], [0, 1, 0], Z6,
                                [0, 0, z_t1 - 1.0, 0, 0, 0], Z6, Z6])
            kinds.append(1); qs.append(q)
            expect.append((z_t1 <= 0 and 0 <= y <= 1, 1 / (1 - z_t1) if z_t1 != 1 else math.inf))
    kinds = np.array(kinds, np.int8)
    qs = np.array(qs, np.float64)
    _, toi, hit = run_gpu(ctx, kinds, qs, toi_tol=1e-6)
    for i, (eh, et) in enumerate(expect):
        assert bool(hit[i]) == eh, i
        if eh:
            assert toi[i] <= et and abs(toi[i] - et) <= max(1e-4, 1.2e-5 * abs(et))
    # and bit-equal to the oracle with the same tolerance
    for i in range(len(kinds)):
        h, t, _ = (O.ee_toi if kinds[i] == 0 else O.fv_toi)(qs[i], 1.0, 1e-6)
        assert int(h) == int(hit[i]) and (not h or t == toi[i]), i


def test_earliest_mode_is_min_of_per_query(ctx):
    kinds, qs = scenes.random_queries(600, 7, rot_scale=1.0)
    best_pq, toi, hit = run_gpu(ctx, kinds, qs, per_query=True, budget=200000, allow_unresolved=True)
    best, _, _ = run_gpu(ctx, kinds, qs, per_query=False, budget=200000, allow_unresolved=True)
    expect = toi[hit == 1].min() if np.any(hit == 1) else math.inf
    assert best_pq == expect
    assert best == expect


def test_empty_candidate_list(ctx):
    kinds, qs = scenes.random_queries(2, 3)
    scene, p0, p1, ee, fv, order = scenes.queries_to_scene(kinds, qs)
    ctx.set_scene(scene)
    best, toi, hit = ctx.narrow_phase(p0, p1, np.zeros((0, 2)), np.zeros((0, 2)))
    assert best == math.inf and len(toi) == 0
