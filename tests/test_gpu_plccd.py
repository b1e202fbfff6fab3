"""GPU parity of the piecewise-linear CCD (rigc_narrow_phase_piecewise_linear,
rigc_compute_earliest_toi_piecewise_linear) against the CPU oracle.

Bar: hit flags and TOIs bit-exact query by query, and the same number of
inclusion boxes visited in total (the breadth-first traversal visits boxes
in the same order), on
  * the reference's closed-form translation scenes
    (tests/ccd/test_rigid_body_time_of_impact.cpp:170-303),
  * random rotating queries (seeded), with and without a minimum separation,
  * a packed multi-body scene through the broad phase (earliest TOI = the
    oracle's minimum over the oracle's candidates).
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import oracle_lib as O  # noqa: E402
import plccd_cases as C  # noqa: E402
import scenes  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import gpu_util
    c = gpu_util.context()
    yield c
    c.close()


def oracle_all(kinds, qs, earliest, msd):
    hit = np.zeros(len(kinds), np.int8)
    toi = np.full(len(kinds), math.inf)
    boxes = 0
    for i, (k, q) in enumerate(zip(kinds, qs)):
        h, t, _, b = (O.pl_ee_toi if k == 0 else O.pl_fv_toi)(q, earliest, msd)
        hit[i] = 1 if h else 0
        toi[i] = t if h else math.inf
        boxes += b
    return hit, toi, boxes


def device_all(ctx, kinds, qs, earliest, msd):
    scene, p0, p1, ee, fv, order = scenes.queries_to_scene(kinds, qs)
    ctx.set_scene(scene)
    best, toi, hit = ctx.narrow_phase_piecewise_linear(p0, p1, ee, fv, earliest_toi=earliest,
                                                       minimum_separation_distance=msd)
    out_toi = np.empty(len(kinds))
    out_hit = np.empty(len(kinds), np.int8)
    out_toi[order] = toi
    out_hit[order] = hit
    return best, out_hit, out_toi, ctx.stats()


def check(ctx, kinds, qs, earliest=1.0, msd=0.0):
    want_hit, want_toi, want_boxes = oracle_all(kinds, qs, earliest, msd)
    best, hit, toi, st = device_all(ctx, kinds, qs, earliest, msd)
    print(f"queries {len(kinds)} hits {int(want_hit.sum())} boxes {want_boxes} device boxes "
          f"{st['box_evaluations']} passes {st['ccd_launches']} kernel {st['ccd_kernel_ms']:.3f} ms", flush=True)
    np.testing.assert_array_equal(hit, want_hit)
    np.testing.assert_array_equal(toi, want_toi)
    assert st['box_evaluations'] == want_boxes
    assert best == (want_toi[want_hit == 1].min() if want_hit.any() else math.inf)


@pytest.mark.parametrize("kind", [0, 1])
def test_translation_closed_forms_bit_exact(ctx, kind):
    cases = C.ee_translation_cases() if kind == 0 else C.fv_translation_cases()
    kinds = np.full(len(cases), kind, np.int8)
    qs = np.array([c[0] for c in cases])
    check(ctx, kinds, qs)


@pytest.mark.parametrize("seed,rot,msd", [(1, 0.3, 0.0), (2, 1.0, 0.0), (3, 2.5, 0.0), (2, 1.0, 1e-4)])
def test_random_queries_bit_exact(ctx, seed, rot, msd):
    kinds, qs = scenes.random_queries(150, seed, rot_scale=rot)
    check(ctx, kinds, qs, msd=msd)


def test_resting_pairs_at_the_separation_edge_bit_exact(ctx):
    """Pairs resting just outside the separation distance, where the
    reference halves the first piece without end: the device stops after
    the same 64 halvings as the oracle, with an impact at t = 0."""
    cases = C.resting_cases()
    kinds = np.array([k for k, _, _ in cases], np.int8)
    qs = np.array([q for _, q, _ in cases])
    check(ctx, kinds, qs, msd=1e-3)


def test_partial_time_domain_bit_exact(ctx):
    kinds, qs = scenes.random_queries(150, 5, rot_scale=1.0)
    check(ctx, kinds, qs, earliest=0.4)


@pytest.fixture(scope="module")
def small_ctx():
    """A context whose piecewise-linear CCD starts in tiny HBM lane slices
    (stack 2 / heap 8, then 4 / 64), so queries take the re-run passes."""
    import gpu_util
    os.environ["RIGC_PLCCD_SMALL_SLICES"] = "1"
    try:
        c = gpu_util.context()
    finally:
        del os.environ["RIGC_PLCCD_SMALL_SLICES"]
    yield c
    c.close()


@pytest.mark.parametrize("seed,rot,msd,earliest", [(2, 1.0, 0.0, 1.0), (3, 2.5, 1e-4, 1.0), (5, 1.0, 0.0, 0.4)])
def test_rerun_passes_bit_exact(small_ctx, seed, rot, msd, earliest):
    """Queries that outgrow their lane slice stop, have their box counts
    taken back, and run again from a re-run list with larger slices; the
    results and box totals must still be the oracle's."""
    kinds, qs = scenes.random_queries(150, seed, rot_scale=rot)
    check(small_ctx, kinds, qs, earliest=earliest, msd=msd)
    assert small_ctx.stats()['ccd_launches'] >= 2


def test_rerun_passes_through_broad_phase(small_ctx):
    from test_gpu_broadphase import cluster, moved
    sc = cluster(4, n=27)
    p0, p1 = moved(sc, 34, dpos=0.05, drot=0.1)
    small_ctx.set_scene(sc)
    got = small_ctx.compute_earliest_toi_piecewise_linear(p0, p1, 0.0)
    assert small_ctx.stats()['ccd_launches'] >= 2
    assert got == _oracle_earliest(sc, p0, p1, 0.0)[0]


def _oracle_earliest(sc, p0, p1, msd):
    """The oracle's minimum TOI over the oracle's candidates (broad phase
    inflated by msd / 2, as compute_earliest_toi does)."""
    from test_gpu_broadphase import query_record
    O.set_scene(sc)
    ee, fv = O.broadphase(p0, p1, msd / 2)
    best = math.inf
    for k, pairs in ((0, ee), (1, fv)):
        for a, b in pairs:
            hit, toi, _, _ = (O.pl_ee_toi if k == 0 else O.pl_fv_toi)(query_record(sc, p0, p1, k, a, b), 1.0, msd)
            if hit:
                best = min(best, toi)
    return best, len(ee) + len(fv)


@pytest.mark.parametrize("msd", [0.0, 1e-4])
def test_earliest_toi_matches_oracle(ctx, msd):
    from test_gpu_broadphase import cluster, moved
    sc = cluster(4, n=27)
    p0, p1 = moved(sc, 34, dpos=0.05, drot=0.1)
    best, n_cand = _oracle_earliest(sc, p0, p1, msd)
    ctx.set_scene(sc)
    got = ctx.compute_earliest_toi_piecewise_linear(p0, p1, msd)
    st = ctx.stats()
    print(f"msd {msd} candidates {n_cand} earliest toi {best} kernel {st['ccd_kernel_ms']:.3f} ms", flush=True)
    assert st['num_candidates'] == n_cand
    assert got == best
