"""The oracle's earliest-TOI call shares the running minimum between its
queries (as distance_barrier_constraint.cpp:185-244 shares earliest_toi),
deterministically: each query still searches t in [0, 1] and only drops
boxes whose t.lo is at or above the running minimum. The minimum over the
call is then the same, bit for bit, as with every query's whole tree, in any
schedule and at any thread count. Checked here against the unshared call and
against a loop over the per-query results, with fewer inclusion-function
evaluations; and the -O3 timing build of the oracle returns the same bits
as the parity build, for single calls and over stepper runs.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from stepper_util import oracle_backend, resting_scene
from test_gpu_broadphase import cluster, moved, query_record
from rigidipc_amd import scenes as S
from rigidipc_amd import sim as M


def per_query_minimum(sc, p0, p1):
    ee, fv = O.broadphase(p0, p1, 0.0)
    best, evals = math.inf, 0
    for k, pairs in ((0, ee), (1, fv)):
        for a, b in pairs:
            hit, toi, n = (O.ee_toi if k == 0 else O.fv_toi)(query_record(sc, p0, p1, k, a, b))
            evals += n
            if hit:
                best = min(best, toi)
    return best, evals


@pytest.mark.parametrize("seed,dpos,drot", [(0, 0.05, 0.1), (1, 0.05, 0.1), (2, 0.1, 0.3), (3, 0.02, 0.05)])
def test_shared_minimum_gives_the_same_bits(seed, dpos, drot):
    sc = cluster(seed, n=27)
    p0, p1 = moved(sc, seed + 30, dpos=dpos, drot=drot)
    O.set_scene(sc)
    want, evals_loop = per_query_minimum(sc, p0, p1)
    try:
        O.set_ccd_pruning(False)
        full, _ = O.earliest_toi(p0, p1)
        evals_full = O.last_box_evaluations()
        O.set_ccd_pruning(True)
        results = []
        for threads in (1, 3, O.max_threads()):
            O.set_num_threads(threads)
            results.append(O.earliest_toi(p0, p1)[0])
        evals_pruned = O.last_box_evaluations()
    finally:
        O.set_ccd_pruning(True)
        O.set_num_threads(O.max_threads())
    print(f"seed {seed}: toi {want}, evaluations {evals_full} whole trees, {evals_pruned} shared", flush=True)
    assert math.isfinite(want)
    assert full == want and evals_full == evals_loop
    assert results == [want] * 3
    assert evals_pruned < evals_full


def test_no_impact_gives_inf_with_sharing():
    sc = cluster(4, n=27)
    p0 = sc.poses.copy()
    O.set_scene(sc)
    assert O.earliest_toi(p0, p0)[0] == math.inf


@pytest.mark.parametrize("scene", ["config1", "resting_friction"])
def test_timing_build_matches_parity_build(scene):
    """Stepper runs over the -O3 build's backend table and over the parity
    build's end in the same bits (CCD, constraint sets, barrier and, on the
    sliding cube, lagged friction)."""
    def make():
        if scene == "config1":
            return S.config1(seed=0), 30
        sc = resting_scene(mu=0.5, gap=4e-4, v=(0.6, 0.0, 0.2))
        sc.velocities[0, 3:] = [0.3, 1.0, -0.2]
        return sc, 10
    runs = []
    for be in (oracle_backend(), oracle_backend(O.lib_o3())):
        sc, n = make()
        sim = M.Simulation(sc, backend=be)
        sim.step(n)
        runs.append((sim.state(), sim.stats()))
        sim.close()
    (x2, v2), s2 = runs[0]
    (x3, v3), s3 = runs[1]
    print(f"{scene}: {s2['newton_iterations']} Newton iterations, {s2['ccd_candidates']} CCD candidates, "
          f"{s2['num_constraints']} constraints at the end", flush=True)
    assert s2["newton_iterations"] == s3["newton_iterations"] and s2["ccd_candidates"] == s3["ccd_candidates"]
    assert s2["ccd_candidates"] > 0 if scene == "config1" else s2["num_constraints"] > 0
    np.testing.assert_array_equal(x2, x3)
    np.testing.assert_array_equal(v2, v3)
