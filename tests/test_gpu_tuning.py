"""The rigid CCD's tunables (rigc_tuning) and its autotuner.

The tunables change only the schedule of the device's interval search
(walk budget before subtrees go to the next pass, speculative levels
evaluated by idle lanes, the first pass's window), never a result: every
setting must return the same earliest TOIs bit for bit, and those must be
the oracle's. The autotuner times a grid of settings on sample calls,
checks that they agree, and sets the fastest.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import oracle_lib as O  # noqa: E402
from test_gpu_broadphase import cluster, moved  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import gpu_util
    c = gpu_util.context()
    yield c
    c.close()


def sequence(sc, n, seed):
    out = []
    for k in range(n):
        p0, p1 = moved(sc, seed + k, dpos=0.05, drot=0.1)
        out.append((p0, p1))
    return out


def test_set_and_get_round_trip_and_bad_values(ctx):
    import rigidipc_amd as R
    ctx.set_scene(cluster(3, n=27))
    t0 = ctx.tuning()
    assert t0 == {"ccd_walk_budget": 4096, "ccd_spec_levels": 5, "ccd_window_factor": 4.0, "ccd_wide_queries": 4096}
    ctx.set_tuning(ccd_walk_budget=256, ccd_spec_levels=2, ccd_window_factor=0.0, ccd_wide_queries=0)
    assert ctx.tuning() == {"ccd_walk_budget": 256, "ccd_spec_levels": 2, "ccd_window_factor": 0.0,
                            "ccd_wide_queries": 0}
    for bad in ({"ccd_spec_levels": 6}, {"ccd_spec_levels": -1}, {"ccd_walk_budget": -5},
                {"ccd_window_factor": -1.0}, {"ccd_window_factor": float("inf")}, {"ccd_wide_queries": -1}):
        with pytest.raises(R.RigcError):
            ctx.set_tuning(**bad)
    ctx.set_tuning(**t0)


@pytest.mark.parametrize("seed", [0, 5])
def test_every_tuning_gives_the_oracle_toi(ctx, seed):
    sc = cluster(seed, n=27)
    ctx.set_scene(sc)
    seq = sequence(sc, 4, 100 + seed)
    O.set_scene(sc)
    O.set_eval_budget(0)
    want = [O.earliest_toi(p0, p1, 0.0)[0] for p0, p1 in seq]
    print("oracle", want, flush=True)
    base = ctx.tuning()
    for wb in (0, 64, 4096):
        for sl in (0, 3, 5):
            for wf in (0.0, 4.0):
                for wq in (0, 1 << 40):
                    ctx.set_tuning(ccd_walk_budget=wb, ccd_spec_levels=sl, ccd_window_factor=wf, ccd_wide_queries=wq)
                    got = [ctx.compute_earliest_toi(p0, p1) for p0, p1 in seq]
                    assert got == want, (wb, sl, wf, wq)
    ctx.set_tuning(**base)


def test_autotuner_sets_a_grid_tuning_and_keeps_results(ctx):
    sc = cluster(7, n=60)
    ctx.set_scene(sc)
    seq = sequence(sc, 3, 300)
    want = [ctx.compute_earliest_toi(p0, p1) for p0, p1 in seq]
    best, ms = ctx.autotune_earliest_toi(seq)
    print("autotuned", best, ms, flush=True)
    assert ms > 0
    assert best["ccd_walk_budget"] in (1024, 4096, 16384)
    assert best["ccd_spec_levels"] in (3, 4, 5)
    assert best["ccd_window_factor"] in (0.0, 4.0, 16.0)
    assert best["ccd_wide_queries"] in (0, 4096, 65536)
    assert ctx.tuning() == best
    assert [ctx.compute_earliest_toi(p0, p1) for p0, p1 in seq] == want


def test_autotuner_leaves_knobs_outside_its_mask(ctx):
    """knobs restricts the search: the other fields keep the context's
    setting, and the results stay the same."""
    sc = cluster(8, n=60)
    ctx.set_scene(sc)
    seq = sequence(sc, 3, 310)
    want = [ctx.compute_earliest_toi(p0, p1) for p0, p1 in seq]
    base = ctx.tuning()
    ctx.set_tuning(ccd_spec_levels=4, ccd_window_factor=16.0)
    try:
        best, ms = ctx.autotune_earliest_toi(seq, knobs=["ccd_walk_budget", "ccd_wide_queries"])
        print("autotuned (walk budget, launch width)", best, ms, flush=True)
        assert best["ccd_spec_levels"] == 4 and best["ccd_window_factor"] == 16.0
        assert best["ccd_walk_budget"] in (1024, 4096, 16384)
        assert best["ccd_wide_queries"] in (0, 4096, 65536)
        assert ctx.tuning() == best
        assert [ctx.compute_earliest_toi(p0, p1) for p0, p1 in seq] == want
    finally:
        ctx.set_tuning(**base)
