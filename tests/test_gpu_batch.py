"""Batched scenes through the C ABI (rigc_set_scenes + rigc_batch_*): every
per-scene result of a call that serves several scenes must equal, bit for
bit, the result of a context that holds that scene alone (same kernels, and
every reduction is done per scene in an order that does not depend on the
other scenes of the call). Checked for the earliest TOI (windowed search
included), the barrier energy / gradient / Hessian, the minimum distance and
the lagged friction sets and potentials, on scene subsets and on all
scenes, plus the single-scene entry points on a multi-scene context."""
import numpy as np
import pytest

from gpu_util import context
from rigidipc_amd import scenes as S

pytestmark = pytest.mark.gpu

SCENE_IDS = (0, 3, 5, 9, 12)
DHAT = 0.05


@pytest.fixture(scope="module")
def setup():
    scenes = [S.config3_scene(i) for i in SCENE_IDS]
    batch = context()
    batch.set_scenes(scenes)
    plain = []
    for sc in scenes:
        c = context()
        c.set_scene(sc)
        plain.append(c)
    print("scene bodies", [len(sc.group_ids) for sc in scenes], flush=True)
    yield scenes, batch, plain
    batch.close()
    for c in plain:
        c.close()


def moved(sc, seed, dpos=0.05, drot=0.1):
    rng = np.random.default_rng(seed)
    p0 = sc.poses.copy()
    p1 = p0.copy()
    dyn = ~sc.is_static.astype(bool)
    p1[dyn] += rng.normal(size=(dyn.sum(), 6)) * [dpos, dpos, dpos, drot, drot, drot]
    # everything also falls towards the plane, so the bottom layer collides
    p1[dyn, 1] -= 0.03
    return p0, p1


def lowered(sc, seed):
    """Poses with every dynamic body moved down by 8 mm (the bottom layer
    starts 10 mm above the plane) and slightly rotated."""
    rng = np.random.default_rng(seed)
    p = sc.poses.copy()
    dyn = ~sc.is_static.astype(bool)
    p[dyn, 1] -= 0.008
    p[dyn, 3:] += rng.normal(size=(dyn.sum(), 3)) * 0.01
    return p


def test_earliest_toi_matches_single_scene_contexts(setup):
    scenes, batch, plain = setup
    finite = 0
    for subset in ((0, 2, 3), tuple(range(len(scenes))), (4,)):
        for s in (1.0, 0.5, 0.02, 1.0):
            p0s, p1s, want = [], [], []
            for k in subset:
                p0, p1 = moved(scenes[k], 40 + k)
                p1 = p0 + s * (p1 - p0)
                p0s.append(p0)
                p1s.append(p1)
                want.append(plain[k].compute_earliest_toi(p0, p1))
            got = batch.batch_earliest_toi(subset, p0s, p1s)
            print(subset, s, want, flush=True)
            assert list(got) == want
            finite += int(np.sum(np.isfinite(want)))
    assert finite > 0


def test_previous_toi_handover_keeps_results(setup):
    """rigc_batch_set_previous_toi only bounds the first pass of the next
    call: whatever previous TOI a scene is handed (tiny, large, none), the
    earliest TOI is the single-scene context's."""
    scenes, batch, plain = setup
    subset = (0, 2, 3)
    p0s, p1s, want = [], [], []
    for k in subset:
        p0, p1 = moved(scenes[k], 40 + k)
        p0s.append(p0)
        p1s.append(p1)
        want.append(plain[k].compute_earliest_toi(p0, p1))
    for prev in (1e-9, 1e-3, 0.2, float("inf")):
        batch.batch_set_previous_toi(subset, [prev] * len(subset))
        assert list(batch.batch_earliest_toi(subset, p0s, p1s)) == want, prev
    assert any(np.isfinite(want))


def test_per_scene_ccd_evaluations_add_up(setup):
    """rigc_batch_ccd_evaluations: the per-scene box evaluations of the last
    earliest-TOI call sum to the call's total, and a scene the call did not
    serve reads 0."""
    scenes, batch, plain = setup
    subset = (0, 2, 3)
    p0s, p1s = [], []
    for k in subset:
        p0, p1 = moved(scenes[k], 40 + k)
        p0s.append(p0)
        p1s.append(p1)
    batch.batch_earliest_toi(subset, p0s, p1s)
    total = batch.stats()["box_evaluations"]
    per = batch.batch_ccd_evaluations(range(len(scenes)))
    print(per, total, flush=True)
    assert total > 0
    assert int(per.sum()) == int(total)
    for k in range(len(scenes)):
        if k not in subset:
            assert per[k] == 0


def test_single_scene_entry_points_on_a_batch_context(setup):
    scenes, batch, plain = setup
    p0 = np.concatenate([moved(sc, 60 + k)[0] for k, sc in enumerate(scenes)])
    p1 = np.concatenate([moved(sc, 60 + k)[1] for k, sc in enumerate(scenes)])
    want = min(plain[k].compute_earliest_toi(*moved(sc, 60 + k)) for k, sc in enumerate(scenes))
    assert batch.compute_earliest_toi(p0, p1) == want
    ee, fv = batch.broad_phase(p0, p1, 0.0)
    # candidates never pair bodies of different scenes
    nbs = np.cumsum([0] + [len(sc.group_ids) for sc in scenes])
    E = np.concatenate([sc.edges + int(np.sum([len(s.vertices) for s in scenes[:k]]))
                        for k, sc in enumerate(scenes)])
    vsc = np.concatenate([np.full(len(sc.vertices), k) for k, sc in enumerate(scenes)])
    assert len(ee) > 0
    assert np.all(vsc[E[ee[:, 0], 0]] == vsc[E[ee[:, 1], 0]])
    assert nbs[-1] == batch.num_bodies


def test_barrier_matches_single_scene_contexts(setup):
    scenes, batch, plain = setup
    for subset in ((1, 2, 4), tuple(range(len(scenes)))):
        build = [lowered(scenes[k], 70 + k) for k in subset]
        evalp = [b.copy() for b in build]
        for e in evalp:
            e[:-1, 1] += 1e-4
        got = batch.batch_barrier_term(subset, build, evalp, DHAT, grad=True, hess=True)
        n_con = 0
        for (e, g, h), k, pb, pe in zip(got, subset, build, evalp):
            plain[k].construct_constraint_set(pb, DHAT)
            n_con += len(plain[k].constraint_set()[0])
            we, wg, wh = plain[k].barrier_term(pe, DHAT, grad=True, hess=True)
            assert e == we
            np.testing.assert_array_equal(g, wg)
            np.testing.assert_array_equal(h[0], wh[0])
            np.testing.assert_array_equal(h[1], wh[1])
            np.testing.assert_array_equal(h[2], wh[2])
        print(subset, "constraints", n_con, flush=True)
        assert n_con > 0
        # energy only, and the per-scene constraint counts
        got_e = batch.batch_barrier_term(subset, build, evalp, DHAT)
        assert [x[0] for x in got_e] == [x[0] for x in got]
        counts = batch.batch_constraint_counts(subset, build, DHAT)
        for row, k, pb in zip(counts, subset, build):
            plain[k].construct_constraint_set(pb, DHAT)
            rec = plain[k].constraint_set()[0]
            assert list(row) == [int(np.sum(rec[:, 0] == t)) for t in range(4)]


def test_minimum_distance_matches_single_scene_contexts(setup):
    scenes, batch, plain = setup
    subset = (0, 1, 3)
    build = [lowered(scenes[k], 80 + k) for k in subset]
    got = batch.batch_minimum_distance(subset, build, build, DHAT)
    want = [plain[k].minimum_distance(pb, DHAT) for k, pb in zip(subset, build)]
    assert list(got) == want
    assert any(np.isfinite(want))


def test_barrier_call_returns_the_minimum_distance(setup):
    # the optional min_distance output of rigc_batch_barrier_term equals
    # rigc_batch_minimum_distance for the same arguments, bit for bit
    scenes, batch, plain = setup
    subset = (0, 2, 3, 4)
    build = [lowered(scenes[k], 90 + k) for k in subset]
    evalp = [b.copy() for b in build]
    for e in evalp:
        e[:-1, 1] += 2e-4
    for pe in (build, evalp):
        e, md = batch.batch_barrier_energy_and_distance(subset, build, pe, DHAT)
        want = batch.batch_minimum_distance(subset, build, pe, DHAT)
        assert list(md) == list(want)
        assert [x[0] for x in batch.batch_barrier_term(subset, build, pe, DHAT)] == list(e)
        assert any(np.isfinite(want))


def test_friction_matches_single_scene_contexts(setup):
    scenes, batch, plain = setup
    kap = {k: 10.0 ** (3 + k) for k in range(len(scenes))}
    mu = {k: 0.1 + 0.1 * k for k in range(len(scenes))}
    eh = {k: 1e-3 * scenes[k].dt for k in range(len(scenes))}
    x0 = {k: lowered(scenes[k], 90 + k) for k in range(len(scenes))}
    # friction sets of all scenes, then scene 2 alone again with other
    # parameters: the other scenes' sets must survive the second update
    allk = tuple(range(len(scenes)))
    batch.batch_update_friction_constraints(allk, [x0[k] for k in allk], DHAT, 0.0, [kap[k] for k in allk],
                                            [mu[k] for k in allk])
    kap[2] *= 3.0
    x0[2] = lowered(scenes[2], 123)
    counts = batch.batch_update_friction_constraints((2,), [x0[2]], DHAT, 0.0, [kap[2]], [mu[2]])
    for k in allk:
        n = plain[k].update_friction_constraints(x0[k], DHAT, 0.0, kap[k], mu[k])
        if k == 2:
            assert counts[0] == n
    for subset in ((0, 2, 4), allk):
        x = []
        for k in subset:
            p = x0[k].copy()
            p[:-1, 0] += 2e-3
            p[:-1, 5] += 1e-3
            x.append(p)
        got = batch.batch_friction_term(subset, x, [x0[k] for k in subset], [eh[k] for k in subset], grad=True,
                                        hess=True)
        nonzero = 0
        for (e, g, h), k, p in zip(got, subset, x):
            we, wg, wh = plain[k].friction_term(p, x0[k], eh[k], grad=True, hess=True)
            assert e == we
            np.testing.assert_array_equal(g, wg)
            for a, b in zip(h, wh):
                np.testing.assert_array_equal(a, b)
            nonzero += e != 0.0
        print(subset, "scenes with friction", nonzero, flush=True)
        assert nonzero > 0
