"""Config 3 (256 independent scenes of 8-512 bodies) against the oracle.

* Seven scenes served by one multi-scene context (rigc_batch_*), each
  result against the oracle run on that scene alone: the earliest TOI
  (bit for bit), the constraint counts per kind (bit for bit), the barrier
  energy and gradient (1e-10 relative) and Hessian blocks (1e-9), the lagged
  friction energy and gradient (1e-10).
* One call over all 256 scenes (earliest TOI and barrier term): every TOI in
  [0, 1] or +inf and every energy finite, each scene's result equal to a
  single-scene context's, bit for bit.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import context
from test_gpu_assembly import compare_blocks, gpu_blocks, oracle_blocks, rel_err
from test_gpu_batch import lowered, moved
from rigidipc_amd import scenes as S

pytestmark = pytest.mark.gpu

SCENE_IDS = (0, 1, 2, 6, 13, 14, 18)
DHAT = 0.05


@pytest.fixture(scope="module")
def batch():
    scenes = [S.config3_scene(i) for i in SCENE_IDS]
    c = context()
    c.set_scenes(scenes)
    print("bodies", [len(sc.group_ids) for sc in scenes], flush=True)
    yield scenes, c
    c.close()


def test_earliest_toi_matches_oracle(batch):
    scenes, c = batch
    ids = tuple(range(len(scenes)))
    p0s, p1s = zip(*[moved(sc, 300 + k) for k, sc in enumerate(scenes)])
    got = c.batch_earliest_toi(ids, list(p0s), list(p1s))
    want = []
    for sc, p0, p1 in zip(scenes, p0s, p1s):
        O.set_scene(sc)
        want.append(O.earliest_toi(p0, p1)[0])
    print("earliest tois", want, flush=True)
    assert list(got) == want
    assert sum(math.isfinite(w) for w in want) >= 3


def test_constraints_and_barrier_match_oracle(batch):
    scenes, c = batch
    ids = tuple(range(len(scenes)))
    build = [lowered(sc, 310 + k) for k, sc in enumerate(scenes)]
    evalp = [b.copy() for b in build]
    for e in evalp:
        e[:-1, 1] += 1e-4
    counts = c.batch_constraint_counts(ids, build, DHAT)
    got = c.batch_barrier_term(ids, build, evalp, DHAT, grad=True, hess=True)
    total = 0
    for k, sc in enumerate(scenes):
        nb = len(sc.group_ids)
        O.set_scene(sc)
        rec, _ = O.constraint_set(build[k], DHAT)
        assert list(counts[k]) == [int(np.sum(rec[:, 0] == t)) for t in range(4)]
        # the oracle evaluates the set built at `build` at the poses `evalp`
        oe, og, ol = O.barrier(evalp[k], DHAT, want_hess=True, n_constraints=len(rec))
        e, g, h = got[k]
        total += len(rec)
        if len(rec) == 0:
            assert e == 0.0 and not np.any(g)
            continue
        assert e == pytest.approx(oe, rel=1e-10)
        assert rel_err(np.ravel(g), np.ravel(og)) <= 1e-10
        assert compare_blocks(gpu_blocks(h), oracle_blocks(rec, ol)) <= 1e-9
    print("constraints", total, flush=True)
    assert total > 40


def test_friction_matches_oracle(batch):
    scenes, c = batch
    ids = tuple(range(len(scenes)))
    x0 = [lowered(sc, 320 + k) for k, sc in enumerate(scenes)]
    kap = [10.0 ** (3 + k % 3) for k in ids]
    mu = [0.2 + 0.05 * k for k in ids]
    eh = [1e-3 * sc.dt for sc in scenes]
    counts = c.batch_update_friction_constraints(ids, x0, DHAT, 0.0, kap, mu)
    x = []
    for p in x0:
        q = p.copy()
        q[:-1, 0] += 2e-3
        q[:-1, 5] += 1e-3
        x.append(q)
    got = c.batch_friction_term(ids, x, x0, eh, grad=True)
    nonzero = 0
    for k, sc in enumerate(scenes):
        O.set_scene(sc)
        O.constraint_set(x0[k], DHAT)
        n = len(O.friction_update(x0[k], DHAT, 0.0, kap[k], mu[k]))
        assert counts[k] == n
        oe, og = O.friction(x[k], x0[k], eh[k])[:2]
        e, g = got[k][0], got[k][1]
        if n == 0:
            assert e == 0.0
            continue
        nonzero += 1
        assert e == pytest.approx(oe, rel=1e-10)
        assert rel_err(np.ravel(g), np.ravel(og)) <= 1e-10
    assert nonzero >= 4


def test_all_256_scenes_in_one_call():
    scenes = [S.config3_scene(i) for i in range(256)]
    print("bodies", sum(len(sc.group_ids) for sc in scenes), flush=True)
    big = context()
    big.set_scenes(scenes)
    one = context()
    ids = tuple(range(256))
    p0s, p1s = zip(*[moved(sc, 400 + k) for k, sc in enumerate(scenes)])
    tois = big.batch_earliest_toi(ids, list(p0s), list(p1s))
    build = [lowered(sc, 500 + k) for k, sc in enumerate(scenes)]
    energies = [x[0] for x in big.batch_barrier_term(ids, build, build, DHAT)]
    assert all(t == math.inf or 0.0 <= t <= 1.0 for t in tois)
    assert all(math.isfinite(e) for e in energies)
    # each scene alone (a sample of 32 plus the largest)
    sample = sorted(set(range(0, 256, 8)) | {int(np.argmax([len(sc.group_ids) for sc in scenes]))})
    for k in sample:
        one.set_scene(scenes[k])
        assert one.compute_earliest_toi(p0s[k], p1s[k]) == tois[k]
        one.construct_constraint_set(build[k], DHAT)
        assert one.barrier_term(build[k], DHAT)[0] == energies[k]
    print("finite tois", int(np.sum(np.isfinite(tois))), "nonzero energies", int(np.sum(np.array(energies) > 0)),
          flush=True)
    assert np.sum(np.isfinite(tois)) > 20 and np.sum(np.array(energies) > 0) > 20
    big.close()
    one.close()
