"""The BVH broad phase on the device (rigc_set_detection_method(BVH): the
reference's default detection_method) against the oracle's restatement
(oracle/oracle_bvh.cpp, pinned in test_oracle_bvh.py).

* Candidate lists bit for bit (EE pairs in the reference's orientation, an
  edge of the body with fewer faces first; FV pairs), discrete and
  continuous, with and without inflation, on packed clusters of spheres,
  tori and cubes and on the reference's 575-body scene.
* The earliest TOI over the BVH's candidates, bit for bit, for rigid and
  piecewise-linear trajectories, with a separation distance, and per scene
  on a multi-scene context.
* Overflow passes (tiny starting capacities) give the same lists; a scene
  with a vertex on no edge is refused.
"""
import os
import sys
import types

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


@pytest.fixture(autouse=True)
def oracle_hash_grid_after():
    yield
    O.set_detection_method("hash_grid")


def compare(ctx, sc, p0, p1, r):
    O.set_scene(sc)
    O.set_detection_method("bvh")
    ctx.set_scene(sc)
    ctx.set_detection_method("bvh")
    ee_o, fv_o = O.broadphase(p0, p1, r)
    ee_g, fv_g = ctx.broad_phase(p0, p1, r)
    print(f"{'continuous' if p1 is not None else 'discrete'} r {r}: ee {len(ee_o)} fv {len(fv_o)}", flush=True)
    assert len(ee_o) + len(fv_o) > 0
    np.testing.assert_array_equal(ee_g, ee_o)
    np.testing.assert_array_equal(fv_g, fv_o)


@pytest.mark.parametrize("seed,r", [(0, 0.0), (1, 1e-3), (2, 0.02)])
def test_cluster_candidates_bit_exact(ctx, seed, r):
    sc = cluster(seed, n=40)
    p0, p1 = moved(sc, 60 + seed, dpos=0.05, drot=0.2)
    compare(ctx, sc, p1, None, r)  # discrete at the moved poses (bodies pushed together)
    compare(ctx, sc, p0, p1, r)


def test_reference_575_body_scene(ctx):
    g = np.load(os.path.join(HERE, "golden", "large_rb_hashgrid.npz"))
    sc = types.SimpleNamespace(**{k: g[k] for k in g.files})
    compare(ctx, sc, sc.poses_t0, sc.poses_t1, 0.0)
    compare(ctx, sc, sc.poses_t0, None, 5e-4)


@pytest.mark.parametrize("seed,msd", [(3, 0.0), (4, 0.0), (3, 1e-4)])
def test_earliest_toi_bit_exact(ctx, seed, msd):
    sc = cluster(seed, n=27)
    p0, p1 = moved(sc, seed + 30, dpos=0.05, drot=0.1)
    O.set_scene(sc)
    O.set_eval_budget(0)
    O.set_detection_method("bvh")
    want, n = O.earliest_toi(p0, p1, msd)
    ctx.set_scene(sc)
    ctx.set_detection_method("bvh")
    got = ctx.compute_earliest_toi(p0, p1, msd)
    print(f"seed {seed} msd {msd}: {n} candidates, earliest toi {want}", flush=True)
    assert np.isfinite(want)
    assert got == want


def test_piecewise_linear_earliest_toi_over_bvh_candidates(ctx):
    import math
    from test_gpu_broadphase import query_record
    sc = cluster(5, n=27)
    p0, p1 = moved(sc, 35, dpos=0.05, drot=0.1)
    O.set_scene(sc)
    O.set_detection_method("bvh")
    ee, fv = O.broadphase(p0, p1, 0.0)
    want = math.inf
    for k, pairs in ((0, ee), (1, fv)):
        for a, b in pairs:
            hit, toi, _, _ = (O.pl_ee_toi if k == 0 else O.pl_fv_toi)(query_record(sc, p0, p1, k, a, b), 1.0, 0.0)
            if hit:
                want = min(want, toi)
    ctx.set_scene(sc)
    ctx.set_detection_method("bvh")
    assert ctx.compute_earliest_toi_piecewise_linear(p0, p1, 0.0) == want


def test_multi_scene_context_per_scene(ctx):
    import gpu_util
    from rigidipc_amd import scenes as S
    scenes = [S.config3_scene(i) for i in (0, 3, 5)]
    batch = gpu_util.context()
    batch.set_scenes(scenes)
    batch.set_detection_method("bvh")
    p0s, p1s, want = [], [], []
    for k, sc in enumerate(scenes):
        p0, p1 = moved(sc, 80 + k, dpos=0.05, drot=0.1)
        p1[~sc.is_static.astype(bool), 1] -= 0.03  # the bottom layer reaches the plane
        p0s.append(p0)
        p1s.append(p1)
        ctx.set_scene(sc)
        ctx.set_detection_method("bvh")
        want.append(ctx.compute_earliest_toi(p0, p1))
    got = batch.batch_earliest_toi((0, 1, 2), p0s, p1s)
    batch.close()
    print("per scene", want, flush=True)
    assert list(got) == want
    assert any(np.isfinite(want))


def test_overflow_passes_give_the_same_lists():
    import gpu_util
    sc = cluster(6, n=40)
    p0, p1 = moved(sc, 66, dpos=0.05, drot=0.2)
    ref = gpu_util.context()
    ref.set_scene(sc)
    ref.set_detection_method("bvh")
    want = ref.broad_phase(p0, p1, 1e-3)
    want_toi = ref.compute_earliest_toi(p0, p1)
    ref.close()
    os.environ["RIGC_BROAD_PHASE_MIN_CAPACITY"] = "4"
    try:
        tiny = gpu_util.context()
    finally:
        del os.environ["RIGC_BROAD_PHASE_MIN_CAPACITY"]
    tiny.set_scene(sc)
    tiny.set_detection_method("bvh")
    got = tiny.broad_phase(p0, p1, 1e-3)
    got_toi = tiny.compute_earliest_toi(p0, p1)
    tiny.close()
    assert len(want[0]) + len(want[1]) > 100
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    assert got_toi == want_toi


def test_codimensional_vertex_is_refused(ctx):
    import rigidipc_amd as R
    from rigidipc_amd import scenes as S
    sc = cluster(0, n=4)
    # one more vertex in the last body, on no edge
    sc2 = types.SimpleNamespace(**vars(sc))
    sc2.vertices = np.concatenate([sc.vertices, [[0.0, 0.0, 0.0]]])
    sc2.body_vertex_start = sc.body_vertex_start.copy()
    sc2.body_vertex_start[-1] += 1
    ctx.set_scene(sc2)
    with pytest.raises(R.RigcError):
        ctx.set_detection_method("bvh")
    ctx.set_scene(sc)
    ctx.set_detection_method("bvh")
    ctx.set_detection_method("hash_grid")
