"""Batched stepping (rigs_batch_*): the fused mode, where every scene's
Newton loop runs on its own host thread and their contact calls are served
by fused multi-scene device calls, must reproduce, bit for bit, the plain
path that steps each scene alone with its own single-scene context (same
poses, velocities, Newton and line-search iteration counts)."""
import numpy as np
import pytest

from gpu_util import rigidipc_amd  # noqa: F401  (puts the package on the path)
from rigidipc_amd import scenes as S
from rigidipc_amd import sim as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ids,steps", [((3, 4, 5, 7, 9, 10), 4), ((0, 2, 8, 12), 3)])
def test_fused_batch_matches_sequential(ids, steps):
    scenes = [S.config3_scene(i) for i in ids]
    print("bodies", [len(sc.group_ids) for sc in scenes], flush=True)
    seq = M.Batch(scenes, mode=M.BATCH_SEQUENTIAL)
    fused = M.Batch(scenes, mode=M.BATCH_FUSED)
    for k in range(steps):
        seq.step(1)
        fused.step(1)
        print("step", k + 1, flush=True)
    bs = fused.batch_stats()
    print("rounds", bs["rounds"], "fused calls", bs["fused_calls"], "requests", bs["requests"], flush=True)
    newton = 0
    for i in range(len(scenes)):
        p0, v0 = seq.state(i)
        p1, v1 = fused.state(i)
        np.testing.assert_array_equal(p1, p0)
        np.testing.assert_array_equal(v1, v0)
        s0, s1 = seq.stats(i), fused.stats(i)
        for key in ("newton_iterations", "line_search_iterations", "ccd_calls", "failed_solves"):
            assert s1[key] == s0[key], (i, key)
        newton += s0["newton_iterations"]
    # several scenes shared each round
    assert bs["requests"] > bs["rounds"]
    assert newton > 0
    seq.close()
    fused.close()


def test_single_scene_stepper_with_helpers_matches_fused():
    # From 200 bodies up the single-scene stepper computes the body energy
    # on a helper thread and the friction term on a second device context
    # beside the barrier term; the fused batch runs the same scene without
    # either. Both must give the same trajectory bit for bit.
    sc = S.config3_scene(13)
    assert len(sc.group_ids) >= 200
    single = M.Simulation(sc)
    fused = M.Batch([sc], mode=M.BATCH_FUSED)
    for k in range(3):
        single.step(1)
        fused.step(1)
        print("step", k + 1, flush=True)
    p0, v0 = single.state()
    p1, v1 = fused.state(0)
    np.testing.assert_array_equal(p1, p0)
    np.testing.assert_array_equal(v1, v0)
    s0, s1 = single.stats(), fused.stats(0)
    for key in ("newton_iterations", "line_search_iterations", "ccd_calls", "failed_solves"):
        assert s1[key] == s0[key], key
    assert s0["newton_iterations"] > 0
    single.close()
    fused.close()


def test_fused_batch_intersection_checks_match_sequential():
    """do_intersection_check in both batch modes: the fused proxies serve the
    check in rounds (and directly for the initial state, before any round),
    with the counts of the plain path; none of the steps ends intersecting."""
    scenes = [S.config3_scene(i) for i in (1, 4, 6)]
    params = []
    for sc in scenes:
        p = M.default_params(sc)
        p.do_intersection_check = 1
        params.append(p)
    seq = M.Batch(scenes, mode=M.BATCH_SEQUENTIAL, params=params)
    fused = M.Batch(scenes, mode=M.BATCH_FUSED, params=params)
    seq.step(2)
    fused.step(2)
    for i in range(len(scenes)):
        s0, s1 = seq.stats(i), fused.stats(i)
        for key in ("initial_intersections", "intersections", "intersecting_steps"):
            assert s1[key] == s0[key] == 0, (i, key, s0[key], s1[key])
        np.testing.assert_array_equal(fused.state(i)[0], seq.state(i)[0])
    seq.close()
    fused.close()


def test_fused_batch_counts_an_intersecting_initial_state():
    import oracle_lib as O
    from test_host_stepper import intersecting_start_scene
    scenes = [intersecting_start_scene(), S.config3_scene(2), intersecting_start_scene()]
    fused = M.Batch(scenes, mode=M.BATCH_FUSED)
    got = [fused.stats(i)["initial_intersections"] for i in range(len(scenes))]
    fused.close()
    want = []
    for sc in scenes:
        O.set_scene(sc)
        want.append(len(O.detect_intersections(sc.poses)))
    print("initial intersections", got, want, flush=True)
    assert got == want and want[0] > 0 and want[1] == 0
