"""GPU intersection check against the oracle.

Bar: the sorted list of intersecting (edge, face) pairs bit-exact with the
oracle's (whose orientations are exact big-integer determinants, pinned by
Fraction arithmetic on the CPU), on interpenetrating dyadic cubes, a cube
resting exactly on a plane and a lattice of cubes whose faces touch (exact
zero orientations: the device's undecided-pair path), same-group and
separated bodies (empty), a cluster of rotated meshes pushed into each
other, and a 200-body mixed scene compressed towards its centre. The
overflow passes (body pairs, undecided pairs, results) are run from tiny
capacities and must give the default context's list.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rigid-ipc_amd"))

import oracle_lib as O  # noqa: E402
from rigidipc_amd import scenes as S  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import gpu_util
    c = gpu_util.context()
    yield c
    c.close()


def static_body(V, F, p, group=None):
    return S.Body(np.asarray(V, float), np.asarray(F), np.eye(3), np.asarray(p, float), np.zeros(3), np.zeros(3),
                  static=True, group=group)


def exact_scenes():
    V, F = S.cube(1.0)
    Vp, Fp = S.plane(8.0)
    grid = [static_body(Vp, Fp, (0, 0, 0))]
    for i in range(3):
        for j in range(3):
            grid.append(static_body(V, F, (i - 1.0, 0.5, j - 1.0)))
    return {
        "overlap": S.assemble([static_body(V, F, (0, 0.5, 0)), static_body(V, F, (0.5, 0.75, 0.25))]),
        "touching": S.assemble([static_body(Vp, Fp, (0, 0, 0)), static_body(V, F, (0.25, 0.5, 0.125))]),
        "touching_grid": S.assemble(grid),
        "apart": S.assemble([static_body(V, F, (0, 0.5, 0)), static_body(V, F, (2, 0.5, 0))]),
        "grouped": S.assemble([static_body(V, F, (0, 0, 0), 3), static_body(V, F, (0.25, 0.25, 0.25), 3)]),
    }


def compare(ctx, sc, poses):
    O.set_scene(sc)
    ctx.set_scene(sc)
    want = O.detect_intersections(poses)
    got = ctx.detect_intersections(poses)
    np.testing.assert_array_equal(got, want)
    assert ctx.count_intersections(poses) == len(want)
    return len(want)


@pytest.mark.parametrize("name", ["overlap", "touching", "touching_grid", "apart", "grouped"])
def test_exact_configurations_match_oracle(ctx, name):
    sc = exact_scenes()[name]
    n = compare(ctx, sc, sc.poses)
    print(name, "intersecting pairs", n, flush=True)
    assert (n > 0) == (name in ("overlap", "touching", "touching_grid"))


def compressed(sc, factor, seed):
    rng = np.random.default_rng(seed)
    p = sc.poses.copy()
    dyn = ~sc.is_static.astype(bool)
    c = p[dyn, :3].mean(axis=0)
    p[dyn, :3] = c + (p[dyn, :3] - c) * factor
    p[dyn, 3:] += rng.normal(size=(dyn.sum(), 3)) * 0.3
    return p


def test_rotated_cluster_matches_oracle(ctx):
    rng = np.random.default_rng(11)
    meshes = {0: S.icosphere(2, 0.5), 1: S.torus(0.5, 0.15), 2: S.cube(1.0)}
    sc = S.assemble(S.lattice_bodies(rng, 40, (5.0, 5.0, 5.0), 1.25, 0.0, 1.0, 2.0, meshes=meshes))
    assert compare(ctx, sc, sc.poses) == 0  # the lattice starts overlap-free
    n = compare(ctx, sc, compressed(sc, 0.8, 12))
    print("cluster intersecting pairs", n, flush=True)
    assert n > 100


@pytest.mark.parametrize("seed", [31, 32, 33, 34])
def test_rotated_clusters_seed_sweep(ctx, seed):
    rng = np.random.default_rng(seed)
    meshes = {0: S.icosphere(1, 0.5), 1: S.torus(0.5, 0.2, 12, 6), 2: S.cube(0.9)}
    sc = S.assemble(S.lattice_bodies(rng, 24, (4.0, 4.0, 4.0), 1.0, 0.0, 1.0, 2.0, meshes=meshes))
    n = compare(ctx, sc, compressed(sc, 0.7 + 0.05 * (seed % 4), seed + 100))
    print("seed", seed, "intersecting pairs", n, flush=True)


def test_multi_scene_context_is_the_union_of_its_scenes(ctx):
    """Scenes of one context never pair with each other: the list is each
    scene's own list, shifted by the scene's edge and face offsets."""
    rng = np.random.default_rng(41)
    meshes = {0: S.icosphere(2, 0.5), 1: S.torus(0.5, 0.15), 2: S.cube(1.0)}
    scs = [S.assemble(S.lattice_bodies(rng, 12, (4.0, 4.0, 4.0), 1.25, 0.0, 1.0, 2.0, meshes=meshes))
           for _ in range(3)]
    poses = [compressed(sc, 0.7, 50 + i) for i, sc in enumerate(scs)]
    want, e0, f0 = [], 0, 0
    for sc, p in zip(scs, poses):
        O.set_scene(sc)
        w = O.detect_intersections(p).astype(np.int64)
        want.append(w + [e0, f0])
        e0 += len(sc.edges)
        f0 += len(sc.faces)
    want = np.concatenate(want)
    want = want[np.lexsort((want[:, 1], want[:, 0]))]
    import gpu_util
    multi = gpu_util.context()
    multi.set_scenes(scs)
    got = multi.detect_intersections(np.concatenate(poses))
    multi.close()
    print("multi-scene intersecting pairs", len(got), flush=True)
    assert len(got) > 0
    np.testing.assert_array_equal(got, want)


def test_mixed_scene_compressed_matches_oracle(ctx):
    sc = S.config2(seed=3, n=200)
    n = compare(ctx, sc, compressed(sc, 0.85, 13))
    print("200-body intersecting pairs", n, "device ms", ctx.stats()["broad_phase_ms"], flush=True)
    assert n > 0


def test_overflow_passes_give_the_same_list():
    import gpu_util
    rng = np.random.default_rng(21)
    meshes = {0: S.icosphere(2, 0.5), 1: S.torus(0.5, 0.15), 2: S.cube(1.0)}
    sc = S.assemble(S.lattice_bodies(rng, 30, (5.0, 5.0, 5.0), 1.25, 0.0, 1.0, 2.0, meshes=meshes))
    p = compressed(sc, 0.75, 22)
    ref = gpu_util.context()
    ref.set_scene(sc)
    want = ref.detect_intersections(p)
    ref.close()
    assert len(want) > 64
    os.environ["RIGC_BROAD_PHASE_MIN_CAPACITY"] = "4"
    try:
        tiny = gpu_util.context()
    finally:
        del os.environ["RIGC_BROAD_PHASE_MIN_CAPACITY"]
    tiny.set_scene(sc)
    np.testing.assert_array_equal(tiny.detect_intersections(p), want)
    tiny.set_scene(exact_scenes()["touching_grid"])
    O.set_scene(exact_scenes()["touching_grid"])
    g = exact_scenes()["touching_grid"]
    np.testing.assert_array_equal(tiny.detect_intersections(g.poses), O.detect_intersections(g.poses))
    tiny.close()


def test_batch_counts_match_single_scene_contexts():
    import gpu_util
    rng = np.random.default_rng(61)
    meshes = {0: S.icosphere(2, 0.5), 1: S.torus(0.5, 0.15), 2: S.cube(1.0)}
    scs = [S.assemble(S.lattice_bodies(rng, 10, (4.0, 4.0, 4.0), 1.25, 0.0, 1.0, 2.0, meshes=meshes))
           for _ in range(4)]
    poses = [compressed(sc, 0.7, 70 + i) for i, sc in enumerate(scs)]
    multi = gpu_util.context()
    multi.set_scenes(scs)
    ids = [0, 2, 3]
    got = multi.batch_detect_intersections(ids, [poses[i] for i in ids])
    multi.close()
    want = []
    for i in ids:
        O.set_scene(scs[i])
        want.append(len(O.detect_intersections(poses[i])))
    print("batch counts", list(got), want, flush=True)
    assert list(got) == want and sum(want) > 0
