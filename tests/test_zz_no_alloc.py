"""rigc_narrow_phase_noalloc (every per-call buffer in the library's static
device array) against rigc_narrow_phase on the same inputs: per-query TOIs
and hit flags bit for bit, the earliest TOI with and without cross-query
pruning; two calls in a row, the second with a smaller input (the array
keeps the first call's contents); and a request larger than the array is
refused with RIGC_ERR_CAPACITY before any device work."""
import math

import numpy as np
import pytest

from gpu_util import PKG  # noqa: F401  (puts the package on sys.path)
from test_gpu_broadphase import cluster, moved
from rigidipc_amd import RIGC_ERR_CAPACITY, RigcError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import gpu_util
    c = gpu_util.context()
    yield c
    c.close()


def case(ctx, seed, n, dpos, drot):
    sc = cluster(seed, n=n)
    p0, p1 = moved(sc, seed + 30, dpos=dpos, drot=drot)
    ctx.set_scene(sc)
    ee, fv = ctx.broad_phase(p0, p1, 0.0)
    return sc, p0, p1, ee, fv


def compare(ctx, p0, p1, ee, fv):
    """Per query (no pruning): every TOI and flag. With pruning, which
    queries are cut short depends on the schedule; the earliest TOI does
    not, and it is the least of the reported hits."""
    want = ctx.narrow_phase(p0, p1, ee, fv, per_query=True)
    got = ctx.narrow_phase_noalloc(p0, p1, ee, fv, per_query=True)
    assert got[0] == want[0]
    np.testing.assert_array_equal(got[2], want[2])
    np.testing.assert_array_equal(got[1][got[2] == 1], want[1][want[2] == 1])
    pruned = ctx.narrow_phase_noalloc(p0, p1, ee, fv, per_query=False)
    assert pruned[0] == want[0] == ctx.narrow_phase(p0, p1, ee, fv, per_query=False)[0]
    h = pruned[2] == 1
    assert (not h.any() and want[0] == math.inf) or pruned[1][h].min() == want[0]
    return want[0]


def test_two_calls_in_a_row_match_the_allocating_entry_point(ctx):
    _, p0, p1, ee, fv = case(ctx, 0, 27, 0.05, 0.1)
    print(f"first call: {len(ee)} ee + {len(fv)} fv", flush=True)
    assert len(ee) + len(fv) > 1000
    t1 = compare(ctx, p0, p1, ee, fv)
    # second call: a smaller input on the same scene
    ee2, fv2 = ee[: len(ee) // 3], fv[: len(fv) // 5]
    print(f"second call: {len(ee2)} ee + {len(fv2)} fv", flush=True)
    t2 = compare(ctx, p0, p1, ee2, fv2)
    assert math.isfinite(t1) and t2 >= t1


def test_smaller_scene_after_a_larger_one(ctx):
    _, p0, p1, ee, fv = case(ctx, 1, 27, 0.05, 0.1)
    compare(ctx, p0, p1, ee, fv)
    _, p0, p1, ee, fv = case(ctx, 2, 8, 0.1, 0.3)
    print(f"smaller scene: {len(ee)} ee + {len(fv)} fv", flush=True)
    compare(ctx, p0, p1, ee, fv)
    # no candidates at all
    e0 = np.zeros((0, 2), np.int32)
    assert ctx.narrow_phase_noalloc(p0, p1, e0, e0)[0] == math.inf


def test_request_beyond_the_array_is_refused(ctx):
    sc, p0, p1, ee, fv = case(ctx, 3, 8, 0.05, 0.1)
    n = (40 << 20) // 41  # about 40 MiB of per-query buffers
    big = np.tile(ee[:1], (n, 1)) if len(ee) else np.zeros((n, 2), np.int32)
    with pytest.raises(RigcError) as e:
        ctx.narrow_phase_noalloc(p0, p1, big, np.zeros((0, 2), np.int32))
    assert e.value.status == RIGC_ERR_CAPACITY
    compare(ctx, p0, p1, ee, fv)  # the context still works
