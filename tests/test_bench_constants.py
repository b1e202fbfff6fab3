"""Constants the benchmark prices its roofline with."""
import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_flops_per_box_matches_op_count():
    """FLOPS_PER_BOX = mean FP64 add/sub/mul/div/sqrt count of one interval
    inclusion-function evaluation, from the op-counting oracle build over
    the recorded CCD queries at random dyadic sub-boxes."""
    import bench
    so = os.path.join(ROOT, "oracle", "_build", "liboracle_flops.so")
    if not os.path.exists(so):
        pytest.skip("op-counting oracle not built")
    L = ctypes.CDLL(so)
    dp = ctypes.POINTER(ctypes.c_double)
    L.oracle_flops_get.restype = ctypes.c_long
    L.oracle_ee_aabb.argtypes = [dp, dp, dp]
    L.oracle_fv_aabb.argtypes = [dp, dp, dp]
    g = np.load(os.path.join(ROOT, "tests", "golden", "ccd_queries.npz"))
    counts = []
    out = np.zeros(6)
    rng = np.random.default_rng(0)
    for kind, q in zip(g["kind"], g["q"]):
        q = np.ascontiguousarray(q)
        for _ in range(4):
            lev = rng.integers(0, 12)
            i = rng.integers(0, 2 ** lev)
            w = 2.0 ** -lev
            box = np.array([i * w, (i + 1) * w, 0, 0.5, 0.25, 0.5])
            L.oracle_flops_reset()
            (L.oracle_ee_aabb if kind == 0 else L.oracle_fv_aabb)(
                q.ctypes.data_as(dp), box.ctypes.data_as(dp), out.ctypes.data_as(dp))
            counts.append(L.oracle_flops_get())
    assert float(np.mean(counts)) == pytest.approx(bench.FLOPS_PER_BOX, rel=1e-3)
