"""Convert the reference's recorded rigid CCD queries into a compact fixture
and record the oracle's per-query results.

Inputs (data files held by the reference's own tests):
  tests/data/wrecking-ball/ccd-test-000..385.json   (EE and FV)
  tests/data/kinematic/ccd-test-000..012.json       (EE)
  tests/data/ccd-test-000..003.json                 (EE and FV)
Format written by src/ccd/save_queries.cpp:40-185.

Output: tests/golden/ccd_queries.npz with
  kind  (n,) int8     0 = EE, 1 = FV
  q     (n, 36) f64   4 body-local vertices, poses A_t0, A_t1, B_t0, B_t1
  hit   (n,) int8     oracle result on t in [0, 1], toi_tol = 1e-4;
                      -1 = more than EVAL_BUDGET inclusion evaluations
  toi   (n,) f64
  evals (n,) int64    inclusion-function evaluations the oracle made
"""
import glob
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

REF = "/root/reference/tests/data"
EVAL_BUDGET = 2_000_000


def pose(p):
    return list(p["position"]) + list(p["rotation"])


def load(path):
    with open(path) as f:
        d = json.load(f)
    if d["type"] == "ee":
        e0, e1 = d["edge0"], d["edge1"]
        q = e0["vertex0"] + e0["vertex1"] + e1["vertex0"] + e1["vertex1"]
        q += pose(e0["pose_t0"]) + pose(e0["pose_t1"]) + pose(e1["pose_t0"]) + pose(e1["pose_t1"])
        return 0, q
    if d["type"] == "fv":
        v, f = d["vertex"], d["face"]
        q = v["vertex"] + f["vertex0"] + f["vertex1"] + f["vertex2"]
        q += pose(v["pose_t0"]) + pose(v["pose_t1"]) + pose(f["pose_t0"]) + pose(f["pose_t1"])
        return 1, q
    raise ValueError(d["type"])


def main():
    files = sorted(glob.glob(f"{REF}/wrecking-ball/ccd-test-*.json"))
    files += sorted(glob.glob(f"{REF}/kinematic/ccd-test-*.json"))
    files += sorted(glob.glob(f"{REF}/ccd-test-*.json"))
    kinds, qs, hits, tois, evals, names = [], [], [], [], [], []
    t0 = time.time()
    O.set_eval_budget(EVAL_BUDGET)
    for path in files:
        k, q = load(path)
        q = np.asarray(q, dtype=np.float64)
        assert q.shape == (36,)
        ts = time.time()
        hit, toi, n = (O.ee_toi if k == 0 else O.fv_toi)(q)
        dt = time.time() - ts
        if dt > 1.0:
            print(f"slow query {path}: {dt:.2f}s evals={n}")
        kinds.append(k); qs.append(q); hits.append(int(hit)); tois.append(toi); evals.append(n)
        names.append(os.path.relpath(path, REF))
    print(f"{len(files)} queries in {time.time() - t0:.1f}s; hits={hits.count(1)} "
          f"unresolved={hits.count(-1)}; "
          f"evals total={sum(evals)} max={max(evals)}")
    np.savez_compressed(os.path.join(HERE, "ccd_queries.npz"), kind=np.array(kinds, np.int8),
                        q=np.array(qs), hit=np.array(hits, np.int8), toi=np.array(tois),
                        evals=np.array(evals, np.int64), names=np.array(names))


if __name__ == "__main__":
    main()
