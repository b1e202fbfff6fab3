"""Pack the reference's large rigid hash-grid fixture and record the oracle's
candidate sets on it.

Input (data file held by the reference's own tests):
  tests/data/large-rb-hashgrid/large-rb-hashgrid-000.json
  575 bodies (world-frame body-local vertices, edges, faces, pose_t0, pose_t1).
Output: tests/golden/large_rb_hashgrid.npz with the packed scene, both
poses, and the oracle's continuous candidates (inflation 0) and discrete
candidates at pose_t0 (inflation (dhat + dmin) / 2 = 5e-4).
Every body is its own group.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

SRC = "/root/reference/tests/data/large-rb-hashgrid/large-rb-hashgrid-000.json"


def pack(bodies):
    V, E, F, FE, bv, be, bf, p0, p1 = [], [], [], [], [0], [0], [0], [], []
    for b in bodies:
        nv0, ne0 = bv[-1], be[-1]
        v = np.asarray(b["vertices"], np.float64).reshape(-1, 3)
        e = np.asarray(b["edges"], np.int64).reshape(-1, 2)
        f = np.asarray(b.get("faces", []), np.int64).reshape(-1, 3)
        emap = {(min(x), max(x)): i for i, x in enumerate(e.tolist())}
        fe = [[emap[(min(t[j], t[(j + 1) % 3]), max(t[j], t[(j + 1) % 3]))] for j in range(3)] for t in f.tolist()]
        V.append(v); E.append(e + nv0); F.append(f + nv0)
        FE.append(np.asarray(fe, np.int64).reshape(-1, 3) + ne0)
        bv.append(nv0 + len(v)); be.append(ne0 + len(e)); bf.append(bf[-1] + len(f))
        p0.append(b["pose_t0"]["position"] + b["pose_t0"]["rotation"])
        p1.append(b["pose_t1"]["position"] + b["pose_t1"]["rotation"])
    nb = len(bodies)
    return dict(vertices=np.concatenate(V), edges=np.concatenate(E).astype(np.int32),
                faces=np.concatenate(F).astype(np.int32), face_edges=np.concatenate(FE).astype(np.int32),
                body_vertex_start=np.array(bv, np.int32), body_edge_start=np.array(be, np.int32),
                body_face_start=np.array(bf, np.int32), group_ids=np.arange(nb, dtype=np.int32),
                is_static=np.zeros(nb, np.uint8), poses_t0=np.array(p0), poses_t1=np.array(p1))


def main():
    import types
    with open(SRC) as fh:
        d = json.load(fh)
    a = pack(d["bodies"])
    sc = types.SimpleNamespace(**a)
    O.set_scene(sc)
    ee_c, fv_c = O.broadphase(a["poses_t0"], a["poses_t1"], 0.0)
    ee_s, fv_s = O.broadphase(a["poses_t0"], None, 5e-4)
    print("continuous", len(ee_c), len(fv_c), "discrete", len(ee_s), len(fv_s))
    np.savez_compressed(os.path.join(HERE, "large_rb_hashgrid.npz"), **a, ee_ccd=ee_c, fv_ccd=fv_c,
                        ee_static=ee_s, fv_static=fv_s)


if __name__ == "__main__":
    main()
