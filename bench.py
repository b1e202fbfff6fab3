"""Benchmark of the MI355X contact hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ccd]

Workload "ccd": the 1,000-body mixed scene (config 2: ico2 / torus / cube,
seeded, plus the ground plane). Each step is one call of the device CCD path
(continuous broad phase + rigid interval narrow phase with cross-query
pruning, rigc_compute_earliest_toi) from the poses at step s to the inertial
prediction of step s+1 (x + h v + h^2 g, r + h w), along a ballistic
trajectory. Prints one JSON line (see the keys below). The CPU baseline
is the oracle (CPU restatement of the same path, OpenMP over all host
cores) on the first steps of the same sequence.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rigid-ipc_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "time steps/sec + CCD candidate pairs/sec at 1k bodies; batched vs plain GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ballistic_steps(sc, n, dt=0.01):
    g = np.asarray(sc.gravity, float)
    dyn = ~sc.is_static.astype(bool)
    p = sc.poses.copy()
    v = sc.velocities.copy()
    out = []
    for _ in range(n):
        p1 = p.copy()
        p1[dyn, :3] += dt * v[dyn, :3] + dt * dt * g
        p1[dyn, 3:] += dt * v[dyn, 3:]
        out.append((p.copy(), p1.copy()))
        p = p1
        v[dyn, :3] += dt * g
    return out


def flops_per_box():
    """FP64 add/sub/mul/div/sqrt of one inclusion-function evaluation,
    counted by the op-counting oracle build over the reference's recorded
    queries' start boxes (EE and FV averaged by their share)."""
    so = os.path.join(ROOT, "oracle", "_build", "liboracle_flops.so")
    if not os.path.exists(so):
        return None
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
    return float(np.mean(counts))


def cpu_baseline(sc, steps, budget_s=20.0):
    import oracle_lib as O
    O.set_scene(sc)
    t0 = time.perf_counter()
    ncand = 0
    done = 0
    for p0, p1 in steps:
        _, n = O.earliest_toi(p0, p1, 0.0)
        ncand += n
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(value=ncand / dt, unit="CCD candidate pairs/s", cores=os.cpu_count(), kind="port",
                sample=f"{done} CCD calls (broad + narrow) of the same sequence, OpenMP "
                       f"{os.environ.get('OMP_NUM_THREADS', os.cpu_count())} threads; "
                       f"{done / dt:.3f} calls/s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="ccd", choices=["ccd"])
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the scaling design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import rigidipc_amd
    from rigidipc_amd import scenes as S

    sc = S.config2(seed=0)
    seq = ballistic_steps(sc, args.warmup + args.steps)
    ctx = rigidipc_amd.Context(0)
    ctx.set_scene(sc)
    for p0, p1 in seq[:args.warmup]:
        ctx.compute_earliest_toi(p0, p1)
    tot_cand = 0
    tot_evals = 0
    bp_ms = np_ms = 0.0
    t0 = time.perf_counter()
    for p0, p1 in seq[args.warmup:]:
        ctx.compute_earliest_toi(p0, p1)
        st = ctx.stats()
        tot_cand += st["num_candidates"]
        tot_evals += st["box_evaluations"]
        bp_ms += st["broad_phase_ms"]
        np_ms += st["narrow_phase_ms"]
    wall = time.perf_counter() - t0
    ms = 1e3 * wall / args.steps
    fbox = flops_per_box()
    roof = None
    if fbox and np_ms > 0:
        achieved = tot_evals * fbox / (np_ms * 1e-3) / 1e12
        roof = dict(bound="fp64", achieved=achieved, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=achieved / FP64_PEAK_TFLOPS, traffic=None, kernel="ccd_kernel",
                    flops_per_box=fbox, box_evaluations=tot_evals, kernel_ms_total=np_ms)
    line = dict(
        metric=METRIC, value=tot_cand / wall, unit="CCD candidate pairs/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=ms, higher_is_better=True, scaling="weak", vs_baseline=None,
        dtype="f64", data="synthetic",
        config=dict(workload="config2 ballistic CCD sequence: 1000 mixed bodies + plane, one "
                             "rigc_compute_earliest_toi per step (poses_t -> inertial prediction)",
                    bodies=int(len(sc.group_ids)), candidates_per_step=tot_cand / args.steps,
                    broad_phase_ms_per_step=bp_ms / args.steps, narrow_phase_ms_per_step=np_ms / args.steps),
        roofline=roof,
    )
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, seq[args.warmup:])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
