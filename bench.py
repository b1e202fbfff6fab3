"""Benchmark of the MI355X contact hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sim|ccd|plccd|batch|intersect|grad] [--no-cpu]

Workload "sim" (default): the 1,000-body mixed scene (ico2 / torus / cube,
seeded, plus the ground plane; dt 0.01, dhat 1e-3, mu 0.2) stepped with
the reference's implicit-Euler IPC Newton solve (host, librigc_sim.so)
over the device contact path (librigc.so: CCD in every line search,
constraint set, barrier and friction assembly, minimum distance). W
warm-up steps run on a separate simulation of the same scene; the K timed
steps start from the initial state. value = K / wall time of the K steps.
The CPU baseline runs the same stepper over the oracle (the CPU
restatement, OpenMP over the host cores) from the same initial state for a
bounded sample of steps.

The sim line also carries "batched": the batching workload (256
independent scenes of 8-512 bodies) advanced --batch-steps steps with the
fused batch and with the plain path scene after scene (same trajectories).
Workload "batch" runs that comparison alone for K steps (default all 10).

Workload "grad" (also "gradient" in the sim line): the differentiation
scene (64 spheres dropped on the plane, 20 steps, mu 0.5), gradient of the
end-centre loss with respect to all 384 initial velocities by the stepper's
implicit-function adjoint; checked against central differences on a sample
of parameters and against the same adjoint over the CPU oracle.

Workload "ccd": one device CCD call (continuous broad phase + rigid
interval narrow phase, rigc_compute_earliest_toi) per step along a
ballistic trajectory of the same scene; value = CCD candidate pairs/s.

Workload "intersect": the edge-triangle intersection check
(rigc_detect_intersections) of the same scene with its dynamic bodies
pulled towards their centroid (many interpenetrating pairs), K checks;
value = checks/s, the CPU baseline one oracle check of the same poses.
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

METRIC = "time steps/sec + CCD candidate pairs/sec at 1k bodies; batched vs plain GPU"
FP64_PEAK_TFLOPS = 78.6
# FP64 add/sub/mul/div/sqrt of one interval inclusion-function evaluation
# (rigid edge-edge / face-vertex trajectory box), counted with the
# op-counting build of the CPU restatement over the recorded CCD queries;
# tests/test_bench_constants.py recomputes it.
FLOPS_PER_BOX = 1015.7
# HBM bytes per ccd_kernel launch from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE runs (profiles/ccd_kernel_traffic_pmc.md): FETCH_SIZE scaled
# by 2.0 and WRITE_SIZE by 1.0, the factors at which the counters read a
# 256 MiB device copy as 256 MiB read and 256 MiB written.
CCD_TRAFFIC = (812.0 * 2.0 + 2185.6) * 1024   # ballistic CCD sequence
SIM_CCD_TRAFFIC = (223.9 * 2.0 + 2126.3) * 1024  # time stepping, first 3 steps


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


def ccd_roofline(evals, kernel_ms, launches, traffic=None):
    """Roofline of the CCD traversal kernel: evaluations x FLOPS_PER_BOX per
    launch over its average launch duration (HIP events recorded around each
    launch on the context stream). traffic: HBM bytes per launch from the
    FETCH_SIZE / WRITE_SIZE counters (None when not measured for the
    workload)."""
    if not kernel_ms or not evals or not launches:
        return None
    avg_s = kernel_ms * 1e-3 / launches
    achieved = evals / launches * FLOPS_PER_BOX / avg_s / 1e12
    return dict(bound="fp64", achieved=achieved, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s", frac=achieved / FP64_PEAK_TFLOPS,
                traffic=traffic, kernel="ccd_kernel", flops_per_box=FLOPS_PER_BOX, box_evaluations=int(evals),
                launches=int(launches), kernel_ms_total=kernel_ms, kernel_us_avg=1e3 * kernel_ms / launches)


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ------------------------------------------------------------------ sim
def cpu_baseline_child(budget_s):
    """Child process: the same stepper over the oracle backend, from the
    initial state. A polling thread prints the Newton iterations completed
    so far every half second (the stepper runs in C without the GIL)."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from rigidipc_amd import scenes as S
    from rigidipc_amd import sim as M
    sc = S.config2(seed=0)
    L = O.lib()
    L.oracle_backend.argtypes = [ctypes.POINTER(M.Backend)]
    L.oracle_backend.restype = None
    be = M.Backend()
    L.oracle_backend(ctypes.byref(be))
    sim = M.Simulation(sc, backend=ctypes.pointer(be))
    print("READY", flush=True)
    t0 = time.perf_counter()

    def poll():
        while True:
            st = sim.stats()
            print(f"ITER {int(st['newton_iterations'])} {int(st['steps'])} {time.perf_counter() - t0:.6f}",
                  flush=True)
            time.sleep(0.5)

    threading.Thread(target=poll, daemon=True).start()
    while time.perf_counter() - t0 < budget_s + 60:
        sim.step(1)


def cpu_baseline_sim(budget_s):
    """Runs cpu_baseline_child in a subprocess (it never touches the GPU) and
    stops it at the deadline. Returns (Newton iterations, steps, seconds) of
    its last progress report before the deadline."""
    import selectors
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--cpu-child", "--cpu-budget", str(budget_s)]
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    sel = selectors.DefaultSelector()
    sel.register(proc.stdout, selectors.EVENT_READ)
    if sel.select(timeout=300.0):  # scene set-up in the child
        proc.stdout.readline()
    t0 = time.perf_counter()
    iters, steps, elapsed = 0, 0, 0.0
    last_log = 0.0
    while time.perf_counter() - t0 < budget_s:
        if not sel.select(timeout=1.0):
            continue
        line = proc.stdout.readline()
        if not line:
            break
        if line.startswith("ITER"):
            _, it, st, t = line.split()
            if float(t) <= budget_s:
                iters, steps, elapsed = int(it), int(st), float(t)
            if elapsed - last_log >= 5.0:
                last_log = elapsed
                log(f"cpu baseline: {iters} Newton iterations, {steps} steps, {elapsed:.1f} s")
    if proc.poll() is None:
        proc.kill()
    proc.wait()
    return iters, steps, elapsed


def bench_sim(args):
    from rigidipc_amd import scenes as S
    from rigidipc_amd import sim as M
    sc = S.config2(seed=0)
    if args.warmup:
        w = M.Simulation(sc)
        for k in range(args.warmup):
            w.step(1)
            log(f"warmup step {k + 1}/{args.warmup}")
        w.close()
    sim = M.Simulation(sc)
    per_step = []
    t_all = time.perf_counter()
    for k in range(args.steps):
        t0 = time.perf_counter()
        sim.step(1)
        per_step.append(time.perf_counter() - t0)
        log(f"step {k + 1}/{args.steps}: {1e3 * per_step[-1]:.1f} ms")
    wall = time.perf_counter() - t_all
    st = sim.stats()
    sim.close()
    ccd_dev_ms = st["ccd_broad_phase_ms"] + st["ccd_narrow_phase_ms"]
    line = dict(
        metric=METRIC, value=args.steps / wall, unit="time steps/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(
            workload="config2 time stepping: 1000 mixed bodies (ico2 / torus 24x12 / cube) + static plane, "
                     "dt 0.01, dhat 1e-3, mu 0.2, implicit-Euler IPC Newton on the host, contact terms on the device",
            bodies=int(len(sc.group_ids)), vertices=int(len(sc.vertices)), faces=int(len(sc.faces)),
            newton_iterations=int(st["newton_iterations"]), line_search_iterations=int(st["line_search_iterations"]),
            ccd_calls=int(st["ccd_calls"]), constraint_builds=int(st["constraint_builds"]),
            constraints_last=int(st["num_constraints"]), failed_solves=int(st["failed_solves"]),
            contact_ms_per_step=st["contact_ms"] / args.steps, host_ms_per_step=st["host_ms"] / args.steps,
            contact_ms_by_call={k: round(st[k + "_ms"] / args.steps, 3)
                                for k in ("ccd", "constraint", "barrier", "friction", "distance")},
            ccd_device_ms_per_step={"broad_phase": round(st["ccd_broad_phase_ms"] / args.steps, 3),
                                    "narrow_phase": round(st["ccd_narrow_phase_ms"] / args.steps, 3),
                                    "ccd_kernel": round(st["ccd_kernel_ms"] / args.steps, 3)},
            ccd_candidate_pairs_per_s=(st["ccd_candidates"] / (ccd_dev_ms * 1e-3)) if ccd_dev_ms > 0 else None,
            ccd_candidate_pairs_per_s_wall=(st["ccd_candidates"] / (st["ccd_ms"] * 1e-3)) if st["ccd_ms"] > 0 else None,
            ccd_candidates=int(st["ccd_candidates"]), step_ms=[round(1e3 * t, 3) for t in per_step]),
        roofline=ccd_roofline(st["ccd_box_evaluations"], st["ccd_kernel_ms"], st["ccd_launches"], SIM_CCD_TRAFFIC),
    )
    if not args.no_cpu:
        iters, steps, t_cpu = cpu_baseline_sim(args.cpu_budget)
        gpu_iters = int(st["newton_iterations"])
        # Both paths take the same Newton iterations (same algorithm, contact
        # terms at parity), so the CPU's iteration rate over the sample is
        # scaled by the timed run's iterations per step.
        value = (iters / t_cpu) * (args.steps / gpu_iters) if iters and t_cpu > 0 and gpu_iters else None
        line["cpu_baseline"] = dict(
            value=value, unit="time steps/s", cores=cpu_threads(), kind="port",
            sample=f"same scene and initial state, same host stepper over the oracle backend (OpenMP, "
                   f"{cpu_threads()} threads), run for {t_cpu:.1f} s: {iters} Newton iterations, {steps} steps "
                   f"completed ({iters / max(t_cpu, 1e-9):.2f} iterations/s); scaled by the timed run's "
                   f"{gpu_iters} Newton iterations over {args.steps} steps (device: "
                   f"{gpu_iters / wall:.1f} iterations/s)")
    return line


# ------------------------------------------------------------------ ccd
def cpu_baseline_ccd(sc, steps, budget_s):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    O.set_scene(sc)
    t0 = time.perf_counter()
    ncand = done = 0
    for p0, p1 in steps:
        _, n = O.earliest_toi(p0, p1, 0.0)
        ncand += n
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return dict(value=ncand / dt, unit="CCD candidate pairs/s", cores=cpu_threads(), kind="port",
                sample=f"{done} CCD calls (broad + narrow) of the same sequence, OpenMP {cpu_threads()} threads; "
                       f"{done / dt:.3f} calls/s")


def bench_ccd(args):
    import rigidipc_amd
    from rigidipc_amd import scenes as S
    sc = S.config2(seed=0)
    seq = ballistic_steps(sc, args.warmup + args.steps)
    ctx = rigidipc_amd.Context(0)
    ctx.set_scene(sc)
    for p0, p1 in seq[:args.warmup]:
        ctx.compute_earliest_toi(p0, p1)
    tot_cand = tot_evals = launches = 0
    bp_ms = np_ms = k_ms = 0.0
    t0 = time.perf_counter()
    for p0, p1 in seq[args.warmup:]:
        ctx.compute_earliest_toi(p0, p1)
        st = ctx.stats()
        tot_cand += st["num_candidates"]
        tot_evals += st["box_evaluations"]
        bp_ms += st["broad_phase_ms"]
        np_ms += st["narrow_phase_ms"]
        k_ms += st["ccd_kernel_ms"]
        launches += st["ccd_launches"]
    wall = time.perf_counter() - t0
    ctx.close()
    line = dict(
        metric=METRIC, value=tot_cand / wall, unit="CCD candidate pairs/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload="config2 ballistic CCD sequence: 1000 mixed bodies + plane, one "
                             "rigc_compute_earliest_toi per step (poses_t -> inertial prediction)",
                    bodies=int(len(sc.group_ids)), candidates_per_step=tot_cand / args.steps,
                    broad_phase_ms_per_step=bp_ms / args.steps, narrow_phase_ms_per_step=np_ms / args.steps),
        roofline=ccd_roofline(tot_evals, k_ms, launches, CCD_TRAFFIC),
    )
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_ccd(sc, seq[args.warmup:], args.cpu_budget)
    return line


# ------------------------------------------------------------------ piecewise-linear CCD
# FP64 operations of one box evaluation of the linear-trajectory function:
# 8 corners x 3 axes x 19 (four lerps of the vertices in t, the two
# interpolations in u, v, one difference), plus the 3 box widths.
PL_FLOPS_PER_BOX = 8 * 3 * 19 + 3


def cpu_baseline_plccd(sc, steps, budget_s):
    """The oracle's piecewise-linear CCD, one thread, over the candidates of
    the same sequence (oracle broad phase), until the budget runs out."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from test_gpu_broadphase import query_record
    O.set_scene(sc)
    t_q = 0.0
    nq = done = 0
    for p0, p1 in steps:
        ee, fv = O.broadphase(p0, p1, 0.0)
        work = [(O.pl_ee_toi if k == 0 else O.pl_fv_toi, query_record(sc, p0, p1, k, a, b))
                for k, pairs in ((0, ee), (1, fv)) for a, b in pairs]
        t0 = time.perf_counter()
        for fn, q in work:
            fn(q)
        nq += len(work)
        t_q += time.perf_counter() - t0
        done += 1
        if t_q > budget_s:
            break
    return dict(value=nq / t_q, unit="CCD candidate pairs/s", cores=1, kind="port",
                sample=f"piecewise-linear narrow phase of the oracle over all {nq} candidates of the first {done} "
                       f"calls of the same sequence, one thread (query packing and broad phase not timed)")


def bench_plccd(args):
    import rigidipc_amd
    from rigidipc_amd import scenes as S
    sc = S.config2(seed=0)
    seq = ballistic_steps(sc, args.warmup + args.steps)
    ctx = rigidipc_amd.Context(0)
    ctx.set_scene(sc)
    for p0, p1 in seq[:args.warmup]:
        ctx.compute_earliest_toi_piecewise_linear(p0, p1)
    tot_cand = tot_evals = launches = 0
    bp_ms = k_ms = 0.0
    t0 = time.perf_counter()
    for p0, p1 in seq[args.warmup:]:
        ctx.compute_earliest_toi_piecewise_linear(p0, p1)
        st = ctx.stats()
        tot_cand += st["num_candidates"]
        tot_evals += st["box_evaluations"]
        bp_ms += st["broad_phase_ms"]
        k_ms += st["ccd_kernel_ms"]
        launches += st["ccd_launches"]
    wall = time.perf_counter() - t0
    ctx.close()
    roof = None
    if k_ms and launches:
        avg_s = k_ms * 1e-3 / launches
        achieved = tot_evals / launches * PL_FLOPS_PER_BOX / avg_s / 1e12
        roof = dict(bound="fp64", achieved=achieved, peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=achieved / FP64_PEAK_TFLOPS, traffic=None, kernel="plccd_wave_kernel",
                    flops_per_box=PL_FLOPS_PER_BOX, box_evaluations=int(tot_evals), launches=int(launches),
                    kernel_ms_total=k_ms, kernel_us_avg=1e3 * k_ms / launches)
    line = dict(
        metric=METRIC, value=tot_cand / wall, unit="CCD candidate pairs/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload="config2 ballistic CCD sequence, piecewise-linear trajectories: 1000 mixed bodies + "
                             "plane, one rigc_compute_earliest_toi_piecewise_linear per step",
                    bodies=int(len(sc.group_ids)), candidates_per_step=tot_cand / args.steps,
                    broad_phase_ms_per_step=bp_ms / args.steps, narrow_phase_ms_per_step=k_ms / args.steps),
        roofline=roof,
    )
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_plccd(sc, seq[args.warmup:], args.cpu_budget)
    return line


# ------------------------------------------------------------------ intersection check
def compressed_poses(sc, factor=0.85, seed=13):
    """Dynamic bodies pulled towards their centroid by `factor` and turned by
    N(0, 0.3) rad about each axis: many interpenetrating body pairs."""
    rng = np.random.default_rng(seed)
    p = sc.poses.copy()
    dyn = ~sc.is_static.astype(bool)
    c = p[dyn, :3].mean(axis=0)
    p[dyn, :3] = c + (p[dyn, :3] - c) * factor
    p[dyn, 3:] += rng.normal(size=(int(dyn.sum()), 3)) * 0.3
    return p


def bench_intersect(args):
    import rigidipc_amd
    from rigidipc_amd import scenes as S
    sc = S.config2(seed=0)
    p = compressed_poses(sc)
    ctx = rigidipc_amd.Context(0)
    ctx.set_scene(sc)
    for _ in range(args.warmup):
        ctx.detect_intersections(p)
    dev_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        pairs = ctx.detect_intersections(p)
        dev_ms += ctx.stats()["broad_phase_ms"]
    wall = time.perf_counter() - t0
    ctx.close()
    line = dict(
        metric=METRIC, value=args.steps / wall, unit="intersection checks/s", n_gpus=1, steps=args.steps,
        warmup=args.warmup, ms_per_step=1e3 * wall / args.steps, higher_is_better=True, scaling="weak",
        vs_baseline=None, dtype="f64", data="synthetic",
        config=dict(workload="intersection check (rigc_detect_intersections, every intersecting edge-face pair, "
                             "sorted) of the 1000-body mixed scene with the dynamic bodies pulled to 0.85 of their "
                             "distance from the centroid and turned by N(0, 0.3) rad",
                    bodies=int(len(sc.group_ids)), intersecting_pairs=int(len(pairs)),
                    device_ms_per_check=dev_ms / args.steps),
    )
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as O
        O.set_scene(sc)
        t = time.perf_counter()
        n = len(O.detect_intersections(p))
        dt = time.perf_counter() - t
        line["cpu_baseline"] = dict(value=1.0 / dt, unit="intersection checks/s", cores=1, kind="port",
                                    sample=f"one oracle check of the same poses (single thread, exact big-integer "
                                           f"orientations): {dt:.2f} s, {n} intersecting pairs")
    return line


# ------------------------------------------------------------------ gradient
def grad_compare(fd_params=6, cpu=True):
    """The differentiation workload: 64 ico1 spheres (r 0.3) dropped from a
    jittered 4^3 lattice onto the plane, 20 steps at dt 0.01, mu 0.5; loss =
    sum over spheres of |end centre - target|^2; gradient with respect to
    all 384 initial linear and angular velocities by the stepper's
    implicit-function adjoint (rigs_rollout_backward, device contact path).
    Reports forward (checkpointed rollout) and backward times, the gradient's
    agreement with central differences (h 1e-6) on fd_params parameters, and
    the same adjoint over the CPU oracle (time and agreement)."""
    from rigidipc_amd import diff as Dm
    from rigidipc_amd import scenes as S
    from rigidipc_amd import sim as M
    sc = S.config4(0)
    sim = M.Simulation(sc, params=Dm.params_for_gradients(sc))

    def loss(v, s=sim):
        s.set_state(sc.poses, v)
        s.rollout(sc.steps)
        x, _ = s.state()
        d = x[:64, :3] - sc.targets
        g = np.zeros_like(x)
        g[:64, :3] = 2 * d
        return float((d * d).sum()), g

    loss(sc.velocities)  # warm-up
    t0 = time.perf_counter()
    L, gx = loss(sc.velocities)
    t_fwd = time.perf_counter() - t0
    st = sim.stats()
    t0 = time.perf_counter()
    g = sim.rollout_backward(gx)
    t_bwd = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    picks = [(int(b), int(k)) for b, k in zip(rng.integers(0, 64, fd_params), rng.integers(0, 6, fd_params))]
    fd = []
    for b, k in picks:
        vp, vm = sc.velocities.copy(), sc.velocities.copy()
        vp[b, k] += 1e-6
        vm[b, k] -= 1e-6
        fd.append((loss(vp)[0] - loss(vm)[0]) / 2e-6)
    ad = np.array([g[b, k] for b, k in picks])
    fd = np.array(fd)
    out = dict(workload="config4: 64 ico1 + plane, 20 steps, mu 0.5, 384 parameters (initial v, omega)",
               loss=L, forward_ms=1e3 * t_fwd, backward_ms=1e3 * t_bwd, gradient_ms=1e3 * (t_fwd + t_bwd),
               newton_iterations=int(st["newton_iterations"]), contacts_at_end=int(st["num_constraints"]),
               fd_params=picks, fd_max_rel_err=float(np.abs(ad - fd).max() / np.abs(fd).max()))
    sim.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from stepper_util import oracle_backend
        orc = M.Simulation(sc, params=Dm.params_for_gradients(sc), backend=oracle_backend())
        t0 = time.perf_counter()
        Lo, gxo = loss(sc.velocities, orc)
        go = orc.rollout_backward(gxo)
        t_cpu = time.perf_counter() - t0
        orc.close()
        out["cpu_oracle"] = dict(gradient_ms=1e3 * t_cpu, cores=cpu_threads(),
                                 max_rel_diff=float(np.abs(g - go).max() / np.abs(go).max()))
    return out


# ------------------------------------------------------------------ batch
def batch_compare(n_scenes, steps, warm=True):
    """The batching workload: n_scenes independent scenes (body counts
    log-uniform in [8, 512], mix as the 1,000-body scene, dt 0.005 / 0.01,
    dhat 1e-3), each advanced `steps` steps from its initial state, once with
    the fused batch (one multi-scene context, every scene's Newton loop on a
    host thread, contact calls served by fused device calls per round) and
    once scene after scene on the plain single-scene path. Both produce the
    same trajectories bit for bit; returns aggregate scene-steps/s of each."""
    from rigidipc_amd import scenes as S
    from rigidipc_amd import sim as M
    scenes = [S.config3_scene(i) for i in range(n_scenes)]
    if warm:
        # device code and allocator warm-up on a small batch
        w = M.Batch(scenes[:4], mode=M.BATCH_FUSED)
        w.step(1)
        w.close()
    out = dict(scenes=n_scenes, steps=steps, bodies=int(sum(len(sc.group_ids) for sc in scenes)))
    for name, mode in (("fused", M.BATCH_FUSED), ("plain", M.BATCH_SEQUENTIAL)):
        b = M.Batch(scenes, mode=mode)
        t0 = time.perf_counter()
        for k in range(steps):
            t = time.perf_counter()
            b.step(1)
            log(f"batch {name} step {k + 1}/{steps}: {time.perf_counter() - t:.2f} s")
        wall = time.perf_counter() - t0
        newton = sum(b.stats(i)["newton_iterations"] for i in range(n_scenes))
        out[name] = dict(scene_steps_per_s=n_scenes * steps / wall, wall_s=wall, newton_iterations=int(newton))
        if mode == M.BATCH_FUSED:
            bs = b.batch_stats()
            out[name].update(rounds=int(bs["rounds"]), fused_calls=int(bs["fused_calls"]),
                             requests=int(bs["requests"]))
        b.close()
    out["fused_over_plain"] = out["fused"]["scene_steps_per_s"] / out["plain"]["scene_steps_per_s"]
    return out


def bench_batch(args):
    r = batch_compare(256, args.steps)
    return dict(metric=METRIC, value=r["fused"]["scene_steps_per_s"], unit="scene time steps/s (batched)", n_gpus=1,
                steps=args.steps, warmup=args.warmup, ms_per_step=1e3 * r["fused"]["wall_s"] / args.steps,
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload="256 independent scenes (8-512 bodies each, log-uniform; mixed ico2 / torus / "
                                     "cube + plane; dt 0.005 or 0.01; dhat 1e-3; mu 0.2), all advanced together",
                            **r))


def bench_grad(args):
    r = grad_compare(cpu=not args.no_cpu)
    line = dict(metric=METRIC, value=1e3 / r["gradient_ms"], unit="gradients/s (384-parameter rollout adjoint)",
                n_gpus=1, steps=20, warmup=1, ms_per_step=r["gradient_ms"] / 20, higher_is_better=True,
                scaling="weak", vs_baseline=None, dtype="f64", data="synthetic", config=r)
    if "cpu_oracle" in r:
        line["cpu_baseline"] = dict(value=1e3 / r["cpu_oracle"]["gradient_ms"], unit="gradients/s",
                                    cores=r["cpu_oracle"]["cores"], kind="port",
                                    sample="one forward rollout and adjoint over the CPU oracle")
    return line


def bench_copy(args):
    """Calibration for the PMC traffic counters: device-to-device copies of a
    known size (reads and writes 256 MiB each per launch)."""
    import torch
    n = 32 * 1024 * 1024  # doubles: 256 MiB
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(args.warmup):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nbytes = 2 * 8 * n
    return dict(metric="device copy bandwidth (PMC calibration)", value=nbytes * args.steps / wall / 1e9,
                unit="GB/s", n_gpus=1, steps=args.steps, warmup=args.warmup, ms_per_step=1e3 * wall / args.steps,
                higher_is_better=True, scaling="weak", vs_baseline=None, dtype="f64", data="synthetic",
                config=dict(workload=f"torch copy_ of {8 * n} bytes (read) into {8 * n} bytes (write)"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", default="sim", choices=["sim", "ccd", "plccd", "batch", "copy", "intersect", "grad"])
    ap.add_argument("--batch-steps", type=int, default=3,
                    help="steps of the batched-vs-plain comparison added to the sim line (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    ap.add_argument("--cpu-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_baseline_child(args.cpu_budget)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not part of the build)", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {"sim": 8, "batch": 10}.get(args.workload, 40)
    if args.warmup is None:
        args.warmup = 1 if args.workload in ("sim", "batch") else 3
    if args.workload == "copy":
        line = bench_copy(args)
    elif args.workload == "batch":
        line = bench_batch(args)
    elif args.workload == "ccd":
        line = bench_ccd(args)
    elif args.workload == "plccd":
        line = bench_plccd(args)
    elif args.workload == "intersect":
        line = bench_intersect(args)
    elif args.workload == "grad":
        line = bench_grad(args)
    else:
        line = bench_sim(args)
        if args.batch_steps > 0:
            line["batched"] = batch_compare(256, args.batch_steps)
            line["batched"]["note"] = (f"first {args.batch_steps} of the batching workload's 10 steps; "
                                       "bench.py --workload batch runs all 10")
        line["gradient"] = grad_compare(cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
