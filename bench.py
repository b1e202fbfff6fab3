"""Benchmark of the MI355X contact hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sim|ccd|plccd|batch|intersect|grad] [--no-cpu]

Workload "sim" (default): the 1,000-body mixed scene (ico2 / torus / cube,
seeded, plus the ground plane; dt 0.01, dhat 1e-3, mu 0.2) stepped with
the reference's implicit-Euler IPC Newton solve (host, librigc_sim.so)
over the device contact path (librigc.so: CCD in every line search,
constraint set, barrier and friction assembly, minimum distance). W
warm-up steps run on a separate simulation of the same scene; the K timed
steps start from the initial state. value = K / wall time of the K steps.
The CPU baseline replays a uniform sample of the run's contact calls (over
all K steps) on the -O3 oracle, the CPU restatement, at all host cores
and at one (see cpu_estimate); "config1" runs the 10-body scene's 100
steps outright on the device and on the oracle.

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
CCD_TRAFFIC = (787.5 * 2.0 + 98.3) * 1024   # ballistic CCD sequence
SIM_CCD_TRAFFIC = (465.0 * 2.0 + 46.3) * 1024  # time stepping, first 3 steps
HBM_PEAK_GBS = 8000.0
# Algorithmic FP64 work of one barrier constraint's Hessian block: the
# vertex-space value / gradient / Hessian with the rotation Jacobians
# (about 2,000, counted from the kernel's code), J^T H J over 12 dofs
# (2 x 12^3 x 2 = 6,912) and a 12 x 12 symmetric eigendecomposition with
# vectors at the textbook 9 n^3 = 15,552 (tridiagonal QR; the device's
# parallel Jacobi does more).
ASSEMBLY_FLOP_PER_CONSTRAINT = 2000 + 6912 + 9 * 12 ** 3


def broad_phase_bytes(sc, calls, candidates):
    """Algorithmic HBM bytes of `calls` continuous broad phases of sc that
    emitted `candidates` pairs in all: per call 96 B of poses per body,
    the mesh (24 B a vertex, 8 an edge, 12 a face), each element's box
    written and read once (96 B) and keyed into the grid (24 B); 8 B per
    candidate pair written."""
    nb, nv, ne, nf = len(sc.group_ids), len(sc.vertices), len(sc.edges), len(sc.faces)
    per_call = 96 * nb + 24 * nv + 8 * ne + 12 * nf + (96 + 24) * (nv + ne + nf)
    return calls * per_call + 8 * candidates


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


def host_cpu():
    """nproc, the usable cores, and lscpu's model name and clocks."""
    import subprocess
    info = {"nproc": None, "affinity": len(os.sched_getaffinity(0)), "model": None, "mhz_max": None}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True, timeout=10).stdout.strip())
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "CPU max MHz":
                info["mhz_max"] = float(v)
            elif k.strip() == "Socket(s)":
                info["sockets"] = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return info


# ------------------------------------------------------- sampled CPU replay
# The CPU baseline of a time-stepping run: the device run is repeated with
# a contact backend that forwards every call to the device and keeps a
# uniform sample of each entry point's calls (with the constraint set and
# lagged friction set in force at the call); the sampled calls are then
# replayed on the -O3 build of the oracle (the CPU restatement, OpenMP
# over the host cores), each timed alone, at all cores and, on a subset of
# the same calls, at one core. CPU time of the run = sum over entry points
# of (calls x mean CPU time per call) + the stepper's own host time.
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp, _d, _i32 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int32
BACKEND_PROTO = {
    "compute_earliest_toi": ctypes.CFUNCTYPE(ctypes.c_int, _vp, _dp, _dp, _d, _dp),
    "construct_constraint_set": ctypes.CFUNCTYPE(ctypes.c_int, _vp, _dp, _d, _d, _i64p),
    "compute_barrier_term": ctypes.CFUNCTYPE(ctypes.c_int, _vp, _dp, _d, _d, _i32, _dp, _dp, _vp),
    "compute_minimum_distance": ctypes.CFUNCTYPE(ctypes.c_int, _vp, _dp, _d, _d, _i32, _dp),
    "update_friction_constraints": ctypes.CFUNCTYPE(ctypes.c_int, _vp, _dp, _d, _d, _d, _d, _i32, _i64p),
    "compute_friction_term": ctypes.CFUNCTYPE(ctypes.c_int, _vp, _dp, _dp, _d, _i32, _dp, _dp, _vp),
}
RIGC_BUILD_CONSTRAINTS, RIGC_HESSIAN, RIGC_ERR_CAPACITY = 4, 2, 4


def device_backend(ctx):
    """The device library's entry points as a contact-backend table over
    the rigidipc_amd.Context ctx (the table rigs_create builds itself)."""
    from rigidipc_amd import sim as M
    L = ctx._L
    be = M.Backend()
    be.self = ctx._ctx.value
    for member, fn in (("set_scene", "rigc_set_scene"), ("compute_earliest_toi", "rigc_compute_earliest_toi"),
                       ("construct_constraint_set", "rigc_construct_constraint_set"),
                       ("compute_barrier_term", "rigc_compute_barrier_term"),
                       ("compute_minimum_distance", "rigc_compute_minimum_distance"),
                       ("update_friction_constraints", "rigc_update_friction_constraints"),
                       ("compute_friction_term", "rigc_compute_friction_term"), ("last_error", "rigc_last_error"),
                       ("get_stats", "rigc_get_stats"), ("set_detection_method", "rigc_set_detection_method")):
        setattr(be, member, ctypes.cast(getattr(L, fn), ctypes.c_void_p).value)
    return be


class CallRecorder:
    """A contact backend that forwards every call to `inner` and keeps, per
    entry point, a uniform sample (reservoir of k, fixed seed) of its calls
    with what a replay needs."""

    def __init__(self, inner, num_bodies, k=64, seed=0):
        import random
        from rigidipc_amd import sim as M
        self.n = 6 * num_bodies
        self.k = k
        self.rng = random.Random(seed)
        self.count = {name: 0 for name in BACKEND_PROTO}
        self.sample = {name: [] for name in BACKEND_PROTO}
        self.hessian_calls = {"calls": 0, "constraints": 0, "assembly_ms": 0.0}
        self.build = None  # (poses, dhat, dmin) of the constraint set in force
        self.fric = None   # (build, poses, dhat, dmin, kappa, mu, flags) of the lagged friction set
        self.table = M.Backend()
        ctypes.memmove(ctypes.byref(self.table), ctypes.byref(inner), ctypes.sizeof(M.Backend))
        self.table.detect_intersections = None
        self.table.compute_earliest_toi_piecewise_linear = None
        self.table.friction_adjoint = None
        from rigidipc_amd import Stats
        self._stats_t = Stats
        self._stats = (ctypes.CFUNCTYPE(ctypes.c_int, _vp, ctypes.POINTER(Stats))(inner.get_stats)
                       if inner.get_stats else None)
        self._callbacks = []
        for name, proto in BACKEND_PROTO.items():
            fwd = proto(getattr(inner, name))
            cb = proto(getattr(self, "_" + name)(fwd))
            self._callbacks.append(cb)
            setattr(self.table, name, ctypes.cast(cb, ctypes.c_void_p).value)

    def _p(self, ptr):
        return np.ctypeslib.as_array(ptr, shape=(self.n,)).copy()

    def _keep(self, name, make):
        self.count[name] += 1
        s = self.sample[name]
        if len(s) < self.k:
            s.append(make())
        else:
            j = self.rng.randrange(self.count[name])
            if j < self.k:
                s[j] = make()

    def _compute_earliest_toi(self, fwd):
        def cb(h, p0, p1, msd, out):
            kept = []

            def make():
                kept.append([self._p(p0), self._p(p1), msd, None])
                return kept[0]
            self._keep("compute_earliest_toi", make)
            st = fwd(h, p0, p1, msd, out)
            if kept and self._stats is not None:
                # the device's inclusion-function evaluations of this call
                s = self._stats_t()
                if self._stats(h, ctypes.byref(s)) == 0:
                    kept[0][3] = int(s.box_evaluations)
            return st
        return cb

    def _construct_constraint_set(self, fwd):
        def cb(h, p, dhat, dmin, counts):
            self.build = (self._p(p), dhat, dmin)
            self._keep("construct_constraint_set", lambda: self.build)
            return fwd(h, p, dhat, dmin, counts)
        return cb

    def _with_build(self, name, p, dhat, dmin, flags, rest):
        before = None if flags & RIGC_BUILD_CONSTRAINTS else self.build
        item = [None]

        def make():
            item[0] = (before, self._p(p), dhat, dmin, flags) + rest
            return item[0]
        self._keep(name, make)
        if flags & RIGC_BUILD_CONSTRAINTS:
            self.build = (self._p(p), dhat, dmin)

    def _compute_barrier_term(self, fwd):
        def cb(h, p, dhat, dmin, flags, e, g, hess):
            self._with_build("compute_barrier_term", p, dhat, dmin, flags, ())
            st = fwd(h, p, dhat, dmin, flags, e, g, hess)
            if flags & RIGC_HESSIAN and self._stats is not None:
                s = self._stats_t()
                if self._stats(h, ctypes.byref(s)) == 0:
                    self.hessian_calls["calls"] += 1
                    self.hessian_calls["constraints"] += int(s.num_constraints)
                    self.hessian_calls["assembly_ms"] += float(s.assembly_ms)
            return st
        return cb

    def _compute_minimum_distance(self, fwd):
        def cb(h, p, dhat, dmin, flags, out):
            self._with_build("compute_minimum_distance", p, dhat, dmin, flags, ())
            return fwd(h, p, dhat, dmin, flags, out)
        return cb

    def _update_friction_constraints(self, fwd):
        def cb(h, p, dhat, dmin, kappa, mu, flags, count):
            before = None if flags & RIGC_BUILD_CONSTRAINTS else self.build
            self.fric = (before, self._p(p), dhat, dmin, kappa, mu, flags)
            if flags & RIGC_BUILD_CONSTRAINTS:
                self.build = (self._p(p), dhat, dmin)
            self._keep("update_friction_constraints", lambda: self.fric)
            return fwd(h, p, dhat, dmin, kappa, mu, flags, count)
        return cb

    def _compute_friction_term(self, fwd):
        def cb(h, p, p0, eh, flags, e, g, hess):
            self._keep("compute_friction_term", lambda: (self.fric, self._p(p), self._p(p0), eh, flags))
            return fwd(h, p, p0, eh, flags, e, g, hess)
        return cb


class Replayer:
    """Replays recorded calls on a contact-backend table (the oracle's),
    timing each call alone; its set-up (the constraint set and lagged
    friction set in force at the call) is not timed."""

    def __init__(self, table, num_bodies):
        from rigidipc_amd import BlockHessian
        self.t = table
        self.nb = num_bodies
        self.fn = {name: proto(getattr(table, name)) for name, proto in BACKEND_PROTO.items()}
        self.grad = np.zeros(6 * num_bodies)
        self.diag = np.zeros(36 * num_bodies)
        self.cap = 64 * num_bodies
        self._alloc()
        self.BlockHessian = BlockHessian

    def _alloc(self):
        self.pairs = np.zeros(2 * self.cap, np.int32)
        self.off = np.zeros(36 * self.cap)

    def _hess(self):
        from rigidipc_amd import _p, _c_double_p, _c_int32_p
        return self.BlockHessian(_p(self.diag, _c_double_p), _p(self.pairs, _c_int32_p), _p(self.off, _c_double_p),
                                 self.cap, 0)

    def _check(self, st, what):
        if st != 0:
            raise RuntimeError(f"oracle {what} failed ({st})")

    def _build(self, b):
        counts = np.zeros(4, np.int64)
        p, dhat, dmin = b
        self._check(self.fn["construct_constraint_set"](self.t.self, p.ctypes.data_as(_dp), dhat, dmin,
                                                        counts.ctypes.data_as(_i64p)), "constraint set")

    def _fupdate(self, f):
        before, p, dhat, dmin, kappa, mu, flags = f
        if before is not None:
            self._build(before)
        n = ctypes.c_int64()
        self._check(self.fn["update_friction_constraints"](self.t.self, p.ctypes.data_as(_dp), dhat, dmin, kappa, mu,
                                                           flags, ctypes.byref(n)), "friction update")

    def run(self, name, item):
        """Seconds of one replayed call."""
        e = ctypes.c_double()
        g = self.grad.ctypes.data_as(_dp)
        for _ in range(3):
            if name == "compute_earliest_toi":
                p0, p1, msd = item[:3]
                out = ctypes.c_double()
                t0 = time.perf_counter()
                st = self.fn[name](self.t.self, p0.ctypes.data_as(_dp), p1.ctypes.data_as(_dp), msd, ctypes.byref(out))
                self.last_toi = out.value
            elif name == "construct_constraint_set":
                t0 = time.perf_counter()
                self._build(item)
                st = 0
            elif name == "update_friction_constraints":
                if item[0] is not None:
                    self._build(item[0])
                t0 = time.perf_counter()
                self._fupdate((None,) + item[1:])
                st = 0
            elif name in ("compute_barrier_term", "compute_minimum_distance"):
                before, p, dhat, dmin, flags = item
                if before is not None:
                    self._build(before)
                pp = p.ctypes.data_as(_dp)
                if name == "compute_minimum_distance":
                    t0 = time.perf_counter()
                    st = self.fn[name](self.t.self, pp, dhat, dmin, flags, ctypes.byref(e))
                else:
                    h = self._hess()
                    t0 = time.perf_counter()
                    st = self.fn[name](self.t.self, pp, dhat, dmin, flags, ctypes.byref(e), g,
                                       ctypes.byref(h) if flags & RIGC_HESSIAN else None)
            else:
                fric, p, p0, eh, flags = item
                self._fupdate(fric)
                h = self._hess()
                t0 = time.perf_counter()
                st = self.fn[name](self.t.self, p.ctypes.data_as(_dp), p0.ctypes.data_as(_dp), eh, flags,
                                   ctypes.byref(e), g, ctypes.byref(h) if flags & RIGC_HESSIAN else None)
            dt = time.perf_counter() - t0
            if st == RIGC_ERR_CAPACITY and name in ("compute_barrier_term", "compute_friction_term"):
                self.cap = 2 * max(self.cap, int(h.count))
                self._alloc()
                continue
            self._check(st, name)
            return dt
        raise RuntimeError(f"oracle {name}: Hessian capacity")


def record_device_run(sc, steps, params=None, k=64):
    """The device run repeated through a CallRecorder: (recorder, stepper
    stats)."""
    import rigidipc_amd as R
    from rigidipc_amd import sim as M
    ctx = R.Context()
    rec = CallRecorder(device_backend(ctx), len(sc.group_ids), k=k)
    sim = M.Simulation(sc, params=params, backend=ctypes.pointer(rec.table))
    for k in range(steps):
        sim.step(1)
        if steps <= 20 or (k + 1) % 10 == 0:
            log(f"recording run: step {k + 1}/{steps}")
    st = sim.stats()
    sim.close()
    ctx.close()
    return rec, st


def replay_cpu(sc, rec, budget_s, params=None):
    """Replays rec's samples on the -O3 oracle. All cores: per entry point,
    samples in order until its share of budget_s is spent (at least 2);
    one core: the first of those same calls until half that share is spent
    (at least 1), giving per entry point the one-core / all-core ratio on
    identical calls. Returns per entry point (calls, mean all-core seconds,
    samples timed, one-core ratio, samples timed at one core)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from rigidipc_amd import sim as M
    L3 = O.lib_o3()
    L3.oracle_backend.argtypes = [ctypes.POINTER(M.Backend)]
    L3.oracle_backend.restype = None
    table = M.Backend()
    L3.oracle_backend(ctypes.byref(table))
    holder = M.Simulation(sc, params=params, backend=ctypes.pointer(table))  # sets the oracle's scene
    rp = Replayer(table, len(sc.group_ids))
    names = [n for n in BACKEND_PROTO if rec.count[n] > 0]
    share = budget_s / max(len(names), 1)
    all_threads = O.max_threads(L3)
    out = {}
    boxes = {"oracle_tree": 0, "device_executed": 0, "calls": 0}
    for name in names:
        times = []
        t_spent = 0.0
        for item in rec.sample[name]:
            if len(times) >= 2 and t_spent >= share:
                break
            times.append(rp.run(name, item))
            t_spent += times[-1]
            if name == "compute_earliest_toi" and item[3] is not None:
                boxes["oracle_tree"] += O.last_box_evaluations(L3)
                boxes["device_executed"] += item[3]
                boxes["calls"] += 1
        one = []
        O.set_num_threads(1, L3)
        try:
            t_spent = 0.0
            for item in rec.sample[name][:len(times)]:
                if one and t_spent >= share / 2:
                    break
                one.append(rp.run(name, item))
                t_spent += one[-1]
        finally:
            O.set_num_threads(all_threads, L3)
        out[name] = dict(calls=rec.count[name], mean_s=float(np.mean(times)), timed=len(times),
                         one_core_ratio=float(sum(one) / max(sum(times[:len(one)]), 1e-12)), timed_one_core=len(one))
        log(f"cpu replay {name}: {rec.count[name]} calls, {len(times)} timed, mean {1e3 * out[name]['mean_s']:.2f} ms "
            f"({all_threads} threads), x{out[name]['one_core_ratio']:.2f} at one core over {len(one)}")
    holder.close()
    out["_boxes"] = boxes
    return out, all_threads


def cpu_estimate(sc, steps, budget_s, params=None, k=64, recorded=None):
    """Sampled-replay CPU baseline of `steps` steps of sc from its initial
    state: dict with all-core and one-core time steps/s and the details.
    recorded: a (recorder, stats) pair from record_device_run of the same
    run, else the run is recorded here."""
    rec, st = recorded if recorded is not None else record_device_run(sc, steps, params, k)
    per, threads = replay_cpu(sc, rec, budget_s, params)
    boxes = per.pop("_boxes")
    host_s = (st["total_ms"] - st["contact_ms"]) * 1e-3
    contact_all = sum(v["calls"] * v["mean_s"] for v in per.values())
    contact_one = sum(v["calls"] * v["mean_s"] * v["one_core_ratio"] for v in per.values())
    return dict(
        all_cores=steps / (contact_all + host_s), one_core=steps / (contact_one + host_s), threads=threads,
        contact_s_all_cores=contact_all, contact_s_one_core=contact_one, host_s=host_s,
        calls={n: v["calls"] for n, v in per.items()},
        cpu_ms_per_call={n: round(1e3 * v["mean_s"], 3) for n, v in per.items()},
        samples_timed={n: v["timed"] for n, v in per.items()},
        one_core_ratio={n: round(v["one_core_ratio"], 2) for n, v in per.items()},
        samples_timed_one_core={n: v["timed_one_core"] for n, v in per.items()},
        newton_iterations=int(st["newton_iterations"]),
        ccd_box_evaluations=dict(boxes, note="over the CCD calls timed: inclusion-function evaluations of the "
                                             "oracle's trees (queries sharing the running minimum) and executed "
                                             "by the device kernel (speculative evaluations included)"))


def config1_outright(steps=100, no_cpu=False):
    """Config 1 (10 ico2 spheres + plane, 100 steps) run outright on the
    device and, all cores, on the -O3 oracle; plus the sampled estimate of
    the same run, checked against the outright CPU time."""
    from rigidipc_amd import scenes as S
    from rigidipc_amd import sim as M
    sc = S.config1(seed=0)
    sim = M.Simulation(sc)
    t0 = time.perf_counter()
    sim.step(steps)
    gpu_s = time.perf_counter() - t0
    x_gpu, _ = sim.state()
    st = sim.stats()
    sim.close()
    out = dict(workload="config1: 10 ico2 spheres + static plane, dt 0.01, dhat 1e-3, mu 0, 100 steps",
               steps=steps, gpu_steps_per_s=steps / gpu_s, gpu_newton_iterations=int(st["newton_iterations"]))
    log(f"config1: device {steps} steps in {gpu_s:.2f} s")
    if no_cpu:
        return out
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    L3 = O.lib_o3()
    L3.oracle_backend.argtypes = [ctypes.POINTER(M.Backend)]
    L3.oracle_backend.restype = None
    table = M.Backend()
    L3.oracle_backend(ctypes.byref(table))
    cpu = M.Simulation(sc, backend=ctypes.pointer(table))
    t0 = time.perf_counter()
    for k in range(steps):
        cpu.step(1)
        if (k + 1) % 20 == 0:
            log(f"config1 cpu: {k + 1} steps, {time.perf_counter() - t0:.1f} s")
    cpu_s = time.perf_counter() - t0
    x_cpu, _ = cpu.state()
    cst = cpu.stats()
    cpu.close()
    est = cpu_estimate(sc, steps, 10.0)
    out.update(cpu_steps_per_s_all_cores=steps / cpu_s, cpu_threads=O.max_threads(L3),
               cpu_newton_iterations=int(cst["newton_iterations"]),
               max_end_pose_difference=float(np.abs(x_gpu - x_cpu).max()),
               gpu_vs_cpu=(steps / gpu_s) / (steps / cpu_s),
               sampled_estimate=dict(all_cores=est["all_cores"], one_core=est["one_core"],
                                     cpu_ms_per_call=est["cpu_ms_per_call"], calls=est["calls"]),
               gpu_vs_cpu_one_core=(steps / gpu_s) / est["one_core"])
    log(f"config1: cpu {steps} steps in {cpu_s:.1f} s (sampled estimate {steps / est['all_cores']:.1f} s)")
    return out


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
    bp_bytes = broad_phase_bytes(sc, st["ccd_calls"], st["ccd_candidates"])
    bp_s = st["ccd_broad_phase_ms"] * 1e-3
    line["broad_phase"] = dict(
        bound="hbm", achieved=bp_bytes / bp_s * 1e-9 if bp_s > 0 else None, peak=HBM_PEAK_GBS, unit="GB/s",
        frac=bp_bytes / bp_s * 1e-9 / HBM_PEAK_GBS if bp_s > 0 else None, bytes_per_call=bp_bytes / st["ccd_calls"],
        calls=int(st["ccd_calls"]), device_ms=st["ccd_broad_phase_ms"],
        note="continuous broad phase of the CCD calls: algorithmic bytes (poses, mesh, element boxes written and "
             "read, grid keys, candidate pairs) over the device time between its events")
    log("recording run (contact calls sampled for the CPU baseline)")
    recorded = record_device_run(sc, args.steps)
    hc = recorded[0].hessian_calls
    flop = hc["constraints"] * ASSEMBLY_FLOP_PER_CONSTRAINT
    line["assembly"] = dict(
        bound="fp64", achieved=flop / (hc["assembly_ms"] * 1e-3) * 1e-12 if hc["assembly_ms"] > 0 else None,
        peak=FP64_PEAK_TFLOPS, unit="TFLOP/s",
        frac=flop / (hc["assembly_ms"] * 1e-3) * 1e-12 / FP64_PEAK_TFLOPS if hc["assembly_ms"] > 0 else None,
        flop_per_constraint=ASSEMBLY_FLOP_PER_CONSTRAINT, calls=hc["calls"], constraints=hc["constraints"],
        device_ms=hc["assembly_ms"],
        note="barrier terms with Hessian of the same run: constraints x algorithmic FLOP per constraint over the "
             "device time of the assembly (constraint terms, PSD projection, block reduction)")
    if not args.no_cpu:
        est = cpu_estimate(sc, args.steps, args.cpu_budget, recorded=recorded)
        hc = host_cpu()
        line["cpu_baseline"] = dict(
            value=est["all_cores"], unit="time steps/s", cores=est["threads"], kind="port",
            value_one_core=est["one_core"], host=hc,
            sample=f"the same {args.steps} steps from the same initial state, repeated on the device through a "
                   f"recording backend; a uniform sample of each contact entry point's calls over all the steps "
                   f"replayed on the -O3 oracle (OpenMP, {est['threads']} threads; at one thread on a subset of "
                   f"the same calls), each timed alone; CPU time = sum of calls x mean time per call + the "
                   f"stepper's host time ({est['host_s']:.2f} s)",
            detail=est)
        line["gpu_vs_cpu"] = dict(all_cores=line["value"] / est["all_cores"], one_core=line["value"] / est["one_core"])
        if not args.sim_only:
            line["config1"] = config1_outright()
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
    iters = None
    for name, mode in (("fused", M.BATCH_FUSED), ("plain", M.BATCH_SEQUENTIAL)):
        b = M.Batch(scenes, mode=mode)
        t0 = time.perf_counter()
        if mode == M.BATCH_FUSED:
            # one call: every scene runs its `steps` steps on its own thread,
            # so a scene that is slow in one step does not hold the others
            # at the step boundary (RIGS_BATCH_LOG: progress lines as scenes
            # finish)
            os.environ.setdefault("RIGS_BATCH_LOG", "1")
            b.step(steps)
            log(f"batch {name}: {steps} steps of every scene in {time.perf_counter() - t0:.2f} s")
        else:
            for k in range(steps):
                t = time.perf_counter()
                b.step(1)
                log(f"batch {name} step {k + 1}/{steps}: {time.perf_counter() - t:.2f} s")
        wall = time.perf_counter() - t0
        it = np.array([b.stats(i)["newton_iterations"] for i in range(n_scenes)], dtype=np.int64)
        out[name] = dict(scene_steps_per_s=n_scenes * steps / wall, wall_s=wall, newton_iterations=int(it.sum()))
        if mode == M.BATCH_FUSED:
            bs = b.batch_stats()
            out[name].update(rounds=int(bs["rounds"]), fused_calls=int(bs["fused_calls"]),
                             requests=int(bs["requests"]), executors=int(bs["executors"]),
                             device_ms_all_executors=bs["device_ms"],
                             autotuned=dict(ccd_walk_budget=int(bs["ccd_walk_budget"]),
                                            ccd_spec_levels=int(bs["ccd_spec_levels"]),
                                            ccd_window_factor=bs["ccd_window_factor"],
                                            ccd_wide_queries=int(bs["ccd_wide_queries"]),
                                            sample_device_ms=bs["autotune_ms"]))
            iters = it
        b.close()
    out["fused_over_plain"] = out["fused"]["scene_steps_per_s"] / out["plain"]["scene_steps_per_s"]
    # Newton iterations per scene over all the steps: the fused run needs at
    # least max(iters) rounds, the plain run does sum(iters) iterations
    out["newton_iterations_per_scene"] = dict(max=int(iters.max()), median=float(np.median(iters)),
                                              sum_over_max=float(iters.sum() / max(1, iters.max())))
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
    ap.add_argument("--batch-steps", type=int, default=10,
                    help="steps of the batched-vs-plain comparison added to the sim line (0: skip; the "
                         "batching workload is 10 steps)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sim-only", action="store_true",
                    help="the stepping workload alone: no config1, batched or gradient entries (for profiles)")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work")
    args = ap.parse_args()
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
        if args.no_cpu and not args.sim_only:
            line["config1"] = config1_outright(no_cpu=True)
        if args.batch_steps > 0 and not args.sim_only:
            line["batched"] = batch_compare(256, args.batch_steps)
            line["batched"]["note"] = (f"{args.batch_steps} of the batching workload's 10 steps, every scene "
                                       "from its initial state, fused then plain")
        if not args.sim_only:
            line["gradient"] = grad_compare(cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
