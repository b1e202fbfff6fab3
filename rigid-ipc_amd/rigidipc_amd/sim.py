"""Host time stepper (``librigc_sim.so``, ``rigid-ipc_amd/host/rigs.h``).

``Simulation(scene)`` steps a scene from :mod:`rigidipc_amd.scenes` with the
reference's implicit-Euler IPC Newton solve on the host and every contact
term (CCD, constraint set, barrier, friction, minimum distance) on the
device. Another contact backend can be passed as a ``Backend`` table
(ctypes pointer), which is how the tests and the CPU baseline run the same
stepper over the CPU restatement.
"""
import ctypes
import os

import numpy as np

from . import Scene, RigcError, RIGC_OK, _p, _c_double_p, _c_int32_p, _c_uint8_p

_HERE = os.path.dirname(os.path.abspath(__file__))
SIM_PATH = os.path.join(os.path.dirname(_HERE), "librigc_sim.so")


class Backend(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in (
        "self", "set_scene", "compute_earliest_toi", "construct_constraint_set", "compute_barrier_term",
        "compute_minimum_distance", "update_friction_constraints", "compute_friction_term", "last_error",
        "get_stats", "detect_intersections", "compute_earliest_toi_piecewise_linear", "friction_adjoint",
        "set_detection_method")]


class Params(ctypes.Structure):
    _fields_ = [
        ("timestep", ctypes.c_double), ("gravity", ctypes.c_double * 3), ("dhat", ctypes.c_double),
        ("minimum_separation_distance", ctypes.c_double), ("coefficient_friction", ctypes.c_double),
        ("friction_iterations", ctypes.c_int32), ("static_friction_speed_bound", ctypes.c_double),
        ("velocity_conv_tol", ctypes.c_double), ("line_search_lower_bound", ctypes.c_double),
        ("max_iterations", ctypes.c_int32), ("dhat_epsilon", ctypes.c_double),
        ("min_barrier_stiffness_scale", ctypes.c_double), ("do_intersection_check", ctypes.c_int32),
        ("trajectory_type", ctypes.c_int32), ("newton_polish", ctypes.c_int32),
        ("detection_method", ctypes.c_int32),
    ]


class Bodies(ctypes.Structure):
    _fields_ = [("num_bodies", ctypes.c_int32), ("poses", _c_double_p), ("velocities", _c_double_p),
                ("mass", _c_double_p), ("inertia", _c_double_p), ("force", _c_double_p),
                ("kinematic", _c_int32_p), ("kinematic_max_time", _c_double_p),
                ("kinematic_pose_start", _c_int32_p), ("kinematic_poses", _c_double_p)]


def _body_extras(scene, keep):
    """Optional rigs_bodies arrays of a scene (external forces, kinematic
    bodies): keeps them alive in `keep` and returns the Bodies field values."""
    out = {}
    if getattr(scene, "force", None) is not None:
        keep["force"] = np.ascontiguousarray(scene.force, np.float64).reshape(-1, 6)
        out["force"] = _p(keep["force"], _c_double_p)
    if getattr(scene, "kinematic", None) is not None:
        keep["kinematic"] = np.ascontiguousarray(scene.kinematic, np.int32)
        keep["kinematic_max_time"] = np.ascontiguousarray(scene.kinematic_max_time, np.float64)
        keep["kinematic_pose_start"] = np.ascontiguousarray(scene.kinematic_pose_start, np.int32)
        keep["kinematic_poses"] = np.ascontiguousarray(scene.kinematic_poses, np.float64).reshape(-1, 6)
        out["kinematic"] = _p(keep["kinematic"], _c_int32_p)
        out["kinematic_max_time"] = _p(keep["kinematic_max_time"], _c_double_p)
        out["kinematic_pose_start"] = _p(keep["kinematic_pose_start"], _c_int32_p)
        out["kinematic_poses"] = _p(keep["kinematic_poses"], _c_double_p)
    return out


class Stats(ctypes.Structure):
    _fields_ = [
        ("steps", ctypes.c_int64), ("newton_iterations", ctypes.c_int64),
        ("line_search_iterations", ctypes.c_int64), ("ccd_calls", ctypes.c_int64),
        ("constraint_builds", ctypes.c_int64), ("objective_evaluations", ctypes.c_int64),
        ("regularizations", ctypes.c_int64), ("failed_solves", ctypes.c_int64),
        ("contact_ms", ctypes.c_double), ("host_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
        ("kappa", ctypes.c_double), ("min_distance", ctypes.c_double), ("num_constraints", ctypes.c_int64),
        ("ccd_candidates", ctypes.c_int64), ("ccd_ms", ctypes.c_double), ("ccd_broad_phase_ms", ctypes.c_double),
        ("ccd_narrow_phase_ms", ctypes.c_double), ("ccd_box_evaluations", ctypes.c_int64),
        ("ccd_kernel_ms", ctypes.c_double), ("ccd_launches", ctypes.c_int64),
        ("constraint_ms", ctypes.c_double), ("barrier_ms", ctypes.c_double), ("friction_ms", ctypes.c_double),
        ("distance_ms", ctypes.c_double), ("initial_intersections", ctypes.c_int64),
        ("intersections", ctypes.c_int64), ("intersecting_steps", ctypes.c_int64),
    ]


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int64), ("fused_calls", ctypes.c_int64), ("requests", ctypes.c_int64),
        ("device_ms", ctypes.c_double), ("ccd_calls", ctypes.c_int64), ("ccd_candidates", ctypes.c_int64),
        ("ccd_kernel_ms", ctypes.c_double), ("ccd_launches", ctypes.c_int64),
        ("ccd_box_evaluations", ctypes.c_int64), ("executors", ctypes.c_int32),
        ("pad", ctypes.c_int32), ("ccd_walk_budget", ctypes.c_int32), ("ccd_spec_levels", ctypes.c_int32),
        ("ccd_window_factor", ctypes.c_double), ("ccd_wide_queries", ctypes.c_int64), ("autotune_ms", ctypes.c_double),
    ]


BATCH_SEQUENTIAL = 0
BATCH_FUSED = 1

_sim = None


def sim_lib():
    global _sim
    if _sim is not None:
        return _sim
    if not os.path.exists(SIM_PATH):
        raise RuntimeError(f"{SIM_PATH} is not built; run rigid-ipc_amd/build.py")
    L = ctypes.CDLL(SIM_PATH)
    vp = ctypes.c_void_p
    L.rigs_default_params.argtypes = [ctypes.POINTER(Params)]
    L.rigs_default_params.restype = None
    L.rigs_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Backend), ctypes.c_int32, ctypes.POINTER(Scene),
                              ctypes.POINTER(Bodies), ctypes.POINTER(Params)]
    L.rigs_destroy.argtypes = [vp]
    L.rigs_last_error.argtypes = [vp]
    L.rigs_last_error.restype = ctypes.c_char_p
    L.rigs_step.argtypes = [vp, ctypes.c_int32]
    L.rigs_get_state.argtypes = [vp, _c_double_p, _c_double_p]
    L.rigs_set_state.argtypes = [vp, _c_double_p, _c_double_p]
    L.rigs_rollout.argtypes = [vp, ctypes.c_int32]
    L.rigs_rollout_backward.argtypes = [vp, _c_double_p, _c_double_p, _c_double_p, _c_double_p]
    L.rigs_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.rigs_block_ldlt_solve.argtypes = [ctypes.c_int32, _c_double_p, ctypes.c_int64, _c_int32_p, _c_double_p,
                                        _c_double_p, _c_double_p]
    L.rigs_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(Scene),
                                    ctypes.POINTER(Bodies), ctypes.POINTER(Params), ctypes.c_int32]
    L.rigs_batch_destroy.argtypes = [vp]
    L.rigs_batch_last_error.argtypes = [vp]
    L.rigs_batch_last_error.restype = ctypes.c_char_p
    L.rigs_batch_step.argtypes = [vp, ctypes.c_int32]
    L.rigs_batch_get_state.argtypes = [vp, ctypes.c_int32, _c_double_p, _c_double_p]
    L.rigs_batch_get_stats.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(Stats)]
    L.rigs_batch_get_batch_stats.argtypes = [vp, ctypes.POINTER(BatchStats)]
    L.rigs_batch_set_gather.argtypes = [vp, ctypes.c_double]
    for name in ("rigs_create", "rigs_destroy", "rigs_step", "rigs_get_state", "rigs_get_stats", "rigs_set_state",
                 "rigs_rollout", "rigs_rollout_backward",
                 "rigs_block_ldlt_solve", "rigs_batch_create", "rigs_batch_destroy", "rigs_batch_step",
                 "rigs_batch_get_state", "rigs_batch_get_stats", "rigs_batch_get_batch_stats",
                 "rigs_batch_set_gather"):
        getattr(L, name).restype = ctypes.c_int32
    _sim = L
    return L


def default_params(scene=None):
    p = Params()
    sim_lib().rigs_default_params(ctypes.byref(p))
    if scene is not None:
        p.timestep = float(getattr(scene, "dt", p.timestep))
        p.dhat = float(getattr(scene, "dhat", p.dhat))
        p.coefficient_friction = float(getattr(scene, "mu", 0.0))
        g = getattr(scene, "gravity", None)
        if g is not None:
            for k in range(3):
                p.gravity[k] = float(g[k])
    return p


def block_ldlt_solve(diag, pairs, off, b):
    """Solve with the stepper's sparse block LDL^T (test hook)."""
    L = sim_lib()
    diag = np.ascontiguousarray(diag, np.float64).reshape(-1, 36)
    pairs = np.ascontiguousarray(pairs, np.int32).reshape(-1, 2)
    off = np.ascontiguousarray(off, np.float64).reshape(-1, 36)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    st = L.rigs_block_ldlt_solve(len(diag), _p(diag, _c_double_p), len(pairs), _p(pairs, _c_int32_p),
                                 _p(off, _c_double_p), _p(b, _c_double_p), _p(x, _c_double_p))
    if st != RIGC_OK:
        raise RigcError(st, "block LDL^T failed")
    return x


def _scene_inputs(scene):
    """(keep-alive arrays, Scene, Bodies) of a scene from rigidipc_amd.scenes."""
    k = dict(
        vertices=np.ascontiguousarray(scene.vertices, np.float64),
        edges=np.ascontiguousarray(scene.edges, np.int32),
        faces=np.ascontiguousarray(scene.faces, np.int32),
        face_edges=np.ascontiguousarray(scene.face_edges, np.int32),
        body_vertex_start=np.ascontiguousarray(scene.body_vertex_start, np.int32),
        body_edge_start=np.ascontiguousarray(scene.body_edge_start, np.int32),
        body_face_start=np.ascontiguousarray(scene.body_face_start, np.int32),
        group_ids=np.ascontiguousarray(scene.group_ids, np.int32),
        is_static=np.ascontiguousarray(scene.is_static, np.uint8),
        poses=np.ascontiguousarray(scene.poses, np.float64),
        velocities=np.ascontiguousarray(scene.velocities, np.float64),
        mass=np.ascontiguousarray(scene.mass, np.float64),
        inertia=np.ascontiguousarray(scene.inertia, np.float64),
    )
    s = Scene()
    s.num_bodies, s.num_vertices = len(k["group_ids"]), len(k["vertices"])
    s.num_edges, s.num_faces = len(k["edges"]), len(k["faces"])
    for name in ("vertices", "edges", "faces", "face_edges", "body_vertex_start", "body_edge_start",
                 "body_face_start", "group_ids", "is_static"):
        a = k[name]
        t = {np.float64: _c_double_p, np.int32: _c_int32_p, np.uint8: _c_uint8_p}[a.dtype.type]
        setattr(s, name, _p(a, t))
    b = Bodies(s.num_bodies, _p(k["poses"], _c_double_p), _p(k["velocities"], _c_double_p),
               _p(k["mass"], _c_double_p), _p(k["inertia"], _c_double_p))
    for name, v in _body_extras(scene, k).items():
        setattr(b, name, v)
    return k, s, b


class Batch:
    """Many independent scenes stepped together (rigs_batch_*): mode
    BATCH_FUSED serves all scenes' contact calls with fused multi-scene device
    calls; BATCH_SEQUENTIAL steps them one after another on the plain path."""

    def __init__(self, scenes, mode=BATCH_FUSED, params=None, device=0):
        self._L = sim_lib()
        ins = [_scene_inputs(sc) for sc in scenes]
        self._keep = [k for k, _, _ in ins]
        n = len(scenes)
        S = (Scene * n)(*[s for _, s, _ in ins])
        B = (Bodies * n)(*[b for _, _, b in ins])
        P = (Params * n)(*[(params[i] if params is not None else default_params(sc)) for i, sc in enumerate(scenes)])
        self._arrays = (S, B, P)
        self.n = n
        self.num_bodies = [int(len(sc.group_ids)) for sc in scenes]
        self._b = ctypes.c_void_p()
        st = self._L.rigs_batch_create(ctypes.byref(self._b), int(device), n, S, B, P, int(mode))
        if st != RIGC_OK:
            raise RigcError(st, "rigs_batch_create failed")

    def close(self):
        if self._b:
            self._L.rigs_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        st = self._L.rigs_batch_step(self._b, int(n))
        if st != RIGC_OK:
            raise RigcError(st, self._L.rigs_batch_last_error(self._b).decode())

    def set_gather(self, gather_us):
        """Fused mode's gather window in microseconds (< 0: a round waits
        until every stepping scene has a request waiting)."""
        st = self._L.rigs_batch_set_gather(self._b, float(gather_us))
        if st != RIGC_OK:
            raise RigcError(st, "rigs_batch_set_gather failed")

    def state(self, i):
        poses = np.zeros((self.num_bodies[i], 6))
        vel = np.zeros((self.num_bodies[i], 6))
        self._L.rigs_batch_get_state(self._b, int(i), _p(poses, _c_double_p), _p(vel, _c_double_p))
        return poses, vel

    def stats(self, i):
        s = Stats()
        self._L.rigs_batch_get_stats(self._b, int(i), ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def batch_stats(self):
        s = BatchStats()
        self._L.rigs_batch_get_batch_stats(self._b, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in BatchStats._fields_}


class Simulation:
    """DistanceBarrierRBProblem time stepping over a contact backend."""

    def __init__(self, scene, params=None, device=0, backend=None):
        self._L = sim_lib()
        self._keep = dict(
            vertices=np.ascontiguousarray(scene.vertices, np.float64),
            edges=np.ascontiguousarray(scene.edges, np.int32),
            faces=np.ascontiguousarray(scene.faces, np.int32),
            face_edges=np.ascontiguousarray(scene.face_edges, np.int32),
            body_vertex_start=np.ascontiguousarray(scene.body_vertex_start, np.int32),
            body_edge_start=np.ascontiguousarray(scene.body_edge_start, np.int32),
            body_face_start=np.ascontiguousarray(scene.body_face_start, np.int32),
            group_ids=np.ascontiguousarray(scene.group_ids, np.int32),
            is_static=np.ascontiguousarray(scene.is_static, np.uint8),
            poses=np.ascontiguousarray(scene.poses, np.float64),
            velocities=np.ascontiguousarray(scene.velocities, np.float64),
            mass=np.ascontiguousarray(scene.mass, np.float64),
            inertia=np.ascontiguousarray(scene.inertia, np.float64),
        )
        k = self._keep
        s = Scene()
        s.num_bodies, s.num_vertices = len(k["group_ids"]), len(k["vertices"])
        s.num_edges, s.num_faces = len(k["edges"]), len(k["faces"])
        for name in ("vertices", "edges", "faces", "face_edges", "body_vertex_start", "body_edge_start",
                     "body_face_start", "group_ids", "is_static"):
            a = k[name]
            t = {np.float64: _c_double_p, np.int32: _c_int32_p, np.uint8: _c_uint8_p}[a.dtype.type]
            setattr(s, name, _p(a, t))
        b = Bodies(s.num_bodies, _p(k["poses"], _c_double_p), _p(k["velocities"], _c_double_p),
                   _p(k["mass"], _c_double_p), _p(k["inertia"], _c_double_p))
        for name, v in _body_extras(scene, k).items():
            setattr(b, name, v)
        self.params = params if params is not None else default_params(scene)
        self.num_bodies = s.num_bodies
        self._sim = ctypes.c_void_p()
        st = self._L.rigs_create(ctypes.byref(self._sim), backend, int(device), ctypes.byref(s), ctypes.byref(b),
                                 ctypes.byref(self.params))
        if st != RIGC_OK:
            raise RigcError(st, "rigs_create failed")

    def close(self):
        if self._sim:
            self._L.rigs_destroy(self._sim)
            self._sim = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        st = self._L.rigs_step(self._sim, int(n))
        if st != RIGC_OK:
            raise RigcError(st, self._L.rigs_last_error(self._sim).decode())

    def state(self):
        poses = np.zeros((self.num_bodies, 6))
        vel = np.zeros((self.num_bodies, 6))
        self._L.rigs_get_state(self._sim, _p(poses, _c_double_p), _p(vel, _c_double_p))
        return poses, vel

    def set_state(self, poses, velocities):
        """Restart from poses / velocities with a fresh solver state."""
        p = np.ascontiguousarray(poses, np.float64).reshape(self.num_bodies, 6)
        v = np.ascontiguousarray(velocities, np.float64).reshape(self.num_bodies, 6)
        st = self._L.rigs_set_state(self._sim, _p(p, _c_double_p), _p(v, _c_double_p))
        if st != RIGC_OK:
            raise RigcError(st, "rigs_set_state failed")

    def rollout(self, n):
        """Advance n steps keeping the checkpoints rollout_backward needs."""
        st = self._L.rigs_rollout(self._sim, int(n))
        if st != RIGC_OK:
            raise RigcError(st, self._L.rigs_last_error(self._sim).decode())

    def rollout_backward(self, dL_dposes, dL_dvelocities=None, want_poses0=False):
        """dL/d(initial velocities) [nb, 6] of the last rollout (and dL/d(initial
        poses) when want_poses0) from dL/d(end poses) and, optionally,
        dL/d(end velocities)."""
        gx = np.ascontiguousarray(dL_dposes, np.float64).reshape(self.num_bodies, 6)
        gv = None if dL_dvelocities is None else np.ascontiguousarray(dL_dvelocities, np.float64).reshape(
            self.num_bodies, 6)
        out = np.zeros((self.num_bodies, 6))
        out_x = np.zeros((self.num_bodies, 6)) if want_poses0 else None
        st = self._L.rigs_rollout_backward(self._sim, _p(gx, _c_double_p),
                                           None if gv is None else _p(gv, _c_double_p), _p(out, _c_double_p),
                                           None if out_x is None else _p(out_x, _c_double_p))
        if st != RIGC_OK:
            raise RigcError(st, self._L.rigs_last_error(self._sim).decode())
        return (out, out_x) if want_poses0 else out

    def stats(self):
        s = Stats()
        self._L.rigs_get_stats(self._sim, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in Stats._fields_}
