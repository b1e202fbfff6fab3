"""ctypes loader for the CPU oracle (oracle/_build/liboracle.so).

The oracle is test infrastructure: it is only ever used as the checker.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
# The -O3 build of the same sources (the CPU baseline's timing build).
ORACLE_O3_SO = os.path.join(ORACLE_DIR, "_build", "liboracle_o3.so")

_lib = None
_lib_o3 = None

c_double_p = ctypes.POINTER(ctypes.c_double)
c_long_p = ctypes.POINTER(ctypes.c_long)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)


def dptr(a):
    return a.ctypes.data_as(c_double_p)


def i32ptr(a):
    return a.ctypes.data_as(c_int32_p)


def i64ptr(a):
    return a.ctypes.data_as(c_int64_p)


def lib():
    """Load the oracle, building it first if it is not there."""
    global _lib
    if _lib is None:
        _lib = _load(ORACLE_SO)
    return _lib


def lib_o3():
    """The -O3 timing build: a second, separate instance of the oracle (its
    own scene and backend state)."""
    global _lib_o3
    if _lib_o3 is None:
        _lib_o3 = _load(ORACLE_O3_SO)
    return _lib_o3


def _load(path):
    if not os.path.exists(path):
        subprocess.check_call(["make", "-C", ORACLE_DIR, "-j8"])
    L = ctypes.CDLL(path)
    for name in ("oracle_interval_cos", "oracle_interval_sin", "oracle_interval_sinc"):
        getattr(L, name).argtypes = [ctypes.c_double, ctypes.c_double, c_double_p]
    for name in ("oracle_sinc", "oracle_cos_rn", "oracle_sin_rn"):
        getattr(L, name).argtypes = [ctypes.c_double]
        getattr(L, name).restype = ctypes.c_double
    L.oracle_interval_binop.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4 + [c_double_p]
    L.oracle_rotation_interval.argtypes = [c_double_p, c_double_p]
    L.oracle_rotation_double.argtypes = [c_double_p, c_double_p]
    L.oracle_root_quadratic.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int, ctypes.c_double, c_double_p]
    L.oracle_root_quadratic.restype = ctypes.c_int
    L.oracle_root_trig.argtypes = [ctypes.c_double] * 3 + [c_double_p]
    L.oracle_root_trig.restype = ctypes.c_int
    for name in ("oracle_ee_toi", "oracle_fv_toi"):
        getattr(L, name).argtypes = [c_double_p, ctypes.c_double, ctypes.c_double, c_double_p, c_long_p]
        getattr(L, name).restype = ctypes.c_int
    for name in ("oracle_pl_ee_toi", "oracle_pl_fv_toi"):
        getattr(L, name).argtypes = [c_double_p, ctypes.c_double, ctypes.c_double, c_double_p, c_long_p,
                                     c_long_p]
        getattr(L, name).restype = ctypes.c_int
    for name in ("oracle_ti_edge_edge", "oracle_ti_vertex_face"):
        getattr(L, name).argtypes = [c_double_p, ctypes.c_double, c_double_p, c_double_p, c_long_p]
        getattr(L, name).restype = ctypes.c_int
    L.oracle_element_distance_sq.argtypes = [ctypes.c_int, c_double_p]
    L.oracle_element_distance_sq.restype = ctypes.c_double
    L.oracle_set_eval_budget.argtypes = [ctypes.c_long]
    L.oracle_set_scene.argtypes = [ctypes.c_int] * 4 + [c_double_p] + [c_int32_p] * 7 + [ctypes.c_void_p]
    L.oracle_orient3d.argtypes = [c_double_p] * 4
    L.oracle_orient3d.restype = ctypes.c_int
    L.oracle_edge_triangle.argtypes = [c_double_p] * 5
    L.oracle_edge_triangle.restype = ctypes.c_int
    L.oracle_detect_intersections.argtypes = [c_double_p, c_int32_p, ctypes.c_long]
    L.oracle_detect_intersections.restype = ctypes.c_long
    L.oracle_broadphase.argtypes = [ctypes.c_int, c_double_p, c_double_p, ctypes.c_double, c_long_p, c_long_p]
    L.oracle_get_candidates.argtypes = [c_int32_p, c_int32_p]
    L.oracle_average_edge_length.restype = ctypes.c_double
    L.oracle_world_vertices.argtypes = [c_double_p, c_double_p]
    L.oracle_last_vertex_aabbs.argtypes = [c_double_p]
    L.oracle_earliest_toi.argtypes = [c_double_p, c_double_p, ctypes.c_double, c_long_p]
    L.oracle_earliest_toi.restype = ctypes.c_double
    for name in ("oracle_ee_aabb", "oracle_fv_aabb"):
        getattr(L, name).argtypes = [c_double_p, c_double_p, c_double_p]
    L.oracle_set_ccd_pruning.argtypes = [ctypes.c_int]
    L.oracle_last_box_evaluations.restype = ctypes.c_long
    L.oracle_set_num_threads.argtypes = [ctypes.c_int]
    L.oracle_max_threads.restype = ctypes.c_int
    L.oracle_set_detection_method.argtypes = [ctypes.c_int]
    L.oracle_bvh_supported.restype = ctypes.c_int
    return L


def set_ccd_pruning(on, L=None):
    """Earliest-TOI queries share the running minimum (on, the default) or
    each runs its whole tree (off)."""
    (L or lib()).oracle_set_ccd_pruning(int(bool(on)))


def last_box_evaluations(L=None):
    """Inclusion-function evaluations of the last earliest-TOI call."""
    return (L or lib()).oracle_last_box_evaluations()


def set_num_threads(n, L=None):
    (L or lib()).oracle_set_num_threads(int(n))


def max_threads(L=None):
    return (L or lib()).oracle_max_threads()


def interval_fn(name, lo, hi):
    out = np.zeros(2)
    getattr(lib(), name)(lo, hi, dptr(out))
    return out


def binop(op, x, y):
    out = np.zeros(2)
    lib().oracle_interval_binop({"+": 0, "-": 1, "*": 2, "/": 3}[op], x[0], x[1], y[0], y[1], dptr(out))
    return out


def ee_toi(q, earliest_toi=1.0, toi_tol=1e-4):
    q = np.ascontiguousarray(q, dtype=np.float64)
    toi = ctypes.c_double()
    n = ctypes.c_long(0)
    hit = lib().oracle_ee_toi(dptr(q), earliest_toi, toi_tol, ctypes.byref(toi), ctypes.byref(n))
    return (hit if hit < 0 else bool(hit)), toi.value, n.value


def fv_toi(q, earliest_toi=1.0, toi_tol=1e-4):
    q = np.ascontiguousarray(q, dtype=np.float64)
    toi = ctypes.c_double()
    n = ctypes.c_long(0)
    hit = lib().oracle_fv_toi(dptr(q), earliest_toi, toi_tol, ctypes.byref(toi), ctypes.byref(n))
    return (hit if hit < 0 else bool(hit)), toi.value, n.value


def pl_ee_toi(q, earliest_toi=1.0, msd=0.0):
    """Piecewise-linear TOI of a packed EE query: (hit, toi, pieces, boxes)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    toi = ctypes.c_double()
    npc, nbx = ctypes.c_long(0), ctypes.c_long(0)
    hit = lib().oracle_pl_ee_toi(dptr(q), float(earliest_toi), float(msd), ctypes.byref(toi), ctypes.byref(npc),
                                 ctypes.byref(nbx))
    return bool(hit), toi.value, npc.value, nbx.value


def pl_fv_toi(q, earliest_toi=1.0, msd=0.0):
    """Piecewise-linear TOI of a packed FV query: (hit, toi, pieces, boxes)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    toi = ctypes.c_double()
    npc, nbx = ctypes.c_long(0), ctypes.c_long(0)
    hit = lib().oracle_pl_fv_toi(dptr(q), float(earliest_toi), float(msd), ctypes.byref(toi), ctypes.byref(npc),
                                 ctypes.byref(nbx))
    return bool(hit), toi.value, npc.value, nbx.value


def ti_linear(V, vf, ms=0.0):
    """Tight-Inclusion linear CCD of 8 vertices (start then end):
    (hit, toi, output_tolerance, boxes)."""
    V = np.ascontiguousarray(V, dtype=np.float64).reshape(8, 3)
    toi, tol = ctypes.c_double(), ctypes.c_double()
    nbx = ctypes.c_long(0)
    fn = lib().oracle_ti_vertex_face if vf else lib().oracle_ti_edge_edge
    hit = fn(dptr(V), float(ms), ctypes.byref(toi), ctypes.byref(tol), ctypes.byref(nbx))
    return bool(hit), toi.value, tol.value, nbx.value


def element_distance_sq(kind, X):
    """Squared EE (kind 0: ea0 ea1 eb0 eb1) or FV (kind 1: v f0 f1 f2) distance."""
    X = np.ascontiguousarray(X, dtype=np.float64).reshape(4, 3)
    return lib().oracle_element_distance_sq(2 if kind == 0 else 3, dptr(X))


def set_eval_budget(n):
    """Cap inclusion-function evaluations per query (0 = unbounded)."""
    lib().oracle_set_eval_budget(int(n))


def set_scene(scene):
    """Load a scene (same fields as the product's Context.set_scene) into the oracle."""
    a = dict(
        V=np.ascontiguousarray(scene.vertices, np.float64), E=np.ascontiguousarray(scene.edges, np.int32),
        F=np.ascontiguousarray(scene.faces, np.int32), FE=np.ascontiguousarray(scene.face_edges, np.int32),
        bv=np.ascontiguousarray(scene.body_vertex_start, np.int32),
        be=np.ascontiguousarray(scene.body_edge_start, np.int32),
        bf=np.ascontiguousarray(scene.body_face_start, np.int32),
        g=np.ascontiguousarray(scene.group_ids, np.int32), st=np.ascontiguousarray(scene.is_static, np.uint8))
    global _scene_keep
    _scene_keep = a
    lib().oracle_set_scene(len(a["g"]), len(a["V"]), len(a["E"]), len(a["F"]), dptr(a["V"]), i32ptr(a["E"]),
                           i32ptr(a["F"]), i32ptr(a["FE"]), i32ptr(a["bv"]), i32ptr(a["be"]), i32ptr(a["bf"]),
                           i32ptr(a["g"]), a["st"].ctypes.data_as(ctypes.c_void_p))


def broadphase(poses_t0, poses_t1=None, inflation_radius=0.0):
    """Continuous (poses_t1 given) or discrete broad phase; returns (ee, fv) int32 arrays."""
    p0 = np.ascontiguousarray(poses_t0, np.float64)
    cont = poses_t1 is not None
    p1 = np.ascontiguousarray(poses_t1 if cont else poses_t0, np.float64)
    ne, nf = ctypes.c_long(), ctypes.c_long()
    lib().oracle_broadphase(int(cont), dptr(p0), dptr(p1), float(inflation_radius), ctypes.byref(ne), ctypes.byref(nf))
    ee = np.zeros((ne.value, 2), np.int32)
    fv = np.zeros((nf.value, 2), np.int32)
    lib().oracle_get_candidates(i32ptr(ee), i32ptr(fv))
    return ee, fv


def set_detection_method(method):
    """The oracle's broad phase and earliest-TOI candidates: "hash_grid" or
    "bvh" (DetectionMethod, src/ccd/ccd.hpp:20-24)."""
    lib().oracle_set_detection_method({"hash_grid": 0, "bvh": 1}[method])


def bvh_supported():
    return bool(lib().oracle_bvh_supported())


def orient3d(a, b, c, d):
    """Exact sign of det [a - d; b - d; c - d]."""
    pts = [np.ascontiguousarray(x, np.float64) for x in (a, b, c, d)]
    return lib().oracle_orient3d(*[dptr(x) for x in pts])


def edge_triangle(p, q, t0, t1, t2):
    pts = [np.ascontiguousarray(x, np.float64) for x in (p, q, t0, t1, t2)]
    return bool(lib().oracle_edge_triangle(*[dptr(x) for x in pts]))


def detect_intersections(poses, capacity=1 << 18):
    """Sorted intersecting (edge, face) pairs at poses (current scene); one
    oracle run unless more than `capacity` pairs intersect."""
    p = np.ascontiguousarray(poses, np.float64)
    out = np.zeros((capacity, 2), np.int32)
    n = lib().oracle_detect_intersections(dptr(p), i32ptr(out), capacity)
    if n > capacity:
        out = np.zeros((n, 2), np.int32)
        lib().oracle_detect_intersections(dptr(p), i32ptr(out), n)
    return out[:n].copy()


def world_vertices(poses, nv):
    out = np.zeros((nv, 3))
    lib().oracle_world_vertices(dptr(np.ascontiguousarray(poses, np.float64)), dptr(out))
    return out


def last_vertex_aabbs(nv):
    """(included bool (nv,), min (nv,3), max (nv,3)) of the last broad phase."""
    out = np.zeros((nv, 7))
    lib().oracle_last_vertex_aabbs(dptr(out))
    return out[:, 0] > 0, out[:, 1:4], out[:, 4:7]


def earliest_toi(poses_t0, poses_t1, msd=0.0):
    """(earliest toi, number of candidates) on the current scene."""
    n = ctypes.c_long()
    t = lib().oracle_earliest_toi(dptr(np.ascontiguousarray(poses_t0, np.float64)),
                                  dptr(np.ascontiguousarray(poses_t1, np.float64)), float(msd), ctypes.byref(n))
    return t, n.value


def _declare_barrier():
    L = lib()
    if getattr(L, "_barrier_declared", False):
        return L
    L.oracle_set_psd_projection.argtypes = [ctypes.c_int]
    L.oracle_constraint_set.argtypes = [c_double_p, ctypes.c_double, ctypes.c_double, c_long_p]
    L.oracle_get_constraints.argtypes = [c_long_p, c_double_p]
    L.oracle_barrier.argtypes = [c_double_p, ctypes.c_double, ctypes.c_double, c_double_p, c_double_p]
    L.oracle_barrier.restype = ctypes.c_double
    L.oracle_min_distance.argtypes = [c_double_p]
    L.oracle_min_distance.restype = ctypes.c_double
    L.oracle_friction_update.argtypes = [c_double_p] + [ctypes.c_double] * 4
    L.oracle_friction_update.restype = ctypes.c_long
    L.oracle_get_friction.argtypes = [c_double_p]
    L.oracle_friction.argtypes = [c_double_p, c_double_p, ctypes.c_double, c_double_p, c_double_p]
    L.oracle_friction.restype = ctypes.c_double
    L.oracle_vertex_derivatives.argtypes = [c_double_p, ctypes.c_int, c_double_p, c_double_p, c_double_p]
    L.oracle_friction_adjoint.argtypes = [c_double_p, c_double_p] + [ctypes.c_double] * 4 + [c_double_p, c_double_p]
    L._barrier_declared = True
    return L


def set_psd_projection(on):
    _declare_barrier().oracle_set_psd_projection(int(bool(on)))


def constraint_set(poses, dhat, dmin=0.0):
    """Build the constraint set at `poses`; returns (records (n,6) int64:
    type, a, b, multiplicity, body0, body1; eps_x (n,))."""
    L = _declare_barrier()
    counts = np.zeros(4, np.int64)
    L.oracle_constraint_set(dptr(np.ascontiguousarray(poses, np.float64)), float(dhat), float(dmin),
                            counts.ctypes.data_as(c_long_p))
    n = int(counts.sum())
    rec = np.zeros((n, 6), np.int64)
    eps = np.zeros(n)
    L.oracle_get_constraints(rec.ctypes.data_as(c_long_p), dptr(eps))
    return rec, eps


def barrier(poses, dhat, dmin=0.0, nb=None, want_hess=False, n_constraints=0):
    L = _declare_barrier()
    p = np.ascontiguousarray(poses, np.float64)
    g = np.zeros(p.size)
    h = np.zeros((max(n_constraints, 1), 144)) if want_hess else None
    e = L.oracle_barrier(dptr(p), float(dhat), float(dmin), dptr(g), dptr(h) if want_hess else None)
    return e, g, (h[:n_constraints] if want_hess else None)


def min_distance(poses):
    return _declare_barrier().oracle_min_distance(dptr(np.ascontiguousarray(poses, np.float64)))


def friction_update(poses, dhat, dmin, kappa, mu):
    L = _declare_barrier()
    n = L.oracle_friction_update(dptr(np.ascontiguousarray(poses, np.float64)), float(dhat), float(dmin),
                                 float(kappa), float(mu))
    rec = np.zeros((max(n, 1), 9))
    L.oracle_get_friction(dptr(rec))
    return rec[:n]


def friction(poses, poses_t0, epsv_times_h, want_hess=False, n_constraints=0):
    L = _declare_barrier()
    p = np.ascontiguousarray(poses, np.float64)
    g = np.zeros(p.size)
    h = np.zeros((max(n_constraints, 1), 144)) if want_hess else None
    e = L.oracle_friction(dptr(p), dptr(np.ascontiguousarray(poses_t0, np.float64)), float(epsv_times_h), dptr(g),
                          dptr(h) if want_hess else None)
    return e, g, (h[:n_constraints] if want_hess else None)


def vertex_derivatives(pose, vertex):
    L = _declare_barrier()
    w, J, H = np.zeros(3), np.zeros(18), np.zeros(27)
    L.oracle_vertex_derivatives(dptr(np.ascontiguousarray(pose, np.float64)), int(vertex), dptr(w), dptr(J), dptr(H))
    return w, J.reshape(3, 6), H.reshape(3, 3, 3)


def assemble_hessian(rec, local_hess, nb):
    """Sum per-constraint 12x12 blocks into a dense (6nb x 6nb) matrix."""
    Hd = np.zeros((6 * nb, 6 * nb))
    for r, h in zip(rec, local_hess):
        h = h.reshape(12, 12)
        bodies = (r[4], r[5])
        for i in range(2):
            for j in range(2):
                Hd[6 * bodies[i]:6 * bodies[i] + 6, 6 * bodies[j]:6 * bodies[j] + 6] += h[6 * i:6 * i + 6, 6 * j:6 * j + 6]
    return Hd


def friction_adjoint(poses, poses_t0, epsv_times_h, dhat, dmin, kappa, lam):
    """(d grad_x D(poses; poses_t0) / d poses_t0)^T lam over the current
    friction set (dual numbers)."""
    L = _declare_barrier()
    p = np.ascontiguousarray(poses, np.float64)
    out = np.zeros(p.size)
    L.oracle_friction_adjoint(dptr(p), dptr(np.ascontiguousarray(poses_t0, np.float64)), float(epsv_times_h),
                              float(dhat), float(dmin), float(kappa), dptr(np.ascontiguousarray(lam, np.float64)),
                              dptr(out))
    return out
