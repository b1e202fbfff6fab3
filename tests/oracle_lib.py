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

_lib = None

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
        if not os.path.exists(ORACLE_SO):
            subprocess.check_call(["make", "-C", ORACLE_DIR, "-j8"])
        _lib = ctypes.CDLL(ORACLE_SO)
        L = _lib
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
        L.oracle_set_eval_budget.argtypes = [ctypes.c_long]
        L.oracle_set_scene.argtypes = [ctypes.c_int] * 4 + [c_double_p] + [c_int32_p] * 7 + [ctypes.c_void_p]
        L.oracle_broadphase.argtypes = [ctypes.c_int, c_double_p, c_double_p, ctypes.c_double, c_long_p, c_long_p]
        L.oracle_get_candidates.argtypes = [c_int32_p, c_int32_p]
        L.oracle_average_edge_length.restype = ctypes.c_double
        L.oracle_world_vertices.argtypes = [c_double_p, c_double_p]
        L.oracle_last_vertex_aabbs.argtypes = [c_double_p]
        L.oracle_earliest_toi.argtypes = [c_double_p, c_double_p, ctypes.c_double, c_long_p]
        L.oracle_earliest_toi.restype = ctypes.c_double
        for name in ("oracle_ee_aabb", "oracle_fv_aabb"):
            getattr(L, name).argtypes = [c_double_p, c_double_p, c_double_p]
    return _lib


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
