"""Python bindings for librigc, the MI355X contact hot path of Rigid IPC.

The module loads ``rigid-ipc_amd/librigc.so`` (built by
``rigid-ipc_amd/build.py``) through its C ABI (``include/rigc.h``). There is
no CPU fallback: if the library or a GPU is missing, construction raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "librigc.so")

RIGC_OK = 0
RIGC_ERR_INVALID_ARGUMENT = 1
RIGC_ERR_DEVICE = 2
RIGC_ERR_NO_SCENE = 3
RIGC_ERR_CAPACITY = 4
RIGC_ERR_UNRESOLVED = 5

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_int32_p = ctypes.POINTER(ctypes.c_int32)
_c_int8_p = ctypes.POINTER(ctypes.c_int8)
_c_uint8_p = ctypes.POINTER(ctypes.c_uint8)
_c_int64_p = ctypes.POINTER(ctypes.c_int64)


class RigcError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"rigc status {status}: {message}")
        self.status = status


class Scene(ctypes.Structure):
    _fields_ = [
        ("num_bodies", ctypes.c_int32), ("num_vertices", ctypes.c_int32),
        ("num_edges", ctypes.c_int32), ("num_faces", ctypes.c_int32),
        ("vertices", _c_double_p), ("edges", _c_int32_p), ("faces", _c_int32_p),
        ("face_edges", _c_int32_p), ("body_vertex_start", _c_int32_p),
        ("body_edge_start", _c_int32_p), ("body_face_start", _c_int32_p),
        ("group_ids", _c_int32_p), ("is_static", _c_uint8_p),
    ]


RIGC_GRADIENT = 1
RIGC_HESSIAN = 2
RIGC_BUILD_CONSTRAINTS = 4
RIGC_UNPROJECTED = 8


class BlockHessian(ctypes.Structure):
    _fields_ = [("diag", _c_double_p), ("pairs", _c_int32_p), ("off", _c_double_p),
                ("capacity", ctypes.c_int64), ("count", ctypes.c_int64)]


class Tuning(ctypes.Structure):
    """rigc_tuning: schedule knobs of the rigid CCD (no effect on results)."""
    _fields_ = [("ccd_walk_budget", ctypes.c_int32), ("ccd_spec_levels", ctypes.c_int32),
                ("ccd_window_factor", ctypes.c_double), ("ccd_wide_queries", ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("broad_phase_ms", ctypes.c_double), ("narrow_phase_ms", ctypes.c_double),
        ("assembly_ms", ctypes.c_double), ("box_evaluations", ctypes.c_int64),
        ("num_candidates", ctypes.c_int64), ("num_constraints", ctypes.c_int64),
        ("ccd_kernel_ms", ctypes.c_double), ("ccd_launches", ctypes.c_int64),
    ]


_lib = None


def lib():
    """Load librigc.so and declare its entry points (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run rigid-ipc_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.rigc_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32]
    L.rigc_destroy.argtypes = [vp]
    L.rigc_last_error.argtypes = [vp]
    L.rigc_last_error.restype = ctypes.c_char_p
    L.rigc_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
    L.rigc_set_scene.argtypes = [vp, ctypes.POINTER(Scene)]
    L.rigc_narrow_phase.argtypes = [vp, _c_double_p, _c_double_p, _c_int32_p, ctypes.c_int64,
                                    _c_int32_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
                                    ctypes.c_int32, _c_double_p, _c_double_p, _c_int8_p]
    L.rigc_narrow_phase_noalloc.argtypes = [vp, _c_double_p, _c_double_p, _c_int32_p, ctypes.c_int64, _c_int32_p,
                                            ctypes.c_int64, ctypes.c_double, ctypes.c_int32, _c_double_p,
                                            _c_double_p, _c_int8_p]
    L.rigc_broad_phase.argtypes = [vp, _c_double_p, _c_double_p, ctypes.c_double, _c_int32_p, ctypes.c_int64,
                                   _c_int64_p, _c_int32_p, ctypes.c_int64, _c_int64_p]
    L.rigc_compute_earliest_toi.argtypes = [vp, _c_double_p, _c_double_p, ctypes.c_double, _c_double_p]
    L.rigc_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.rigc_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.rigc_autotune_earliest_toi.argtypes = [vp, ctypes.c_int32, _c_double_p, _c_double_p, ctypes.c_double,
                                             ctypes.c_int32, ctypes.POINTER(Tuning), _c_double_p]
    L.rigc_set_detection_method.argtypes = [vp, ctypes.c_int32]
    for name in ("rigc_get_tuning", "rigc_set_tuning", "rigc_autotune_earliest_toi", "rigc_set_detection_method"):
        getattr(L, name).restype = ctypes.c_int32
    L.rigc_narrow_phase_piecewise_linear.argtypes = [vp, _c_double_p, _c_double_p, _c_int32_p, ctypes.c_int64,
                                                     _c_int32_p, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                                     _c_double_p, _c_double_p, _c_int8_p]
    L.rigc_compute_earliest_toi_piecewise_linear.argtypes = [vp, _c_double_p, _c_double_p, ctypes.c_double,
                                                             _c_double_p]
    L.rigc_detect_intersections.argtypes = [vp, _c_double_p, _c_int32_p, ctypes.c_int64, _c_int64_p]
    d, i32, i64 = ctypes.c_double, ctypes.c_int32, ctypes.c_int64
    bh = ctypes.POINTER(BlockHessian)
    L.rigc_construct_constraint_set.argtypes = [vp, _c_double_p, d, d, _c_int64_p]
    L.rigc_get_constraint_set.argtypes = [vp, i64, _c_int32_p, _c_double_p, _c_int64_p]
    L.rigc_compute_barrier_term.argtypes = [vp, _c_double_p, d, d, i32, _c_double_p, _c_double_p, bh]
    L.rigc_compute_minimum_distance.argtypes = [vp, _c_double_p, d, d, i32, _c_double_p]
    L.rigc_update_friction_constraints.argtypes = [vp, _c_double_p, d, d, d, d, i32, _c_int64_p]
    L.rigc_get_friction_constraints.argtypes = [vp, i64, _c_double_p, _c_int64_p]
    L.rigc_compute_friction_term.argtypes = [vp, _c_double_p, _c_double_p, d, i32, _c_double_p, _c_double_p, bh]
    L.rigc_friction_adjoint.argtypes = [vp, _c_double_p, _c_double_p, d, d, d, d, _c_double_p, _c_double_p]
    L.rigc_set_scenes.argtypes = [vp, i32, ctypes.POINTER(Scene)]
    L.rigc_get_scene_count.argtypes = [vp, _c_int32_p]
    L.rigc_batch_earliest_toi.argtypes = [vp, i32, _c_int32_p, _c_double_p, _c_double_p, d, _c_double_p]
    L.rigc_batch_ccd_evaluations.argtypes = [vp, i32, _c_int32_p, ctypes.POINTER(ctypes.c_int64)]
    L.rigc_batch_set_previous_toi.argtypes = [vp, i32, _c_int32_p, _c_double_p]
    L.rigc_batch_barrier_term.argtypes = [vp, i32, _c_int32_p, _c_double_p, _c_double_p, d, d, i32, _c_double_p,
                                          _c_double_p, bh, _c_int64_p, _c_double_p]
    L.rigc_batch_minimum_distance.argtypes = [vp, i32, _c_int32_p, _c_double_p, _c_double_p, d, d, _c_double_p]
    L.rigc_batch_detect_intersections.argtypes = [vp, i32, _c_int32_p, _c_double_p, _c_int64_p]
    L.rigc_batch_update_friction_constraints.argtypes = [vp, i32, _c_int32_p, _c_double_p, d, d, _c_double_p,
                                                         _c_double_p, _c_int64_p]
    L.rigc_batch_friction_term.argtypes = [vp, i32, _c_int32_p, _c_double_p, _c_double_p, _c_double_p, i32,
                                           _c_double_p, _c_double_p, bh]
    for name in ("rigc_set_scenes", "rigc_get_scene_count", "rigc_batch_earliest_toi", "rigc_batch_ccd_evaluations",
                 "rigc_batch_set_previous_toi",
                 "rigc_batch_barrier_term",
                 "rigc_batch_minimum_distance", "rigc_batch_update_friction_constraints",
                 "rigc_batch_detect_intersections",
                 "rigc_batch_friction_term"):
        getattr(L, name).restype = ctypes.c_int32
    for name in ("rigc_create", "rigc_destroy", "rigc_get_stats", "rigc_set_scene", "rigc_narrow_phase",
                 "rigc_broad_phase", "rigc_compute_earliest_toi", "rigc_construct_constraint_set",
                 "rigc_get_constraint_set", "rigc_compute_barrier_term", "rigc_compute_minimum_distance",
                 "rigc_update_friction_constraints", "rigc_get_friction_constraints",
                 "rigc_compute_friction_term", "rigc_friction_adjoint", "rigc_detect_intersections",
                 "rigc_narrow_phase_piecewise_linear", "rigc_compute_earliest_toi_piecewise_linear"):
        getattr(L, name).restype = ctypes.c_int32
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


class Context:
    """One device context (rigc_ctx) bound to one scene at a time."""

    def __init__(self, device=0):
        self._L = lib()
        self._ctx = ctypes.c_void_p()
        st = self._L.rigc_create(ctypes.byref(self._ctx), int(device))
        if st != RIGC_OK:
            raise RigcError(st, "rigc_create failed (no usable HIP device?)")
        self._keep = []

    def close(self):
        if self._ctx:
            self._L.rigc_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, st):
        if st != RIGC_OK:
            raise RigcError(st, self._L.rigc_last_error(self._ctx).decode())

    def stats(self):
        s = Stats()
        self._check(self._L.rigc_get_stats(self._ctx, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    @staticmethod
    def _scene_struct(scene, keep):
        arrs = dict(
            vertices=np.ascontiguousarray(scene.vertices, np.float64),
            edges=np.ascontiguousarray(scene.edges, np.int32),
            faces=np.ascontiguousarray(scene.faces, np.int32),
            face_edges=np.ascontiguousarray(scene.face_edges, np.int32),
            body_vertex_start=np.ascontiguousarray(scene.body_vertex_start, np.int32),
            body_edge_start=np.ascontiguousarray(scene.body_edge_start, np.int32),
            body_face_start=np.ascontiguousarray(scene.body_face_start, np.int32),
            group_ids=np.ascontiguousarray(scene.group_ids, np.int32),
            is_static=np.ascontiguousarray(scene.is_static, np.uint8),
        )
        s = Scene()
        s.num_bodies = len(arrs["group_ids"])
        s.num_vertices = len(arrs["vertices"])
        s.num_edges = len(arrs["edges"])
        s.num_faces = len(arrs["faces"])
        for k, a in arrs.items():
            t = {np.float64: _c_double_p, np.int32: _c_int32_p, np.uint8: _c_uint8_p}[a.dtype.type]
            setattr(s, k, _p(a, t))
        keep.append(arrs)
        return s

    def set_scene(self, scene):
        """scene: an object with numpy fields vertices (nv,3) f64, edges (ne,2),
        faces (nf,3), face_edges (nf,3), body_vertex_start, body_edge_start,
        body_face_start (nb+1), group_ids (nb), is_static (nb)."""
        keep = []
        s = self._scene_struct(scene, keep)
        self._check(self._L.rigc_set_scene(self._ctx, ctypes.byref(s)))
        self.num_bodies = s.num_bodies
        self.scene_bodies = [s.num_bodies]

    def set_scenes(self, scenes):
        """Several independent scenes in one context (rigc_set_scenes)."""
        keep = []
        arr = (Scene * len(scenes))(*[self._scene_struct(sc, keep) for sc in scenes])
        self._check(self._L.rigc_set_scenes(self._ctx, len(scenes), arr))
        self.scene_bodies = [int(len(sc.group_ids)) for sc in scenes]
        self.num_bodies = sum(self.scene_bodies)

    # ---------------------------------------------------------- batched scenes
    def _ids_poses(self, ids, *pose_lists):
        ids = np.ascontiguousarray(ids, np.int32)
        out = [ids]
        for pl in pose_lists:
            out.append(np.ascontiguousarray(np.concatenate([np.asarray(p, np.float64).reshape(-1, 6) for p in pl])))
        return out

    def batch_earliest_toi(self, ids, poses_t0, poses_t1, minimum_separation_distance=0.0):
        ids, p0, p1 = self._ids_poses(ids, poses_t0, poses_t1)
        toi = np.zeros(len(ids))
        self._check(self._L.rigc_batch_earliest_toi(self._ctx, len(ids), _p(ids, _c_int32_p), _p(p0, _c_double_p),
                                                    _p(p1, _c_double_p), float(minimum_separation_distance),
                                                    _p(toi, _c_double_p)))
        return toi

    def batch_set_previous_toi(self, ids, toi):
        """rigc_batch_set_previous_toi: each scene's previous earliest TOI
        (the bound of the first pass of its next call on this context)."""
        ids = np.ascontiguousarray(np.asarray(ids, np.int32))
        toi = np.ascontiguousarray(np.asarray(toi, np.float64))
        self._check(self._L.rigc_batch_set_previous_toi(self._ctx, len(ids), _p(ids, _c_int32_p),
                                                        _p(toi, _c_double_p)))

    def batch_ccd_evaluations(self, ids):
        """Box evaluations per scene of the last earliest-TOI call of this
        multi-scene context (0 for scenes it did not serve)."""
        ids = np.ascontiguousarray(np.asarray(ids, np.int32))
        out = np.zeros(len(ids), np.int64)
        self._check(self._L.rigc_batch_ccd_evaluations(self._ctx, len(ids), _p(ids, _c_int32_p),
                                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        return out

    def _batch_potential(self, fn, ids, args, flags):
        n = len(ids)
        nbs = [self.scene_bodies[i] for i in ids]
        starts = np.concatenate([[0], np.cumsum(nbs)]).astype(int)
        nb = int(starts[-1])
        energy = np.zeros(n)
        grad = np.zeros(6 * nb) if flags & (RIGC_GRADIENT | RIGC_HESSIAN) else None
        h = None
        diag = pairs = off = None
        if flags & RIGC_HESSIAN:
            cap = getattr(self, "_cap_off", 1024)
            for _ in range(2):
                diag = np.zeros((nb, 36))
                pairs = np.zeros((max(cap, 1), 2), np.int32)
                off = np.zeros((max(cap, 1), 36))
                h = BlockHessian(_p(diag, _c_double_p), _p(pairs, _c_int32_p), _p(off, _c_double_p), cap, 0)
                st = fn(self._ctx, n, _p(ids, _c_int32_p), *args, int(flags), _p(energy, _c_double_p),
                        _p(grad, _c_double_p), ctypes.byref(h))
                if st == RIGC_ERR_CAPACITY:
                    cap = self._cap_off = int(h.count)
                    continue
                self._check(st)
                break
            else:
                raise RigcError(RIGC_ERR_CAPACITY, "Hessian capacity retry failed")
        else:
            self._check(fn(self._ctx, n, _p(ids, _c_int32_p), *args, int(flags), _p(energy, _c_double_p),
                           _p(grad, _c_double_p), None))
        out = []
        for i in range(n):
            a, b = starts[i], starts[i + 1]
            g = grad[6 * a:6 * b].copy() if grad is not None else None
            hh = None
            if h is not None:
                P = pairs[:h.count]
                m = (P[:, 0] >= a) & (P[:, 0] < b)
                hh = (diag[a:b].reshape(-1, 6, 6).copy(), (P[m] - a).astype(np.int32),
                      off[:h.count][m].reshape(-1, 6, 6).copy())
            out.append((float(energy[i]), g, hh))
        return out

    def batch_barrier_term(self, ids, build_poses, poses, dhat, dmin=0.0, grad=False, hess=False):
        """Per-scene (energy, grad, hess) with each scene's constraint set
        built at build_poses[i] and evaluated at poses[i]."""
        ids, pb, p = self._ids_poses(ids, build_poses, poses)
        flags = (RIGC_GRADIENT if grad else 0) | (RIGC_HESSIAN if hess else 0)
        self._keep = [ids, pb, p]
        fn = lambda ctx, n, i, a0, a1, a2, a3, f, e, g, hh: self._L.rigc_batch_barrier_term(  # noqa: E731
            ctx, n, i, a0, a1, a2, a3, f, e, g, hh, None, None)
        return self._batch_potential(fn, ids, (_p(pb, _c_double_p), _p(p, _c_double_p), float(dhat), float(dmin)),
                                     flags)

    def batch_constraint_counts(self, ids, poses, dhat, dmin=0.0):
        ids, p = self._ids_poses(ids, poses)
        e = np.zeros(len(ids))
        counts = np.zeros((len(ids), 4), np.int64)
        self._check(self._L.rigc_batch_barrier_term(self._ctx, len(ids), _p(ids, _c_int32_p), _p(p, _c_double_p),
                                                    _p(p, _c_double_p), float(dhat), float(dmin), 0,
                                                    _p(e, _c_double_p), None, None, _p(counts, _c_int64_p), None))
        return counts

    def batch_barrier_energy_and_distance(self, ids, build_poses, poses, dhat, dmin=0.0):
        """Per-scene barrier energy and minimum distance from one
        rigc_batch_barrier_term call."""
        ids, pb, p = self._ids_poses(ids, build_poses, poses)
        e = np.zeros(len(ids))
        md = np.zeros(len(ids))
        self._check(self._L.rigc_batch_barrier_term(self._ctx, len(ids), _p(ids, _c_int32_p), _p(pb, _c_double_p),
                                                    _p(p, _c_double_p), float(dhat), float(dmin), 0,
                                                    _p(e, _c_double_p), None, None, None, _p(md, _c_double_p)))
        return e, md

    def batch_detect_intersections(self, ids, poses):
        """Per-scene intersecting edge-face pair counts (rigc_batch_detect_intersections)."""
        ids, p = self._ids_poses(ids, poses)
        counts = np.zeros(len(ids), np.int64)
        self._check(self._L.rigc_batch_detect_intersections(self._ctx, len(ids), _p(ids, _c_int32_p),
                                                            _p(p, _c_double_p), _p(counts, _c_int64_p)))
        return counts

    def batch_minimum_distance(self, ids, build_poses, poses, dhat, dmin=0.0):
        ids, pb, p = self._ids_poses(ids, build_poses, poses)
        out = np.zeros(len(ids))
        self._check(self._L.rigc_batch_minimum_distance(self._ctx, len(ids), _p(ids, _c_int32_p), _p(pb, _c_double_p),
                                                        _p(p, _c_double_p), float(dhat), float(dmin),
                                                        _p(out, _c_double_p)))
        return out

    def batch_update_friction_constraints(self, ids, poses, dhat, dmin, barrier_stiffness, mu):
        ids, p = self._ids_poses(ids, poses)
        kap = np.ascontiguousarray(barrier_stiffness, np.float64)
        mus = np.ascontiguousarray(mu, np.float64)
        counts = np.zeros(len(ids), np.int64)
        self._check(self._L.rigc_batch_update_friction_constraints(
            self._ctx, len(ids), _p(ids, _c_int32_p), _p(p, _c_double_p), float(dhat), float(dmin),
            _p(kap, _c_double_p), _p(mus, _c_double_p), _p(counts, _c_int64_p)))
        return counts

    def batch_friction_term(self, ids, poses, poses_t0, epsv_times_h, grad=False, hess=False):
        ids, p, p0 = self._ids_poses(ids, poses, poses_t0)
        eh = np.ascontiguousarray(epsv_times_h, np.float64)
        flags = (RIGC_GRADIENT if grad else 0) | (RIGC_HESSIAN if hess else 0)
        self._keep = [ids, p, p0, eh]
        return self._batch_potential(self._L.rigc_batch_friction_term, ids,
                                     (_p(p, _c_double_p), _p(p0, _c_double_p), _p(eh, _c_double_p)), flags)

    def narrow_phase_enqueue(self, stream, poses_t0, poses_t1, ee, n_ee, fv, n_fv, out, toi_tolerance=1e-4):
        """rigc_narrow_phase_enqueue on `stream` (a hipStream_t handle, e.g.
        torch.cuda.Stream().cuda_stream); the other pointer arguments are
        device addresses (e.g. tensor.data_ptr())."""
        L = self._L
        L.rigc_narrow_phase_enqueue.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                                ctypes.c_double, ctypes.c_void_p]
        L.rigc_narrow_phase_enqueue.restype = ctypes.c_int32
        self._check(L.rigc_narrow_phase_enqueue(self._ctx, stream, poses_t0, poses_t1, ee, int(n_ee), fv, int(n_fv),
                                                float(toi_tolerance), out))

    def narrow_phase(self, poses_t0, poses_t1, ee, fv, toi_tolerance=1e-4, eval_budget=0,
                     per_query=False, allow_unresolved=False):
        """Rigid CCD over candidates; returns (earliest_toi, toi[n], hit[n])."""
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        p1 = np.ascontiguousarray(poses_t1, np.float64).reshape(-1, 6)
        ee = np.ascontiguousarray(np.asarray(ee, np.int32).reshape(-1, 2))
        fv = np.ascontiguousarray(np.asarray(fv, np.int32).reshape(-1, 2))
        n = len(ee) + len(fv)
        toi = np.empty(max(n, 1))
        hit = np.empty(max(n, 1), np.int8)
        best = ctypes.c_double()
        st = self._L.rigc_narrow_phase(self._ctx, _p(p0, _c_double_p), _p(p1, _c_double_p),
                                       _p(ee, _c_int32_p), len(ee), _p(fv, _c_int32_p), len(fv),
                                       float(toi_tolerance), int(eval_budget), int(bool(per_query)),
                                       ctypes.byref(best), _p(toi, _c_double_p), _p(hit, _c_int8_p))
        if not (allow_unresolved and st == RIGC_ERR_UNRESOLVED):
            self._check(st)
        return best.value, toi[:n], hit[:n]

    def narrow_phase_noalloc(self, poses_t0, poses_t1, ee, fv, toi_tolerance=1e-4, per_query=False):
        """rigc_narrow_phase_noalloc: rigid CCD over candidates with every
        per-call buffer in the library's static device array; returns
        (earliest_toi, toi[n], hit[n])."""
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        p1 = np.ascontiguousarray(poses_t1, np.float64).reshape(-1, 6)
        ee = np.ascontiguousarray(np.asarray(ee, np.int32).reshape(-1, 2))
        fv = np.ascontiguousarray(np.asarray(fv, np.int32).reshape(-1, 2))
        n = len(ee) + len(fv)
        toi = np.empty(max(n, 1))
        hit = np.empty(max(n, 1), np.int8)
        best = ctypes.c_double()
        self._check(self._L.rigc_narrow_phase_noalloc(self._ctx, _p(p0, _c_double_p), _p(p1, _c_double_p),
                                                      _p(ee, _c_int32_p), len(ee), _p(fv, _c_int32_p), len(fv),
                                                      float(toi_tolerance), int(bool(per_query)), ctypes.byref(best),
                                                      _p(toi, _c_double_p), _p(hit, _c_int8_p)))
        return best.value, toi[:n], hit[:n]

    def broad_phase(self, poses_t0, poses_t1=None, inflation_radius=0.0):
        """Hash-grid candidates, continuous (poses_t1 given) or discrete;
        returns sorted (ee (n,2), fv (n,2)) int32 arrays."""
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        p1 = None if poses_t1 is None else np.ascontiguousarray(poses_t1, np.float64).reshape(-1, 6)
        cap_ee, cap_fv = getattr(self, "_cap_ee", 1 << 16), getattr(self, "_cap_fv", 1 << 16)
        for _ in range(2):
            ee = np.empty((cap_ee, 2), np.int32)
            fv = np.empty((cap_fv, 2), np.int32)
            ne, nf = ctypes.c_int64(), ctypes.c_int64()
            st = self._L.rigc_broad_phase(self._ctx, _p(p0, _c_double_p), _p(p1, _c_double_p),
                                          float(inflation_radius), _p(ee, _c_int32_p), cap_ee, ctypes.byref(ne),
                                          _p(fv, _c_int32_p), cap_fv, ctypes.byref(nf))
            if st == RIGC_ERR_CAPACITY:
                cap_ee, cap_fv = max(cap_ee, ne.value), max(cap_fv, nf.value)
                self._cap_ee, self._cap_fv = cap_ee, cap_fv
                continue
            self._check(st)
            return ee[:ne.value].copy(), fv[:nf.value].copy()
        raise RigcError(RIGC_ERR_CAPACITY, "broad phase capacity retry failed")

    def detect_intersections(self, poses):
        """Edge-triangle intersections between close bodies at poses; returns
        the sorted (edge, face) int32 pairs (n, 2). The reference's
        RigidBodyProblem::detect_intersections is len(result) > 0."""
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        cap = getattr(self, "_cap_isect", 1024)
        for _ in range(2):
            out = np.empty((cap, 2), np.int32)
            n = ctypes.c_int64()
            st = self._L.rigc_detect_intersections(self._ctx, _p(p, _c_double_p), _p(out, _c_int32_p), cap,
                                                   ctypes.byref(n))
            if st == RIGC_ERR_CAPACITY:
                cap = self._cap_isect = max(cap, n.value)
                continue
            self._check(st)
            return out[:n.value].copy()
        raise RigcError(RIGC_ERR_CAPACITY, "intersection capacity retry failed")

    def count_intersections(self, poses):
        """Number of intersecting (edge, face) pairs at poses (no list copied)."""
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        n = ctypes.c_int64()
        self._check(self._L.rigc_detect_intersections(self._ctx, _p(p, _c_double_p), None, 0, ctypes.byref(n)))
        return n.value

    def set_detection_method(self, method):
        """rigc_set_detection_method: "hash_grid" or "bvh" (the broad phase of
        broad_phase() and of the earliest-TOI calls)."""
        self._check(self._L.rigc_set_detection_method(self._ctx, {"hash_grid": 0, "bvh": 1}[method]))

    def tuning(self):
        """The CCD tunables in force (rigc_get_tuning) as a dict."""
        t = Tuning()
        self._check(self._L.rigc_get_tuning(self._ctx, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        """rigc_set_tuning with the given fields changed."""
        cur = self.tuning()
        cur.update(kw)
        t = Tuning(**cur)
        self._check(self._L.rigc_set_tuning(self._ctx, ctypes.byref(t)))

    def autotune_earliest_toi(self, samples, minimum_separation_distance=0.0, knobs=None):
        """rigc_autotune_earliest_toi over samples = [(poses_t0, poses_t1), ...]
        (poses of the whole context); knobs: names of the rigc_tuning fields
        to tune (default all); returns (tuning dict, device ms)."""
        bits = {"ccd_spec_levels": 1, "ccd_walk_budget": 2, "ccd_window_factor": 4, "ccd_wide_queries": 8}
        mask = 0 if knobs is None else sum(bits[k] for k in set(knobs))
        p0 = np.ascontiguousarray(np.stack([np.asarray(a, np.float64).reshape(-1, 6) for a, _ in samples]))
        p1 = np.ascontiguousarray(np.stack([np.asarray(b, np.float64).reshape(-1, 6) for _, b in samples]))
        t = Tuning()
        ms = ctypes.c_double()
        self._check(self._L.rigc_autotune_earliest_toi(self._ctx, len(samples), _p(p0, _c_double_p),
                                                       _p(p1, _c_double_p), float(minimum_separation_distance),
                                                       mask, ctypes.byref(t), ctypes.byref(ms)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}, ms.value

    def compute_earliest_toi(self, poses_t0, poses_t1, minimum_separation_distance=0.0):
        """Broad + narrow phase on the device; +inf when nothing collides."""
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        p1 = np.ascontiguousarray(poses_t1, np.float64).reshape(-1, 6)
        toi = ctypes.c_double()
        self._check(self._L.rigc_compute_earliest_toi(self._ctx, _p(p0, _c_double_p), _p(p1, _c_double_p),
                                                      float(minimum_separation_distance), ctypes.byref(toi)))
        return toi.value

    def narrow_phase_piecewise_linear(self, poses_t0, poses_t1, ee, fv, earliest_toi=1.0,
                                      minimum_separation_distance=0.0):
        """Piecewise-linear CCD over candidates (TrajectoryType::PIECEWISE_LINEAR);
        returns (earliest_toi, toi[n], hit[n]), EE queries first."""
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        p1 = np.ascontiguousarray(poses_t1, np.float64).reshape(-1, 6)
        ee = np.ascontiguousarray(np.asarray(ee, np.int32).reshape(-1, 2))
        fv = np.ascontiguousarray(np.asarray(fv, np.int32).reshape(-1, 2))
        n = len(ee) + len(fv)
        toi = np.empty(max(n, 1))
        hit = np.empty(max(n, 1), np.int8)
        best = ctypes.c_double()
        self._check(self._L.rigc_narrow_phase_piecewise_linear(
            self._ctx, _p(p0, _c_double_p), _p(p1, _c_double_p), _p(ee, _c_int32_p), len(ee), _p(fv, _c_int32_p),
            len(fv), float(earliest_toi), float(minimum_separation_distance), ctypes.byref(best),
            _p(toi, _c_double_p), _p(hit, _c_int8_p)))
        return best.value, toi[:n], hit[:n]

    def compute_earliest_toi_piecewise_linear(self, poses_t0, poses_t1, minimum_separation_distance=0.0):
        """Broad phase + piecewise-linear narrow phase; +inf when nothing collides."""
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        p1 = np.ascontiguousarray(poses_t1, np.float64).reshape(-1, 6)
        toi = ctypes.c_double()
        self._check(self._L.rigc_compute_earliest_toi_piecewise_linear(
            self._ctx, _p(p0, _c_double_p), _p(p1, _c_double_p), float(minimum_separation_distance),
            ctypes.byref(toi)))
        return toi.value

    # ------------------------------------------------------------ assembly
    def construct_constraint_set(self, poses, dhat, dmin=0.0):
        """Build the constraint set at poses on the device; returns the
        (VV, EV, EE, FV) counts."""
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        counts = np.zeros(4, np.int64)
        self._check(self._L.rigc_construct_constraint_set(self._ctx, _p(p, _c_double_p), float(dhat), float(dmin),
                                                          _p(counts, _c_int64_p)))
        return counts

    def constraint_set(self):
        """Current constraint set: (records (n,6) int32: type, a, b,
        multiplicity, body0, body1; eps_x (n,))."""
        n = ctypes.c_int64()
        st = self._L.rigc_get_constraint_set(self._ctx, 0, None, None, ctypes.byref(n))
        if st not in (RIGC_OK, RIGC_ERR_CAPACITY):
            self._check(st)
        rec = np.zeros((max(n.value, 1), 6), np.int32)
        eps = np.zeros(max(n.value, 1))
        self._check(self._L.rigc_get_constraint_set(self._ctx, n.value, _p(rec, _c_int32_p), _p(eps, _c_double_p),
                                                    ctypes.byref(n)))
        return rec[:n.value], eps[:n.value]

    def _potential(self, fn, args, flags):
        nb = self.num_bodies
        energy = ctypes.c_double()
        grad = np.zeros(6 * nb) if flags & (RIGC_GRADIENT | RIGC_HESSIAN) else None
        hess = None
        if flags & RIGC_HESSIAN:
            cap = getattr(self, "_cap_off", 1024)
            for _ in range(2):
                diag = np.zeros((nb, 36))
                pairs = np.zeros((max(cap, 1), 2), np.int32)
                off = np.zeros((max(cap, 1), 36))
                h = BlockHessian(_p(diag, _c_double_p), _p(pairs, _c_int32_p), _p(off, _c_double_p), cap, 0)
                st = fn(self._ctx, *args, int(flags), ctypes.byref(energy), _p(grad, _c_double_p), ctypes.byref(h))
                if st == RIGC_ERR_CAPACITY:
                    cap = self._cap_off = int(h.count)
                    continue
                self._check(st)
                hess = (diag.reshape(nb, 6, 6), pairs[:h.count].copy(), off[:h.count].reshape(-1, 6, 6))
                break
            else:
                raise RigcError(RIGC_ERR_CAPACITY, "Hessian capacity retry failed")
        else:
            self._check(fn(self._ctx, *args, int(flags), ctypes.byref(energy), _p(grad, _c_double_p), None))
        return energy.value, grad, hess

    def barrier_term(self, poses, dhat, dmin=0.0, grad=False, hess=False, rebuild=False, unprojected=False):
        """Barrier energy over the current constraint set (rebuilt at poses
        if rebuild); grad (6 nb) and hess (diag (nb,6,6), pairs (m,2),
        off (m,6,6)) when asked, without project_to_psd when unprojected."""
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        flags = (RIGC_GRADIENT if grad else 0) | (RIGC_HESSIAN if hess else 0) | (RIGC_BUILD_CONSTRAINTS if rebuild else 0)
        flags |= RIGC_UNPROJECTED if unprojected else 0
        self._keep = [p]
        return self._potential(self._L.rigc_compute_barrier_term, (_p(p, _c_double_p), float(dhat), float(dmin)), flags)

    def minimum_distance(self, poses, dhat, dmin=0.0, rebuild=True):
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        out = ctypes.c_double()
        flags = RIGC_BUILD_CONSTRAINTS if rebuild else 0
        self._check(self._L.rigc_compute_minimum_distance(self._ctx, _p(p, _c_double_p), float(dhat), float(dmin),
                                                          flags, ctypes.byref(out)))
        return out.value

    def update_friction_constraints(self, poses, dhat, dmin, barrier_stiffness, mu, rebuild=True):
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        n = ctypes.c_int64()
        flags = RIGC_BUILD_CONSTRAINTS if rebuild else 0
        self._check(self._L.rigc_update_friction_constraints(self._ctx, _p(p, _c_double_p), float(dhat), float(dmin),
                                                             float(barrier_stiffness), float(mu), flags,
                                                             ctypes.byref(n)))
        return n.value

    def friction_constraints(self):
        """(n, 9): normal force, coords (2), basis (3x2 row-major)."""
        n = ctypes.c_int64()
        st = self._L.rigc_get_friction_constraints(self._ctx, 0, None, ctypes.byref(n))
        if st not in (RIGC_OK, RIGC_ERR_CAPACITY):
            self._check(st)
        rec = np.zeros((max(n.value, 1), 9))
        self._check(self._L.rigc_get_friction_constraints(self._ctx, n.value, _p(rec, _c_double_p), ctypes.byref(n)))
        return rec[:n.value]

    def friction_term(self, poses, poses_t0, epsv_times_h, grad=False, hess=False, unprojected=False):
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        flags = (RIGC_GRADIENT if grad else 0) | (RIGC_HESSIAN if hess else 0) | (RIGC_UNPROJECTED if unprojected else 0)
        return self._potential(self._L.rigc_compute_friction_term,
                               (_p(p, _c_double_p), _p(p0, _c_double_p), float(epsv_times_h)), flags)

    def friction_adjoint(self, poses, poses_t0, epsv_times_h, dhat, dmin, barrier_stiffness, lam):
        """(d grad_x D(poses; poses_t0) / d poses_t0)^T lam over the current
        lagged friction set (rigc_friction_adjoint)."""
        p = np.ascontiguousarray(poses, np.float64).reshape(-1, 6)
        p0 = np.ascontiguousarray(poses_t0, np.float64).reshape(-1, 6)
        lm = np.ascontiguousarray(lam, np.float64).reshape(-1, 6)
        out = np.zeros_like(p)
        self._check(self._L.rigc_friction_adjoint(self._ctx, _p(p, _c_double_p), _p(p0, _c_double_p),
                                                  float(epsv_times_h), float(dhat), float(dmin),
                                                  float(barrier_stiffness), _p(lm, _c_double_p), _p(out, _c_double_p)))
        return out.reshape(-1)


def dense_hessian(hess, nb):
    """Dense (6 nb x 6 nb) matrix from the block form (test helper)."""
    diag, pairs, off = hess
    H = np.zeros((6 * nb, 6 * nb))
    for b in range(nb):
        H[6 * b:6 * b + 6, 6 * b:6 * b + 6] = diag[b]
    for (i, j), B in zip(pairs, off):
        H[6 * i:6 * i + 6, 6 * j:6 * j + 6] += B
        H[6 * j:6 * j + 6, 6 * i:6 * i + 6] += B.T
    return H
