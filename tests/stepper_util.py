"""Shared helpers for the host-stepper tests: the oracle as a contact
backend, and small scenes with known motion."""
import ctypes

import numpy as np

import oracle_lib as O
from gpu_util import PKG  # noqa: F401  (puts the package on sys.path)
from rigidipc_amd import scenes as S
from rigidipc_amd import sim as M


def oracle_backend(L=None):
    """The oracle's contact-backend table (of the library L, by default the
    oracle's parity build)."""
    L = L or O.lib()
    L.oracle_backend.argtypes = [ctypes.POINTER(M.Backend)]
    L.oracle_backend.restype = None
    be = M.Backend()
    L.oracle_backend(ctypes.byref(be))
    return ctypes.pointer(be)


def free_body_scene(w=(0.0, 0.0, 0.0), v=(0.3, 2.0, -0.1)):
    """One cube and a far-away static plane (no contact within the run)."""
    Vc, Fc = S.cube(1.0)
    body = S.Body(Vc, Fc, np.eye(3), np.array([0.0, 50.0, 0.0]), np.array(v), np.array(w))
    sc = S.assemble([body, S._ground()])
    sc.dt, sc.dhat, sc.mu, sc.gravity = 0.01, 1e-3, 0.0, np.array([0, -9.81, 0])
    return sc


def resting_scene(mu=0.0, gap=4e-4, v=(0.0, 0.0, 0.0)):
    """A unit cube with its bottom face `gap` above the plane (inside the
    barrier band dhat = 1e-3), away from the plane's diagonal edge so that
    only face-vertex contacts are active."""
    Vc, Fc = S.cube(1.0)
    body = S.Body(Vc, Fc, np.eye(3), np.array([4.0, 0.5 + gap, 1.0]), np.array(v), np.zeros(3))
    sc = S.assemble([body, S._ground()])
    sc.dt, sc.dhat, sc.mu, sc.gravity = 0.01, 1e-3, mu, np.array([0, -9.81, 0])
    return sc


def min_vertex_height(sc, poses):
    O.set_scene(sc)
    W = O.world_vertices(poses, len(sc.vertices))
    dyn = np.repeat(~sc.is_static.astype(bool), np.diff(sc.body_vertex_start))
    return W[dyn, 1].min()
