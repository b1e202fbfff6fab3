"""Closed-form contact derivatives of the device code, built for the host
(tests/native/closed_forms.cpp over rigid-ipc_amd/csrc/rigc_barrier.hpp),
against the oracle's forward-mode dual-number derivatives of the same
formulas (oracle_barrier.cpp: constraint_local, point_derivatives).

Two independent differentiations of one function agree to rounding: the
tolerances below (1e-12 relative to the largest entry) are four orders
tighter than the device parity bar (1e-10 energy / gradient, 1e-9
Hessian), and a wrong sign, index or dropped term in either shows up at
O(1)."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tests", "native", "_build", "libclosed_forms.so")
dp = ctypes.POINTER(ctypes.c_double)


def ptr(a):
    return a.ctypes.data_as(dp)


@pytest.fixture(scope="module")
def libs():
    if not os.path.exists(LIB):
        pytest.skip("closed-form host library not built (run __graft_entry__.build())")
    L = ctypes.CDLL(LIB)
    L.cf_barrier_local.restype = ctypes.c_double
    L.cf_barrier_local.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_longlong, dp, dp]
    L.cf_body_frame.argtypes = [dp, dp, dp]
    L.cf_distance.restype = ctypes.c_double
    L.cf_distance.argtypes = [ctypes.c_int, dp, ctypes.POINTER(ctypes.c_int)]
    L.cf_friction_frame.argtypes = [ctypes.c_int, dp, dp, dp]
    OL = O._declare_barrier()
    OL.oracle_point_derivatives.argtypes = [dp, dp, dp, dp, dp]
    OL.oracle_constraint_local.restype = ctypes.c_double
    OL.oracle_constraint_local.argtypes = [ctypes.c_int, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_long, dp, dp]
    return L, OL


SYM = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (1, 1): 3, (1, 2): 4, (2, 2): 5}


@pytest.mark.parametrize("mag", [0.0, 1e-9, 1e-5, 1e-3, 0.05, 0.5, 2.0, 3.1, 6.0, 20.0])
def test_body_frame_matches_dual_derivatives(libs, mag):
    """dR/dr_c x and d2R/dr_c dr_d x against the oracle's second-order duals
    over Rodrigues' formula (rigid_body.tpp:26-78), on both sides of the
    sinc Taylor switches (1e-4, 0.1) and past pi."""
    L, OL = libs
    rng = np.random.default_rng(int(mag * 1000) + 7)
    for _ in range(20):
        r = rng.normal(size=3)
        r = r / np.linalg.norm(r) * mag if mag > 0 else np.zeros(3)
        xb = rng.normal(size=3)
        pose = np.concatenate([rng.normal(size=3), r])
        w, J, H = np.zeros(3), np.zeros(18), np.zeros(27)
        OL.oracle_point_derivatives(ptr(pose), ptr(xb), ptr(w), ptr(J), ptr(H))
        J, H = J.reshape(3, 6), H.reshape(3, 3, 3)
        dR, d2R = np.zeros(27), np.zeros(54)
        L.cf_body_frame(ptr(r), ptr(dR), ptr(d2R))
        dR, d2R = dR.reshape(3, 3, 3), d2R.reshape(6, 3, 3)
        Jc = np.stack([dR[c] @ xb for c in range(3)], axis=1)
        Hc = np.zeros((3, 3, 3))
        for c in range(3):
            for d in range(3):
                Hc[:, c, d] = d2R[SYM[(min(c, d), max(c, d))]] @ xb
        assert np.abs(Jc - J[:, 3:]).max() <= 1e-12 * max(np.abs(J[:, 3:]).max(), 1e-300)
        assert np.abs(Hc - H).max() <= 1e-12 * max(np.abs(H).max(), 1e-300)


def contact_config(rng, ctype, parallel=False):
    X = np.zeros((4, 3))
    if ctype == 0:
        X[0] = rng.normal(size=3)
        X[1] = X[0] + rng.normal(size=3) * 0.02
    elif ctype == 1:
        X[1] = rng.normal(size=3)
        X[2] = X[1] + rng.normal(size=3)
        X[0] = X[1] + rng.uniform(-0.2, 1.2) * (X[2] - X[1]) + rng.normal(size=3) * 0.02
    elif ctype == 2:
        X[0] = rng.normal(size=3)
        X[1] = X[0] + rng.normal(size=3)
        c = X[0] + rng.uniform(-0.2, 1.2) * (X[1] - X[0]) + rng.normal(size=3) * 0.02
        dv = (X[1] - X[0]) * rng.uniform(0.3, 1.5) + rng.normal(size=3) * 1e-3 if parallel else rng.normal(size=3)
        X[2] = c - rng.uniform(0, 1) * dv
        X[3] = X[2] + dv
    else:
        X[1] = rng.normal(size=3)
        X[2] = X[1] + rng.normal(size=3)
        X[3] = X[1] + rng.normal(size=3)
        a, b = rng.uniform(0, 1, 2)
        if a + b > 1:
            a, b = 1 - a, 1 - b
        X[0] = X[1] + a * (X[2] - X[1]) + b * (X[3] - X[1]) + rng.normal(size=3) * 0.02
    return X.reshape(-1).copy()


@pytest.mark.parametrize("ctype,parallel", [(0, False), (1, False), (2, False), (2, True), (3, False)])
def test_barrier_local_matches_dual_derivatives(libs, ctype, parallel):
    """Energy, vertex gradient and unprojected vertex Hessian of
    m * [mollifier] * b(d) for every distance type the classifiers reach
    (point-point, point-line, line-line, point-plane), with the edge-edge
    mollifier active on the nearly parallel cases."""
    L, OL = libs
    rng = np.random.default_rng(100 + 10 * ctype + parallel)
    dhat, dmin = 0.03, 0.0
    kinds = set()
    mollified = 0
    for _ in range(400):
        X = contact_config(rng, ctype, parallel)
        eps = -1.0
        if ctype == 2:
            ea, eb = X[3:6] - X[0:3], X[9:12] - X[6:9]
            eps = 1e-3 * ea.dot(ea) * eb.dot(eb)
            mollified += int(np.sum(np.cross(ea, eb) ** 2) < eps)
        kind = ctypes.c_int(0)
        L.cf_distance(ctype, ptr(X), ctypes.byref(kind))
        kinds.add(kind.value)
        mult = int(rng.integers(1, 4))
        go, Ho, gc, Hc = np.zeros(12), np.zeros(144), np.zeros(12), np.zeros(144)
        eo = OL.oracle_constraint_local(ctype, ptr(X), eps, dhat, dmin, mult, ptr(go), ptr(Ho))
        ec = L.cf_barrier_local(ctype, ptr(X), eps, dhat, dmin, mult, ptr(gc), ptr(Hc))
        assert ec == pytest.approx(eo, rel=1e-14, abs=0)
        if eo == 0:
            continue
        assert np.abs(gc - go).max() <= 1e-12 * np.abs(go).max()
        assert np.abs(Hc - Ho).max() <= 1e-12 * np.abs(Ho).max()
    if ctype in (1, 2, 3):
        assert len(kinds) >= 2, kinds
    if parallel:
        assert mollified > 20


def test_friction_frame_matches_oracle(libs):
    """Tangent bases and closest-point coordinates of the device friction
    frame against what the geometry fixes (the device friction records are
    also compared with the oracle's in the GPU tests): the basis is
    orthonormal and spans the plane normal to the contact direction (VV: the
    vertex difference; EV: the vertex minus its projection on the edge's
    line; EE: ea x eb; FV: the face normal), and the coordinates are the
    closest-point parameters on the lines / plane (EV: the edge parameter;
    EE: both line parameters; FV: the barycentric pair of the projection),
    solved here by numpy."""
    L, OL = libs
    rng = np.random.default_rng(5)
    for ctype in range(4):
        for _ in range(50):
            X = contact_config(rng, ctype)
            coords, basis = np.zeros(2), np.zeros(6)
            L.cf_friction_frame(ctype, ptr(X), ptr(coords), ptr(basis))
            B = basis.reshape(3, 2)
            assert np.allclose(B.T @ B, np.eye(2), atol=1e-12)
            X = X.reshape(4, 3)
            if ctype == 0:
                n = X[1] - X[0]
            elif ctype == 1:
                e = X[2] - X[1]
                a = (X[0] - X[1]) @ e / (e @ e)
                assert coords[0] == pytest.approx(a, rel=1e-10, abs=1e-12)
                n = X[0] - (X[1] + a * e)
            elif ctype == 2:
                ea, eb = X[1] - X[0], X[3] - X[2]
                M = np.array([[ea @ ea, -(ea @ eb)], [-(ea @ eb), eb @ eb]])
                st = np.linalg.solve(M, [(X[2] - X[0]) @ ea, (X[0] - X[2]) @ eb])
                np.testing.assert_allclose(coords, st, rtol=1e-8, atol=1e-10)
                n = np.cross(ea, eb)
            else:
                e0, e1 = X[2] - X[1], X[3] - X[1]
                M = np.array([[e0 @ e0, e0 @ e1], [e0 @ e1, e1 @ e1]])
                uv = np.linalg.solve(M, [(X[0] - X[1]) @ e0, (X[0] - X[1]) @ e1])
                np.testing.assert_allclose(coords, uv, rtol=1e-10, atol=1e-12)
                n = np.cross(e0, e1)
            n = n / np.linalg.norm(n)
            assert np.abs(B.T @ n).max() <= 1e-10
