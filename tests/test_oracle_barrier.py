"""Pin the oracle's constraint set, barrier and friction assembly (CPU).

The ipc-toolkit pieces behind these terms are not in the reference tree,
so they are pinned by what the mathematics fixes: central finite
differences of energy -> gradient -> (unprojected) Hessian at the body
dofs (the reference's own runtime check, distance_barrier_rb_problem.cpp:
1651-1689, uses the same idea), the closed-form IPC barrier on an
edge-edge pair, invariants of the constraint set (every constraint is
active, multiplicities count duplicates), PSD projection properties, and
the vertex Jacobian of the Rodrigues map. Distances, constraint-set
membership and multiplicities, the mollifier, the normal force and the
friction potential are pinned against independent geometry and closed
forms in test_oracle_ipc_pins.py.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "rigid-ipc_amd"))
import oracle_lib as O  # noqa: E402
from rigidipc_amd import scenes as S  # noqa: E402


def contact_scene(seed=0):
    """A cube resting just above the plane and a sphere and torus touching
    it, all within a large dhat so that every constraint type appears."""
    rng = np.random.default_rng(seed)
    bodies = []
    Vc, Fc = S.cube(1.0)
    bodies.append(S.Body(Vc, Fc, np.eye(3), np.array([0, 0.56, 0]), np.zeros(3), np.zeros(3)))
    Vs, Fs = S.icosphere(1, 0.4)
    bodies.append(S.Body(Vs, Fs, S.random_rotation(rng), np.array([0.2, 1.53, 0.1]), np.zeros(3), np.zeros(3)))
    Vt, Ft = S.torus(0.4, 0.1, 12, 6)
    bodies.append(S.Body(Vt, Ft, np.eye(3), np.array([1.05, 0.6, 0.2]), np.zeros(3), np.zeros(3)))
    bodies.append(S._ground())
    return S.assemble(bodies)


DHAT = 0.12


def setup(seed=0):
    sc = contact_scene(seed)
    O.set_scene(sc)
    rng = np.random.default_rng(seed + 100)
    poses = sc.poses.copy()
    poses[:-1, 3:] += rng.normal(size=(len(poses) - 1, 3)) * 0.02
    rec, eps = O.constraint_set(poses, DHAT)
    return sc, poses, rec, eps


def test_constraint_set_covers_all_types_and_is_active():
    sc, poses, rec, eps = setup(0)
    types = set(rec[:, 0].tolist())
    assert len(types) >= 3, types
    d = O.min_distance(poses)
    assert 0 < d < DHAT ** 2
    # bodies of a constraint always differ; multiplicities are positive
    assert np.all(rec[:, 4] != rec[:, 5])
    assert np.all(rec[:, 3] >= 1)


def fd_grad(f, x, h=1e-7):
    g = np.zeros_like(x)
    for i in range(x.size):
        xp, xm = x.copy(), x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        g.flat[i] = (f(xp) - f(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("seed", [0, 1])
def test_barrier_gradient_and_hessian_match_finite_differences(seed):
    sc, poses, rec, eps = setup(seed)
    nb = len(sc.group_ids)
    O.set_psd_projection(False)
    try:
        e0, g0, lh = O.barrier(poses, DHAT, want_hess=True, n_constraints=len(rec))
        H = O.assemble_hessian(rec, lh, nb)
        dyn = np.arange(6 * (nb - 1))  # the plane is static
        gfd = fd_grad(lambda x: O.barrier(x.reshape(-1, 6), DHAT)[0], poses.ravel())
        scale = np.abs(g0[dyn]).max()
        assert np.abs(gfd[dyn] - g0[dyn]).max() <= 1e-5 * scale
        Hfd = np.stack([fd_grad(lambda x, i=i: O.barrier(x.reshape(-1, 6), DHAT)[1][i], poses.ravel(), 1e-6)
                        for i in dyn])
        hs = np.abs(H[np.ix_(dyn, dyn)]).max()
        assert np.abs(Hfd[:, dyn] - H[np.ix_(dyn, dyn)]).max() <= 1e-4 * hs
    finally:
        O.set_psd_projection(True)


def test_projected_local_hessians_are_psd_and_idempotent():
    sc, poses, rec, eps = setup(0)
    _, _, lh = O.barrier(poses, DHAT, want_hess=True, n_constraints=len(rec))
    for h in lh:
        h = h.reshape(12, 12)
        assert np.allclose(h, h.T, atol=1e-9 * max(1, np.abs(h).max()))
        w = np.linalg.eigvalsh((h + h.T) / 2)
        assert w.min() >= -1e-8 * max(1.0, np.abs(w).max())


def test_edge_edge_barrier_value_closed_form():
    # two crossing unit edges (bodies 0, 1) at height z: one EE constraint
    import types
    z = 0.03
    V = np.array([[-0.5, 0, 0], [0.5, 0, 0], [0, -0.5, 0], [0, 0.5, 0]], float)
    sc = types.SimpleNamespace(
        vertices=V, edges=np.array([[0, 1], [2, 3]], np.int32), faces=np.zeros((0, 3), np.int32),
        face_edges=np.zeros((0, 3), np.int32), body_vertex_start=np.array([0, 2, 4], np.int32),
        body_edge_start=np.array([0, 1, 2], np.int32), body_face_start=np.zeros(3, np.int32),
        group_ids=np.array([0, 1], np.int32), is_static=np.zeros(2, np.uint8))
    O.set_scene(sc)
    poses = np.array([[0, 0, 0, 0, 0, 0], [0, 0, z, 0, 0, 0]], float)
    dhat = 0.1
    rec, eps = O.constraint_set(poses, dhat)
    assert len(rec) == 1 and rec[0, 0] == 2
    e, g, _ = O.barrier(poses, dhat)
    d, dh = z * z, dhat * dhat
    assert e == pytest.approx(-(d - dh) ** 2 * math.log(d / dh), rel=1e-12)
    # the force pushes the edges apart along z
    assert g[2] > 0 and g[8] < 0 and g[2] == pytest.approx(-g[8], rel=1e-12)


def test_vertex_jacobian_and_hessian_match_finite_differences():
    sc = contact_scene(1)
    O.set_scene(sc)
    rng = np.random.default_rng(3)
    for vertex in (0, 5, 30):
        b = int(np.searchsorted(sc.body_vertex_start, vertex, side="right") - 1)
        pose = sc.poses[b] + np.concatenate([np.zeros(3), rng.normal(size=3)])
        w, J, H = O.vertex_derivatives(pose, vertex)
        Jfd = np.stack([fd_grad(lambda p, j=j: O.vertex_derivatives(p, vertex)[0][j], pose) for j in range(3)])
        assert np.allclose(J, Jfd, atol=1e-6)
        for j in range(3):
            Hfd = np.stack([fd_grad(lambda p, k=k: O.vertex_derivatives(p, vertex)[1][j, 3 + k], pose, 1e-6)[3:]
                            for k in range(3)])
            assert np.allclose(H[j], Hfd, atol=1e-5)


def test_friction_gradient_and_hessian_match_finite_differences():
    sc, poses, rec, eps = setup(2)
    nb = len(sc.group_ids)
    fr = O.friction_update(poses, DHAT, 0.0, 1e3, 0.5)
    assert len(fr) == len(rec) and np.all(fr[:, 0] > 0)
    rng = np.random.default_rng(7)
    x = poses.copy()
    x[:-1] += rng.normal(size=(nb - 1, 6)) * 1e-3  # sliding a little, beyond eps_v h
    eh = 1e-5
    O.set_psd_projection(False)
    try:
        e0, g0, lh = O.friction(x, poses, eh, want_hess=True, n_constraints=len(rec))
        H = O.assemble_hessian(rec, lh, nb)
        dyn = np.arange(6 * (nb - 1))
        gfd = fd_grad(lambda y: O.friction(y.reshape(-1, 6), poses, eh)[0], x.ravel(), 1e-8)
        assert np.abs(gfd[dyn] - g0[dyn]).max() <= 1e-5 * np.abs(g0[dyn]).max()
        Hfd = np.stack([fd_grad(lambda y, i=i: O.friction(y.reshape(-1, 6), poses, eh)[1][i], x.ravel(), 1e-8)
                        for i in dyn])
        assert np.abs(Hfd[:, dyn] - H[np.ix_(dyn, dyn)]).max() <= 1e-3 * np.abs(H[np.ix_(dyn, dyn)]).max()
    finally:
        O.set_psd_projection(True)
