"""Differentiable rollouts over the CPU oracle (CPU only).

* The oracle's friction adjoint (dual numbers over the record's vertices at
  x and x0) against central differences of the friction gradient with the
  lagged friction set rebuilt at each perturbed start pose, on a cube
  resting on the static plane (face-vertex contacts with the plane, whose
  tangent basis is the plane's, so the adjoint's only frozen quantities do
  not move).
* The stepper's implicit-function adjoint (rigs_rollout_backward) over the
  oracle backend against central differences (h = 1e-6) of whole rollouts,
  with and without friction, sliding and spinning, for a loss of the end
  poses and end velocities.
* The torch.autograd.Function gives the same gradient, and the backward
  pass refuses a rollout whose steps did not converge.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from stepper_util import oracle_backend, resting_scene
from rigidipc_amd import RigcError
from rigidipc_amd import diff as Dm
from rigidipc_amd import sim as M


def friction_grad(sc, x, x0, dhat, kappa, mu, eh):
    O.constraint_set(x0, dhat)
    O.friction_update(x0, dhat, 0.0, kappa, mu)
    return O.friction(x, x0, eh)[1]


@pytest.mark.parametrize("slide", [0.0, 2e-4, 3e-3])
def test_friction_adjoint_matches_finite_differences(slide):
    """slide below and above eps_v h = 1e-3: both branches of the smoothed
    friction; slide 0 (a body that has not moved since the set was built):
    the friction gradient vanishes there but its derivative does not."""
    sc = resting_scene(mu=0.5, gap=3e-4)
    O.set_scene(sc)
    dhat, kappa, mu, eh = 1e-3, 2e3, 0.5, 1e-3
    x0 = sc.poses.copy()
    x0[0, 3:] = [0.002, -0.001, 0.0015]
    x = x0.copy()
    if slide > 0:
        x[0, :3] += slide * np.array([0.8, -0.05, 0.6])
        x[0, 3:] += [0.001, 0.003, -0.002]
    rng = np.random.default_rng(3)
    lam = rng.normal(size=x.size)
    lam[6:] = 0.0  # the plane's dofs are fixed
    friction_grad(sc, x, x0, dhat, kappa, mu, eh)
    adj = O.friction_adjoint(x, x0, eh, dhat, 0.0, kappa, lam)
    fd = np.zeros(6)
    h = 1e-7
    for j in range(6):
        xp, xm = x0.copy(), x0.copy()
        xp[0, j] += h
        xm[0, j] -= h
        gp = friction_grad(sc, x, xp, dhat, kappa, mu, eh)
        gm = friction_grad(sc, x, xm, dhat, kappa, mu, eh)
        fd[j] = (lam @ gp - lam @ gm) / (2 * h)
    # at slide 0 the cubic term of f0 makes the gradient go like |u| u, so
    # the central difference itself errs by about h / (eps_v h) = 1e-4
    tol = 2e-4 if slide == 0 else 1e-5
    assert np.abs(adj[:6] - fd).max() <= tol * np.abs(fd).max(), (adj[:6], fd)
    assert np.abs(fd).max() > 0


def cube_rollout(mu, n_steps=8):
    sc = resting_scene(mu=mu, gap=4e-4, v=(0.6, 0.0, 0.2))
    sc.velocities[0, 3:] = [0.3, 1.0, -0.2]
    sim = M.Simulation(sc, params=Dm.params_for_gradients(sc), backend=oracle_backend())
    return sc, sim


def cube_loss(x, v):
    target = np.array([4.3, 0.6, 1.1])
    return float(((x[0, :3] - target) ** 2).sum() + 0.1 * (x[0, 3:] ** 2).sum() + 0.05 * (v[0] ** 2).sum())


def cube_loss_grads(x, v):
    target = np.array([4.3, 0.6, 1.1])
    gx, gv = np.zeros_like(x), np.zeros_like(v)
    gx[0, :3] = 2 * (x[0, :3] - target)
    gx[0, 3:] = 0.2 * x[0, 3:]
    gv[0] = 0.1 * v[0]
    return gx, gv


@pytest.mark.parametrize("mu", [0.0, 0.5])
def test_rollout_gradient_matches_central_differences(mu):
    sc, sim = cube_rollout(mu)
    P0, V0 = sc.poses.copy(), sc.velocities.copy()

    def run(v):
        sim.set_state(P0, v)
        f0 = sim.stats()["failed_solves"]
        sim.rollout(8)
        assert sim.stats()["failed_solves"] == f0
        return sim.state()

    x, v = run(V0)
    assert sim.stats()["num_constraints"] > 0  # the cube touches the plane
    gx, gv = cube_loss_grads(x, v)
    g = sim.rollout_backward(gx, gv)
    fd = np.zeros(6)
    for i in range(6):
        vp, vm = V0.copy(), V0.copy()
        vp[0, i] += 1e-6
        vm[0, i] -= 1e-6
        fd[i] = (cube_loss(*run(vp)) - cube_loss(*run(vm))) / 2e-6
    assert np.abs(g[0] - fd).max() <= 1e-4 * np.abs(fd).max(), (g[0], fd)
    assert np.all(g[1] == 0)  # the static plane


def test_torch_function_gives_the_adjoint_gradient():
    sc, sim = cube_rollout(0.5)
    P0 = sc.poses.copy()
    v0 = torch.tensor(sc.velocities, dtype=torch.float64, requires_grad=True)
    x, v = Dm.rollout(sim, P0, v0, 8)
    target = torch.tensor([4.3, 0.6, 1.1], dtype=torch.float64)
    loss = ((x[0, :3] - target) ** 2).sum() + 0.1 * (x[0, 3:] ** 2).sum() + 0.05 * (v[0] ** 2).sum()
    loss.backward()
    assert loss.item() == pytest.approx(cube_loss(x.detach().numpy(), v.detach().numpy()), rel=1e-15)
    gx, gv = cube_loss_grads(x.detach().numpy(), v.detach().numpy())
    sim.set_state(P0, sc.velocities)
    sim.rollout(8)
    want = sim.rollout_backward(gx, gv)
    np.testing.assert_allclose(v0.grad.numpy(), want, rtol=1e-12, atol=1e-15)


def test_backward_refuses_unconverged_steps():
    sc = resting_scene(mu=0.0, gap=4e-4, v=(0.6, 0.0, 0.2))
    p = Dm.params_for_gradients(sc)
    p.max_iterations = 1  # no solve can converge in one iteration
    sim = M.Simulation(sc, params=p, backend=oracle_backend())
    sim.rollout(2)
    assert sim.stats()["failed_solves"] > 0
    with pytest.raises(RigcError, match="did not converge"):
        sim.rollout_backward(np.ones((2, 6)))
