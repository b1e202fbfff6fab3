"""Differentiable rollouts as a torch.autograd.Function.

``rollout(sim, poses0, velocities0, n_steps)`` steps ``sim`` (a
:class:`rigidipc_amd.sim.Simulation`) n_steps from (poses0, velocities0) and
returns the end poses and velocities as float64 tensors. Its backward is the
stepper's implicit-function adjoint (rigs_rollout_backward): every step is
rebuilt from the checkpoint the forward pass kept (its start state and the
barrier stiffnesses its solve used; the constraint and lagged friction sets
are recomputed, not stored) and differentiated through its converged Newton
solve with the unprojected Hessian. Gradients flow to the initial
velocities (linear in world coordinates, angular in the body frame, the
layout rigs_bodies takes); the initial poses are a constant of the rollout.

PyTorch is the interface only: the forward and backward passes run in the
stepper over its contact backend (the device library by default). Tensors
may live on the CPU or on a HIP device; results come back on the device of
velocities0.

The rollout is differentiated as a function of its inputs when every step's
solve ends at the minimiser: use :func:`params_for_gradients` (a tight
Newton tolerance and a few polishing Newton steps) for the simulation's
parameters.
"""
import numpy as np
import torch

from . import sim as M


def params_for_gradients(scene, velocity_conv_tol=1e-7, polish=2):
    """Simulation parameters of `scene` whose solves end at the minimiser to
    rounding, so that a rollout is a smooth function of its inputs."""
    p = M.default_params(scene)
    p.velocity_conv_tol = velocity_conv_tol
    p.newton_polish = polish
    return p


class Rollout(torch.autograd.Function):
    """(velocities0 [nb, 6]) -> (poses [nb, 6], velocities [nb, 6]) after
    n_steps steps of `sim` from poses0."""

    @staticmethod
    def forward(ctx, velocities0, sim, poses0, n_steps):
        v0 = velocities0.detach().to("cpu", torch.float64).numpy()
        sim.set_state(np.asarray(poses0, np.float64), v0)
        sim.rollout(int(n_steps))
        x, v = sim.state()
        ctx.sim = sim
        dev = velocities0.device
        return torch.from_numpy(x).to(dev), torch.from_numpy(v).to(dev)

    @staticmethod
    def backward(ctx, grad_poses, grad_velocities):
        nb = ctx.sim.num_bodies
        gx = np.zeros((nb, 6)) if grad_poses is None else grad_poses.detach().to("cpu", torch.float64).numpy()
        gv = None if grad_velocities is None else grad_velocities.detach().to("cpu", torch.float64).numpy()
        g = ctx.sim.rollout_backward(gx, gv)
        dev = grad_poses.device if grad_poses is not None else grad_velocities.device
        return torch.from_numpy(g).to(dev), None, None, None


def rollout(sim, poses0, velocities0, n_steps):
    """End poses and velocities of an n_steps rollout from (poses0,
    velocities0); differentiable with respect to velocities0."""
    return Rollout.apply(velocities0, sim, poses0, n_steps)
