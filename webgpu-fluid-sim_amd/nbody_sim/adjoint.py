"""Differentiable trajectories: the hand-written adjoint as a torch.autograd.Function.

``trajectory(pos0, vel0, mass, sim, nsteps, every)`` returns the state after
``nsteps`` Barnes-Hut steps of ``sim`` (a Simulation with constant
softening); its backward runs nbody_adjoint_backward, which differentiates
every step with the tree topology and opening decisions frozen at their
forward values.  Tensors live on the host (the C ABI takes host arrays); the
steps and the adjoint run on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch


class BarnesHutTrajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos0, vel0, mass, sim, nsteps, every):
        sim.setBodies(mass.detach().cpu().numpy(), pos0.detach().cpu().numpy(),
                      vel0.detach().cpu().numpy())
        sim.adjointForward(int(nsteps), int(every))
        _, pos, vel = sim.getBodies()
        ctx.sim = sim
        ctx.dtype = pos0.dtype
        return torch.from_numpy(pos).to(pos0.dtype), torch.from_numpy(vel).to(pos0.dtype)

    @staticmethod
    def backward(ctx, gpos, gvel):
        gp = None if gpos is None else gpos.detach().cpu().float().numpy()
        gv = None if gvel is None else gvel.detach().cpu().float().numpy()
        dx, dv, dm = ctx.sim.adjointBackward(gp, gv)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(ctx.dtype)  # noqa: E731
        return t(dx), t(dv), t(dm), None, None, None


def trajectory(pos0, vel0, mass, sim, nsteps: int, every: int = 20):
    return BarnesHutTrajectory.apply(pos0, vel0, mass, sim, nsteps, every)
