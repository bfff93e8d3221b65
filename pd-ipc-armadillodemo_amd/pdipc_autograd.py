"""torch.autograd.Function over the C ABI: a differentiable PD-IPC rollout.

    x_T = PdIpcRollout.apply(k_tri, v0, sim, x0, nsteps, checkpoint_every)

runs `nsteps` FP64 steps of `sim` (a pdipc.Simulator holding one cloth as
simulated model 0) from positions x0 and velocity v0 with per-triangle strain
stiffness k_tri, and returns the final simulated positions.  The backward pass
is the library's hand-written adjoint (pdipc_backward), with checkpoints every
`checkpoint_every` steps.  Tensors are float64 on the CPU: the simulation state
lives in the library's device memory, so only seeds and gradients cross the
boundary.
"""

import numpy as np
import torch


class PdIpcRollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, k_tri, v0, sim, x0, nsteps, checkpoint_every):
        sim.set_cloth_stiffness(0, k_tri.detach().cpu().numpy().astype(np.float64))
        sim.set_x(np.asarray(x0, np.float64))
        sim.set_v(v0.detach().cpu().numpy().astype(np.float64))
        sim.rollout(int(nsteps), int(checkpoint_every))
        ctx.sim = sim
        ctx.n_tris = int(k_tri.numel())
        return torch.from_numpy(sim.x()[: sim.sim].copy())

    @staticmethod
    def backward(ctx, grad_x):
        _, dv0, dk = ctx.sim.backward(dx=grad_x.detach().cpu().numpy(), n_tris=ctx.n_tris)
        return torch.from_numpy(dk.copy()), torch.from_numpy(dv0.copy()), None, None, None, None
