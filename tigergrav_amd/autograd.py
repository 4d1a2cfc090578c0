"""torch.autograd wrapper of the batched force solve and its adjoint.

    acc = tree_accelerations(pos, mass, ewald=True, soft_len=h, theta=0.7, leaf=64)
    loss = (w * acc).sum(); loss.backward()      # pos.grad, mass.grad from the hand-written adjoint

`pos` is an (N, 3) float64 tensor of positions in the unit box, `mass` an (N,) float64 tensor.
The forward builds the tree, interaction lists and expansion centres from `pos` and solves on the
GPU; the backward differentiates that solve with its tree held fixed (tg_solver_adjoint), so the
gradient is exact for the function the forward evaluated.  Double backward is not supported.
"""
import numpy as np
import torch

from .ics import double_to_pos
from .solver import Solver


class TreeAccelerations(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, mass, ewald, soft_len, theta, leaf, device):
        x = double_to_pos(pos.detach().cpu().numpy())
        m = mass.detach().cpu().numpy().astype(np.float32)
        s = Solver(ewald=ewald, soft_len=soft_len, m=float(np.mean(m)), theta=theta, parts_per_node=leaf, device=device)
        s.set_particles(x, m)
        s.solve()
        s.synchronize()
        ctx.solver = s
        return torch.from_numpy(s.forces()[:, 1:].copy()).to(dtype=pos.dtype, device=pos.device)

    @staticmethod
    def backward(ctx, grad_acc):
        g = ctx.solver.adjoint(grad_acc.detach().cpu().numpy())
        gp = torch.from_numpy(g[:, :3].copy()).to(dtype=grad_acc.dtype, device=grad_acc.device)
        gm = torch.from_numpy(g[:, 3].copy()).to(dtype=grad_acc.dtype, device=grad_acc.device)
        return gp, gm, None, None, None, None, None


def tree_accelerations(pos, mass, ewald=True, soft_len=1e-3, theta=0.7, leaf=64, device=0):
    """Accelerations (N, 3) of the batched tree solve, differentiable in pos and mass."""
    return TreeAccelerations.apply(pos, mass, bool(ewald), float(soft_len), float(theta), int(leaf), int(device))
