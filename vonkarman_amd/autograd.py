"""torch.autograd wrappers of the hand-written discrete adjoint.

The forward runs the ETD-RK4 steps on the gfx950 kernels; the backward runs
the adjoint sweep of the same C ABI (vk_solver_adjoint_device: step VJPs with
binomial checkpointing and in-step stage recomputation).  Tensors are
complex128 CUDA tensors shaped like u_hat, [3, n, n, n//2+1], passed to the
library by device pointer; torch's stream is synchronised before each call
and every library call is complete when it returns.

Device pointers pass between torch and the library only if both use one HIP
runtime: import torch (or this module) before anything loads the library.
"""
from __future__ import annotations

import torch

from ._lib import VK_OBJ_DISSIPATION, VK_OBJ_FINAL_ENSTROPHY, VK_OBJ_NONE, hip_runtimes
from .periodic import Periodic3D


def _check(t: torch.Tensor, solver: Periodic3D) -> torch.Tensor:
    rt = hip_runtimes()
    if len(rt) > 1:
        raise RuntimeError("two HIP runtimes are loaded (%s): the vonkarman_amd library was loaded before torch; "
                           "import torch first" % ", ".join(sorted(rt)))
    n = solver.n
    if t.dtype != torch.complex128 or not t.is_cuda or tuple(t.shape) != (3, n, n, n // 2 + 1):
        raise ValueError("expected a complex128 CUDA tensor of shape (3, n, n, n//2+1)")
    return t.contiguous()


class Trajectory(torch.autograd.Function):
    """u_T = step^nsteps(u_0) on `solver` (fixed dt); backward gives
    (d u_T / d u_0)^T grad_u_T."""

    @staticmethod
    def forward(ctx, u0: torch.Tensor, solver: Periodic3D, nsteps: int, max_snapshots: int = 40):
        u0 = _check(u0, solver)
        torch.cuda.synchronize()
        solver.set_u_hat_device(u0.data_ptr())
        solver.steps(nsteps)
        uT = torch.empty_like(u0)
        solver.u_hat_device(uT.data_ptr())
        ctx.save_for_backward(u0)
        ctx.solver, ctx.nsteps, ctx.max_snapshots = solver, nsteps, max_snapshots
        return uT

    @staticmethod
    def backward(ctx, grad_uT: torch.Tensor):
        (u0,) = ctx.saved_tensors
        solver = ctx.solver
        g_in = _check(grad_uT, solver)
        torch.cuda.synchronize()
        solver.set_u_hat_device(u0.data_ptr())
        grad = torch.empty_like(u0)
        solver.adjoint_device(ctx.nsteps, VK_OBJ_NONE, g_in.data_ptr(), ctx.max_snapshots, grad.data_ptr())
        return grad, None, None, None


class Objective(torch.autograd.Function):
    """J(u_0) for a built-in objective over nsteps fixed-dt steps:
    "final_enstrophy" (enstrophy of u_T) or "dissipation" (sum over n=1..T of
    dt 2 nu enstrophy(u_n)).  The gradient comes from the adjoint sweep."""

    @staticmethod
    def forward(ctx, u0: torch.Tensor, solver: Periodic3D, nsteps: int, objective: str = "dissipation",
                max_snapshots: int = 40):
        u0 = _check(u0, solver)
        obj = {"final_enstrophy": VK_OBJ_FINAL_ENSTROPHY, "dissipation": VK_OBJ_DISSIPATION}[objective]
        torch.cuda.synchronize()
        solver.set_u_hat_device(u0.data_ptr())
        grad = torch.empty_like(u0)
        J, _ = solver.adjoint_device(nsteps, obj, None, max_snapshots, grad.data_ptr())
        ctx.save_for_backward(grad)
        return torch.tensor(J, dtype=torch.float64, device=u0.device)

    @staticmethod
    def backward(ctx, gJ: torch.Tensor):
        (grad,) = ctx.saved_tensors
        return gJ * grad, None, None, None, None


def trajectory(u0: torch.Tensor, solver: Periodic3D, nsteps: int, max_snapshots: int = 40) -> torch.Tensor:
    return Trajectory.apply(u0, solver, nsteps, max_snapshots)


def objective(u0: torch.Tensor, solver: Periodic3D, nsteps: int, kind: str = "dissipation",
              max_snapshots: int = 40) -> torch.Tensor:
    return Objective.apply(u0, solver, nsteps, kind, max_snapshots)
