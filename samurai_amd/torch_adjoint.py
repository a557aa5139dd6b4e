"""torch.autograd binding of the frozen-mesh adjoint (test binding).

    u_T = advect(u0, a, problem)

u0: float64 CUDA tensor of the compact cells of all instances; a: float64
tensor (n, dim) of velocities.  Backward runs the hand-written adjoint
(recomputing each checkpoint segment) and returns dL/du0 and dL/da.  torch
only provides the device buffers: both directions run in the library on its
own stream, which is synchronized against torch's around each call.
"""
import numpy as np
import torch

from . import AdjointProblem  # noqa: F401  (re-exported for callers)


class _Advect(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, a, problem):
        if u0.dtype != torch.float64 or not u0.is_cuda or not u0.is_contiguous():
            raise TypeError("u0 must be a contiguous float64 CUDA tensor")
        n_cells = problem.nb_cells()
        if u0.numel() != n_cells:
            raise ValueError(f"u0 has {u0.numel()} values, the instances have {n_cells} cells")
        problem.set_velocity(a.detach().cpu().numpy())
        uT = torch.empty_like(u0)
        torch.cuda.synchronize(u0.device)
        problem.forward_ptr(u0.data_ptr(), uT.data_ptr())
        ctx.problem = problem
        ctx.a_shape, ctx.a_dtype, ctx.a_device = a.shape, a.dtype, a.device
        return uT

    @staticmethod
    def backward(ctx, gT):
        problem = ctx.problem
        gT = gT.contiguous()
        g0 = torch.empty_like(gT)
        torch.cuda.synchronize(gT.device)
        ga = problem.backward_ptr(gT.data_ptr(), g0.data_ptr())
        ga_t = torch.from_numpy(np.ascontiguousarray(ga[:, :ctx.a_shape[1]])).to(ctx.a_device, ctx.a_dtype)
        return g0, ga_t, None


def advect(u0, a, problem):
    """Cells after problem.steps frozen-mesh upwind steps, differentiable in u0 and a."""
    return _Advect.apply(u0, a, problem)
