"""PyTorch binding of the discrete adjoint: the calibration loss of a Solver as a
differentiable function of its initial state and aux fields.

    solver.adjoint_setup(nsteps, gauges, obs_every, obs)
    L = calibration_loss(solver, X0, A0)          # X0: (ncells, n_q), A0: (ncells, n_aux)
    L.backward()                                   # X0.grad[:, 0] = dL/d(bed), A0.grad = dL/d(aux)

Forward and backward run the HIP library through the solver's C ABI (the forward records its
time steps, the backward is one reverse sweep over the checkpoints); the tensors are
float64 on any device and are staged through host memory.
"""
import numpy as np
import torch


class CalibrationLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X0, A0, solver, dts=None):
        X = X0.detach().to("cpu", torch.float64).contiguous().numpy()
        A = None if A0 is None else A0.detach().to("cpu", torch.float64).contiguous().numpy()
        L = solver.adjoint_forward(X, A, dts=dts)
        solver._adj_runs = getattr(solver, "_adj_runs", 0) + 1
        ctx.solver = solver
        ctx.run = solver._adj_runs
        ctx.dev = X0.device
        ctx.has_a = A0 is not None
        return torch.tensor(L, dtype=torch.float64, device=X0.device)

    @staticmethod
    def backward(ctx, grad_out):
        s = ctx.solver
        if s._adj_runs != ctx.run:
            raise RuntimeError("the solver ran another forward since this loss was computed")
        gX, gA = s.adjoint_gradient()
        go = grad_out.to("cpu", torch.float64)
        gx = torch.from_numpy(gX) * go
        ga = torch.from_numpy(gA) * go if ctx.has_a else None
        return gx.to(ctx.dev), (ga.to(ctx.dev) if ga is not None else None), None, None


def calibration_loss(solver, X0, A0=None, dts=None):
    """Differentiable calibration loss (see Solver.adjoint_setup for the loss)."""
    return CalibrationLoss.apply(X0, A0, solver, dts)
