"""PyTorch binding of the discrete adjoint: the calibration loss of a Solver as a
differentiable function of its initial state and aux fields.

    solver.adjoint_setup(nsteps, gauges, obs_every, obs)
    L = calibration_loss(solver, X0, A0)          # X0: (ncells, n_q), A0: (ncells, n_aux)
    L.backward()                                   # X0.grad[:, 0] = dL/d(bed), A0.grad = dL/d(aux)

Forward and backward run the HIP library through the solver's C ABI (the forward records its
time steps, the backward is one reverse sweep over the checkpoints).  Tensors on the GPU are
handed over as device addresses (zdm_adjoint_forward_device / zdm_adjoint_gradient_device):
the solver's stream first waits for torch's current stream, and torch's current stream waits
for the solver's before the gradients are used, so the exchange stays on the device and in
order with the caller's work.  CPU tensors go through host memory.
"""
import numpy as np
import torch


def _solver_stream(solver, device):
    return torch.cuda.ExternalStream(solver.stream_handle(), device=device)


class CalibrationLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, X0, A0, solver, dts=None):
        ctx.solver = solver
        ctx.dev = X0.device
        ctx.has_a = A0 is not None
        if X0.is_cuda:
            X = X0.detach().to(torch.float64).contiguous()
            A = None if A0 is None else A0.detach().to(device=X.device, dtype=torch.float64).contiguous()
            ss = _solver_stream(solver, X.device)
            ss.wait_stream(torch.cuda.current_stream(X.device))   # X, A are written
            # returns after the forward run has read X and A (the loss needs the whole run)
            L = solver.adjoint_forward_device(X.data_ptr(), A.data_ptr() if A is not None else 0, dts=dts)
        else:
            X = X0.detach().to(torch.float64).contiguous().numpy()
            A = None if A0 is None else A0.detach().to(torch.float64).contiguous().numpy()
            L = solver.adjoint_forward(X, A, dts=dts)
        solver._adj_runs = getattr(solver, "_adj_runs", 0) + 1
        ctx.run = solver._adj_runs
        return torch.tensor(L, dtype=torch.float64, device=X0.device)

    @staticmethod
    def backward(ctx, grad_out):
        s = ctx.solver
        if s._adj_runs != ctx.run:
            raise RuntimeError("the solver ran another forward since this loss was computed")
        if ctx.dev.type == "cuda":
            n, nq, na = s.mesh.ncells, s.info.n_q, s.info.n_aux
            gx = torch.empty((n, nq), dtype=torch.float64, device=ctx.dev)
            ga = torch.empty((n, na), dtype=torch.float64, device=ctx.dev) if ctx.has_a else None
            cur = torch.cuda.current_stream(ctx.dev)
            ss = _solver_stream(s, ctx.dev)
            ss.wait_stream(cur)       # gx, ga are allocated on the current stream
            s.adjoint_gradient_device(gx.data_ptr(), ga.data_ptr() if ga is not None else 0)
            cur.wait_stream(ss)       # the gradients are written before torch reads them
            go = grad_out.to(device=ctx.dev, dtype=torch.float64)
            gx.mul_(go)
            if ga is not None:
                ga.mul_(go)
            return gx, ga, None, None
        gX, gA = s.adjoint_gradient()
        go = grad_out.to("cpu", torch.float64)
        gx = torch.from_numpy(np.ascontiguousarray(gX)) * go
        ga = torch.from_numpy(np.ascontiguousarray(gA)) * go if ctx.has_a else None
        return gx.to(ctx.dev), (ga.to(ctx.dev) if ga is not None else None), None, None


def calibration_loss(solver, X0, A0=None, dts=None):
    """Differentiable calibration loss (see Solver.adjoint_setup for the loss)."""
    return CalibrationLoss.apply(X0, A0, solver, dts)
