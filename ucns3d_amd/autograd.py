"""torch.autograd bindings of the residual and of the SSP-RK3 march.

Forward and backward both run in libucnsgpu.so through the C ABI on device
pointers; torch only owns the tensors.  The backward of the march calls
ucns_rk3_adjoint_device, which re-runs the steps from the saved initial
state with every stage stored on the device and then sweeps back, so the
adjoint cost includes that recomputation.

    Residual.apply(u, solver)                        R(u)      (ncell, 5)
    RK3March.apply(u0, bc, solver, nsteps, dt, bt)   U_N       (ncell, 5)

u, u0: float64 CUDA tensors of shape (ncell, 5) (AOS cell averages); bc: a
float64 tensor of 10 (the conserved inflow and outflow states of
ucns_set_boundary) or None when the mesh has no inflow/outflow/farfield
faces; bt: the BOUNDTYPE of those faces.  The step size dt is fixed.
"""
from __future__ import annotations

import torch


def _dev(x: torch.Tensor) -> torch.Tensor:
    if x.dtype != torch.float64 or not x.is_cuda:
        raise TypeError("expected a float64 CUDA tensor")
    return x.contiguous()


def _sync() -> None:
    # the library works on its own stream: what torch queued must be done
    torch.cuda.current_stream().synchronize()


class Residual(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, solver):
        u = _dev(u)
        rhs = torch.empty_like(u)
        _sync()
        solver.set_state_device(u.data_ptr())
        solver.residual_device(rhs.data_ptr())
        ctx.solver = solver
        ctx.save_for_backward(u)
        return rhs

    @staticmethod
    def backward(ctx, g):
        (u,) = ctx.saved_tensors
        g = _dev(g)
        out = torch.empty_like(u)
        _sync()
        ctx.solver.set_state_device(u.data_ptr())
        ctx.solver.residual_vjp_device(g.data_ptr(), out.data_ptr())
        return out, None


class RK3March(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, bc, solver, nsteps: int, dt: float, boundtype: int = 1):
        u0 = _dev(u0)
        b = None
        if bc is not None:
            b = bc.detach().cpu().double().numpy().copy()
            solver.set_boundary(boundtype, b[:5], b[5:])
        _sync()
        solver.set_state_device(u0.data_ptr())
        solver.rk3_steps(nsteps, dt_fixed=dt)
        uN = torch.empty_like(u0)
        solver.get_state_device(uN.data_ptr())
        ctx.solver, ctx.nsteps, ctx.dt, ctx.has_bc = solver, nsteps, dt, bc is not None
        ctx.bc_device = bc.device if bc is not None else None
        # the free-stream data of this forward: a later forward on the same
        # solver may have set others before this backward runs
        ctx.bc_host, ctx.boundtype = b, boundtype
        ctx.save_for_backward(u0)
        return uN

    @staticmethod
    def backward(ctx, g):
        (u0,) = ctx.saved_tensors
        g = _dev(g)
        u0b = torch.empty_like(u0)
        bcb = torch.zeros(10, dtype=torch.float64, device=u0.device)
        _sync()
        if ctx.bc_host is not None:
            ctx.solver.set_boundary(ctx.boundtype, ctx.bc_host[:5], ctx.bc_host[5:])
        ctx.solver.set_state_device(u0.data_ptr())
        ctx.solver.rk3_adjoint_device(ctx.nsteps, g.data_ptr(), u0b.data_ptr(), bcb.data_ptr(), dt_fixed=ctx.dt)
        return u0b, (bcb.to(ctx.bc_device) if ctx.has_bc else None), None, None, None, None
