"""torch.autograd wrapper of the right-hand side and its hand-written adjoint (a thin harness).

    R = rhs_function(ctx, mesh, params, ucont, ucat, theta, cs, scale=1.0, buoy_scale=1.0)
    R.backward(rbar)      # ucont.grad, ucat.grad, theta.grad, cs.grad from tosca_rhs_vjp

R(ucont, ucat, theta, cs) = FormU(ucont, ucat, nu_t) + Buoyancy(theta) with nu_t = UpdateNut(ucat, cs),
evaluated by the library's own entry points (tosca_update_nut, tosca_formu onto a zero rhs,
tosca_buoyancy); the backward pass calls tosca_rhs_vjp, which recomputes nu_t and the face
quantities instead of keeping anything from the forward pass.  All tensors are float64 CUDA
tensors in the DMDA local layout; R has the owned-range shape (mz, my, mx, 3).
"""
from __future__ import annotations

import torch


class _RhsFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, ctx, mesh, params, scale, buoy_scale, ucont, ucat, theta, cs):
        info = ctx.info
        shape = (info.mz, info.my, info.mx, 3)
        nut = torch.zeros_like(cs)
        ctx.UpdateNut(mesh, params, ucat, cs, nut)
        rhs = torch.zeros(shape, dtype=torch.float64, device=ucat.device)
        ctx.FormU(mesh, params, ucont, ucat, nut, rhs, scale, 0)
        bt = torch.zeros_like(rhs)
        ctx.Buoyancy(mesh, params, theta, buoy_scale, bt)
        fctx.save_for_backward(ucont, ucat, theta, cs)
        fctx.rhs_ctx, fctx.mesh, fctx.params, fctx.scale, fctx.buoy_scale = ctx, mesh, params, scale, buoy_scale
        return rhs + bt

    @staticmethod
    def backward(fctx, rbar):
        ucont, ucat, theta, cs = fctx.saved_tensors
        bars = [torch.empty_like(t) for t in (ucont, ucat, theta, cs)]
        fctx.rhs_ctx.rhs_vjp(fctx.mesh, fctx.params, ucont, ucat, theta, cs, rbar.contiguous(), fctx.scale,
                             fctx.buoy_scale, *bars)
        return (None, None, None, None, None) + tuple(bars)


def rhs_function(ctx, mesh, params, ucont, ucat, theta, cs, scale=1.0, buoy_scale=1.0):
    return _RhsFunction.apply(ctx, mesh, params, float(scale), float(buoy_scale), ucont, ucat, theta, cs)
