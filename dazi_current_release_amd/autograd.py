"""torch.autograd.Function over the stage's ragged layered solver.

forward: Rayleigh fundamental-mode phase velocities c (float64, before the
reference's float32 rounding) of a LayeredBatch for given layer vs, vp, rho.
backward: the hand-written VJP kernel (implicit differentiation of the secular
function at each root, dc/dm = -(dF/dm)/(dF/dc)) returns the gradients with
respect to vs, vp and rho of every layer.  Periods whose root was not found
(c = 0) contribute no gradient, as no kernel exists there.
"""
from __future__ import annotations

import torch

from . import stage


class Dispersion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vs, vp, rho, batch, tune=None):
        out = stage.layered_dev(batch, vs=vs.detach().contiguous(), vp=vp.detach().contiguous(),
                                rho=rho.detach().contiguous(), tune=tune)
        c = out["c_unrounded"]
        ctx.batch = batch
        ctx.save_for_backward(vs, vp, rho, c)
        ctx.mark_non_differentiable(out["status"])
        return c, out["status"]

    @staticmethod
    def backward(ctx, grad_c, grad_status):
        vs, vp, rho, c = ctx.saved_tensors
        gvs, gvp, grho = stage.layered_vjp(ctx.batch, c, grad_c, vs=vs.detach().contiguous(),
                                           vp=vp.detach().contiguous(),
                                           rho=rho.detach().contiguous())
        return gvs.to(vs.dtype), gvp.to(vp.dtype), grho.to(rho.dtype), None, None


def dispersion(vs, vp, rho, batch, tune=None):
    """c (nsolves,) float64 and status (ncol,) for layer tensors laid out like batch.vs."""
    return Dispersion.apply(vs, vp, rho, batch, tune)
