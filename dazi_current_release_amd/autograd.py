"""torch.autograd.Function over the stage's ragged layered solver.

forward: Rayleigh fundamental-mode phase velocities c (float64, before the
reference's float32 rounding) of a LayeredBatch for given layer vs, vp, rho and
thickness.  backward: the hand-written VJP kernel (implicit differentiation of
the secular function at each root, dc/dm = -(dF/dm)/(dF/dc)) returns the
gradients with respect to vs, vp, rho and thickness of every layer (the
thickness gradient carries the earth flattening of the deeper layers and is 0
for the half-space).  Periods whose root was not found (c = 0) contribute no
gradient, as no kernel exists there.
"""
from __future__ import annotations

import torch

from . import stage


class Dispersion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vs, vp, rho, thk, batch, tune=None):
        m = [t.detach().contiguous() for t in (vs, vp, rho, thk)]
        out = stage.layered_dev(batch, vs=m[0], vp=m[1], rho=m[2], thk=m[3], tune=tune)
        c = out["c_unrounded"]
        ctx.batch = batch
        ctx.save_for_backward(*m, c)
        ctx.mark_non_differentiable(out["status"])
        return c, out["status"]

    @staticmethod
    def backward(ctx, grad_c, grad_status):
        vs, vp, rho, thk, c = ctx.saved_tensors
        g = stage.layered_vjp(ctx.batch, c, grad_c, vs=vs, vp=vp, rho=rho, thk=thk)
        return (g[0].to(vs.dtype), g[1].to(vp.dtype), g[2].to(rho.dtype), g[3].to(thk.dtype),
                None, None)


def dispersion(vs, vp, rho, batch, thk=None, tune=None):
    """c (nsolves,) float64 and status (ncol,) for layer tensors laid out like batch.vs;
    thk defaults to the batch's thicknesses."""
    return Dispersion.apply(vs, vp, rho, batch.thk if thk is None else thk, batch, tune)


class TIDispersion(torch.autograd.Function):
    """Phase velocities of node models with a first-order transversely
    isotropic perturbation of the refined layers' moduli:
        c(k, col) = c_iso(k, col) + sum_m dc/dA_m dA(m, col) + dc/dL_m dL(m, col),
    where c_iso is the stage's Rayleigh phase velocity and dc/dA, dc/dL are
    tregn96's kernels chained as depthkernelTI chains them
    (depthkernelTI.f90:152-162), at the isotropic root.  dA, dL are changes of
    A (= C) and L (= N) of every refined layer (L_layers, ncol), float64.
    backward: the TI VJP kernel, grad_dA = sum_k grad_c dc/dA and grad_dL likewise
    (periods without a phase velocity contribute nothing)."""

    @staticmethod
    def forward(ctx, dA, dL, vel, depz, periods, minthk):
        o = stage.ti_kernels_dev(vel, depz, periods, minthk, lsen=False)
        kA, kL = o["dc_dA"], o["dc_dL"]
        ok = o["pv"] > 0
        c = o["pv"] + torch.where(ok, (kA.double() * dA.detach()[:, None, :]).sum(0) +
                                  (kL.double() * dL.detach()[:, None, :]).sum(0),
                                  torch.zeros_like(o["pv"]))
        ctx.save_for_backward(kA, kL, ok)
        ctx.mark_non_differentiable(o["status"])
        return c, o["status"]

    @staticmethod
    def backward(ctx, grad_c, grad_status):
        kA, kL, ok = ctx.saved_tensors
        gA, gL = stage.ti_vjp_dev(kA, kL, grad_c * ok)
        return gA, gL, None, None, None, None


def ti_dispersion(dA, dL, vel, depz, periods, minthk):
    """c (K, ncol) float64 and status (ncol,) of node models (vel (nz, ncol)
    float32) with first-order TI perturbations dA, dL (L_layers, ncol) float64
    of the refined layers' moduli; differentiable in dA and dL."""
    return TIDispersion.apply(dA, dL, vel, depz, periods, minthk)
