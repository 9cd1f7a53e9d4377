"""Forward and backward of the stage over a population of node models: the
Voronoi realizations of one inversion step (the reference projects each
realization with voronoiproj_layer*.f90 and calls depthkernel/depthkernelTI per
realization, CalSurfGAniso_Joint.f90:478-486).

Columns are processed in memory-bounded chunks of fixed size.  One chunk is one
replay of a captured graph:
  stage (c, isotropic node kernels, Lsen_Gsc)  ->  grad_c = c - c_obs (0 where
  either is 0)  ->  effective-dVs VJP  g(z, col) = sum_k grad_c dc/dVs_eff.
The gradient with respect to the Voronoi cell values is the sum of g over the
nodes each cell covers (node z of column j takes cell(j, layer(z))).
"""
from __future__ import annotations

from . import stage
from .batched import GraphRunner


class RealizationStep:
    def __init__(self, nz, depz, periods, minthk, chunk, tune=None):
        import torch
        dev = depz.device
        K = periods.numel()
        self.nz, self.K, self.chunk = nz, K, chunk
        self.depz, self.periods, self.minthk, self.tune = depz, periods, minthk, tune
        self.vel = torch.full((nz, chunk), 3.5, dtype=torch.float32, device=dev)
        self.c_obs = torch.zeros((K, chunk), dtype=torch.float64, device=dev)
        self.gc = torch.zeros((K, chunk), dtype=torch.float64, device=dev)
        self.out = dict(pv=torch.zeros((K, chunk), dtype=torch.float64, device=dev),
                        status=torch.zeros(chunk, dtype=torch.int32, device=dev),
                        lsen=torch.zeros((nz - 1, K, chunk), dtype=torch.float32, device=dev))
        for k in ("sen_vs", "sen_vp", "sen_rho"):
            self.out[k] = torch.zeros((nz, K, chunk), dtype=torch.float64, device=dev)
        self.grad = dict(grad_vs=torch.zeros((nz, chunk), dtype=torch.float64, device=dev))
        self.graph = GraphRunner(self._step)

    def _step(self):
        import torch
        o = self.out
        stage.stage_dev(self.vel, self.depz, self.periods, self.minthk, iso=True, ti=True,
                        tune=self.tune, out=o)
        torch.sub(o["pv"], self.c_obs, out=self.gc)
        self.gc.mul_((o["pv"] > 0) & (self.c_obs > 0))
        stage.effective_dev(self.vel, o["sen_vs"], o["sen_vp"], o["sen_rho"], self.gc,
                            out=self.grad)

    def run(self, vel, c_obs, consume):
        """vel (nz, N) f32 and c_obs (K, N) f64 on the device; consume(c0, n, step)
        is called after each chunk (columns c0..c0+n-1) and reads step.out,
        step.gc and step.grad before the next chunk overwrites them."""
        N = vel.shape[1]
        for c0 in range(0, N, self.chunk):
            n = min(self.chunk, N - c0)
            self.vel[:, :n].copy_(vel[:, c0:c0 + n])
            self.c_obs[:, :n].copy_(c_obs[:, c0:c0 + n])
            if n < self.chunk:
                self.c_obs[:, n:].zero_()
            self.graph.replay()
            consume(c0, n, self)


def cell_gradient(grad_vs, layer_of_node, cell, ncells, c0, n, ncol_real, acc):
    """Adds the node gradient of chunk columns c0..c0+n-1 (grad_vs (nz, >= n)) to
    acc (nreal, nlayers, ncells) f64: the sum over the nodes of every cell.
    cell (nreal, nlayers, ncol_real) int64.  On the GPU the per-cell sums are
    the library's dazi_dk_cell_sum_dev (fixed summation order); CPU tensors
    take torch's index_add."""
    import torch
    nl = acc.shape[1]
    g = grad_vs[:, :n]
    # node -> layer sums as a product with the (nl, nz) 0/1 membership matrix
    # (a GEMM in a fixed order, where an index_add over the columns was an
    # atomic scatter of nz x n values)
    member = torch.nn.functional.one_hot(layer_of_node, nl).to(torch.float64).t().contiguous()
    gl = (member @ g).contiguous()
    if gl.is_cuda:
        from . import _lib
        if cell.dtype != torch.int64 or not cell.is_contiguous() or not acc.is_contiguous():
            raise ValueError("cell must be contiguous int64 and acc contiguous")
        rc = _lib.lib().dazi_dk_cell_sum_dev(
            gl.data_ptr(), n, nl, c0, n, ncol_real, acc.shape[0], cell.data_ptr(), ncells,
            acc.data_ptr(), stage._stream_ptr(None))
        _lib.check(rc)
        return
    cols = torch.arange(c0, c0 + n, device=g.device)
    r, j = cols // ncol_real, cols % ncol_real
    idx = (r[None, :] * nl + torch.arange(nl, device=g.device)[:, None]) * ncells + \
        cell[r[None, :], torch.arange(nl, device=g.device)[:, None], j[None, :]]
    acc.view(-1).index_add_(0, idx.reshape(-1), gl.reshape(-1))
