"""Regression vectors of the kernel passes, written by one build of this
library on a GPU: node kernels (dc/dVs, dc/dVp, dc/drho at the nodes) of 32
plain_1024 columns, layer kernels of 8 diff_4096 layered columns (30 periods)
and the layered VJP for a seeded grad_c on them.

The inputs are stored with the outputs, so the fixture does not depend on the
workload generators.  Run on a GPU box:

    python tests/golden/make_kernel_regression.py --pkg-root <checkout> --out <file.npz>

where <checkout> holds the build to record (its dazi_current_release_amd
package with libdazi_dk.so built in place).
"""
import argparse
import os
import sys

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pkg-root", required=True)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    sys.path.insert(0, os.path.abspath(a.pkg_root))
    import torch
    import dazi_current_release_amd as dk
    from dazi_current_release_amd import configs
    print("package:", dk.__file__, flush=True)
    assert os.path.abspath(dk.__file__).startswith(os.path.abspath(a.pkg_root))
    out = {}
    # node kernels: 32 columns of plain_1024
    W = configs.node_workload("plain_1024")
    nz = W["vel"].shape[-1]
    vel = np.ascontiguousarray(W["vel"].reshape(-1, nz)[:32])
    pv, svs, svp, srho = dk.run_depthkernel(vel[:, None, :], W["periods"], W["depz"],
                                            float(W["minthk"]))
    out.update(node_vel=vel, node_depz=np.asarray(W["depz"], np.float32),
               node_periods=np.asarray(W["periods"], np.float64),
               node_minthk=np.float32(W["minthk"]), node_pv=pv, node_sen_vs=svs,
               node_sen_vp=svp, node_sen_rho=srho)
    # layer kernels and VJP: 8 columns of diff_4096
    L = configs.layered_workload("diff_4096")
    n = 8
    nl, nk = L["nl"][:n], L["nk"][:n]
    nlay = int(nl.sum())
    per = np.asarray(L["periods"])
    per = per[:int(nk.sum())] if per.ndim == 1 and per.size == int(L["nk"].sum()) else per
    sl = slice(0, nlay)
    thk, vp, vs, rho = (np.asarray(L[k])[sl] for k in ("thk", "vp", "vs", "rho"))
    b = dk.LayeredBatch(nl, nk, thk, vp, vs, rho, per)
    o = dk.layered_dev(b, kernels=True)
    gc = np.random.default_rng(31).standard_normal(b.nsolves)
    g = dk.layered_vjp(b, o["c_unrounded"], torch.as_tensor(gc, device="cuda"))
    out.update(lay_nl=nl, lay_nk=nk, lay_thk=thk, lay_vp=vp, lay_vs=vs, lay_rho=rho,
               lay_periods=per, lay_c=o["c_unrounded"].cpu().numpy(),
               lay_dc_dvs=o["dc_dvs"].cpu().numpy(), lay_dc_dvp=o["dc_dvp"].cpu().numpy(),
               lay_dc_drho=o["dc_drho"].cpu().numpy(), vjp_grad_c=gc,
               vjp_vs=g[0].cpu().numpy(), vjp_vp=g[1].cpu().numpy(), vjp_rho=g[2].cpu().numpy())
    np.savez_compressed(a.out, **out)
    print("wrote", a.out, {k: np.shape(v) for k, v in out.items()}, flush=True)


if __name__ == "__main__":
    main()
