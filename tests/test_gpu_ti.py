"""GPU parity tests of the TI part (Lsen_Gsc, the dc/dA, dc/dL splice).

The reference quantity is tregn96's kernels at the float32 phase velocity of
surfdisp96, chained and spliced in float32 by depthkernelTI.  Bars: the
reference's TI A/B bands (run_ti_depthkernel_ab.py:297-304: lsen rel_l2 <= 1e-5,
max_abs <= 1e-4; pv rel_l2 <= 1e-6, max_abs <= 1e-5) and <= 1e-4 relative to the
per-(column, period) max-norm.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import port as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    m.lib()
    return m


def check_ti(pv, ls, pv_ref, ls_ref):
    assert np.array_equal(pv == 0, pv_ref == 0)
    assert np.abs(pv - pv_ref).max() <= 1e-5
    assert np.linalg.norm(pv - pv_ref) / np.linalg.norm(pv_ref) <= 1e-6
    d = ls - ls_ref
    assert np.linalg.norm(d) / np.linalg.norm(ls_ref) <= 1e-5
    assert np.abs(d).max() <= 1e-4
    scale = np.abs(ls_ref).max(axis=-1, keepdims=True)
    assert (np.abs(d) / scale).max() <= 1e-4


@pytest.mark.parametrize("fixture", ["ref_plain_1024.npz", "ref_plain_4096.npz"])
def test_ti_vs_reference_fixture(dk, fixture):
    f = load_golden(fixture)
    n = f["ti_pv"].shape[0]
    vel = f["vel"][:n, None, :]
    pv, ls = dk.run_depthkernelTI(vel, f["periods"], f["depz"], float(f["minthk"]))
    # fixture lsen is (col, K, nz-1)
    check_ti(pv, ls, f["ti_pv"], f["ti_lsen"])


@pytest.mark.skipif(not P.ref_available(), reason="reference library not built")
def test_ti_vs_live_reference_64_columns(dk):
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_4096", ncol=64)
    pv, ls = dk.run_depthkernelTI(W["vel"], W["periods"], W["depz"], W["minthk"])
    ref = [P.ti_column(W["vel"][j, 0], W["periods"], W["depz"], W["minthk"]) for j in range(64)]
    check_ti(pv, ls, np.array([r[0] for r in ref]), np.array([r[1] for r in ref]))


@pytest.mark.skipif(not P.ref_available(), reason="reference library not built")
def test_ti_deep_refinement_vs_live_reference(dk):
    """157 refined layers (three 64-layer chunks in the TI model prep)."""
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_1024", ncol=16)
    minthk = 12.0
    pv, ls = dk.run_depthkernelTI(W["vel"], W["periods"], W["depz"], minthk)
    nx = W["vel"].shape[0]
    ref = [P.ti_column(W["vel"][j % nx, j // nx], W["periods"], W["depz"], minthk)
           for j in range(16)]
    check_ti(pv, ls, np.array([r[0] for r in ref]), np.array([r[1] for r in ref]))


def test_fused_stage_equals_separate_calls(dk):
    import torch
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_4096", ncol=512)
    nz = W["vel"].shape[-1]
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a.reshape(-1, nz, order="F").T),  # noqa
                                  device="cuda")
    args = (T(W["vel"]), torch.as_tensor(W["depz"], device="cuda"),
            torch.as_tensor(W["periods"], device="cuda"), W["minthk"])
    a = dk.stage_dev(*args, iso=True, ti=True)
    a = {k: v.clone() for k, v in a.items()}
    b = dk.stage_dev(*args, iso=False, ti=True)
    c = dk.iso_dev(*args, kernels=True)
    assert torch.equal(a["pv"], b["pv"]) and torch.equal(a["pv"], c["pv"])
    assert torch.equal(a["lsen"], b["lsen"])
    for k in ("sen_vs", "sen_vp", "sen_rho"):
        assert torch.equal(a[k], c[k])
    assert torch.isfinite(a["lsen"]).all()
