"""The kernel passes against recorded outputs of an earlier build of this
library (tests/golden/kernel_regression.npz, written by
tests/golden/make_kernel_regression.py): the evaluation of the layer algebra
may change (faster exponentials, sparse partials), the kernels may not,
beyond 1e-12 of each solve's largest kernel value at the same root.  The root
search may change how it evaluates the secular function (a positive multiple
of the reference's): the float32 phase velocities stay bit-identical, the
unrounded root moves inside its refinement bracket (2e-6 km/s wide; the chord
root through it is within about 1e-11 km/s of the exact root either way), and
the kernels, evaluated at that root, move by dK/dc times that shift (a root
shift of 2e-12 km/s moved layer kernels by up to 7e-10 of the row maximum).
The bound for them is then 1e-8 of the row maximum, four orders below the
1e-4 parity bar and far below any change of the kernel algebra."""
import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu
TOL = 1e-8
TOL_C = 1e-9  # km/s, unrounded root


@pytest.fixture(scope="module")
def dk():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    m.lib()
    return m


def rel_to_max(got, want, axis):
    s = np.abs(want).max(axis=axis, keepdims=True)
    s[s == 0] = 1.0
    return float((np.abs(got - want) / s).max())


def test_node_kernels_match_recorded_build(dk):
    f = load_golden("kernel_regression.npz")
    vel = f["node_vel"][:, None, :]
    pv, svs, svp, srho = dk.run_depthkernel(vel, f["node_periods"], f["node_depz"],
                                            float(f["node_minthk"]))
    assert np.array_equal(pv, f["node_pv"])
    for name, got in (("node_sen_vs", svs), ("node_sen_vp", svp), ("node_sen_rho", srho)):
        err = rel_to_max(got, f[name], axis=-1)
        print(name, err)
        assert err <= TOL, (name, err)


def test_layer_kernels_and_vjp_match_recorded_build(dk):
    import torch
    f = load_golden("kernel_regression.npz")
    b = dk.LayeredBatch(f["lay_nl"], f["lay_nk"], f["lay_thk"], f["lay_vp"], f["lay_vs"],
                        f["lay_rho"], f["lay_periods"])
    o = dk.layered_dev(b, kernels=True)
    c = o["c_unrounded"].cpu().numpy()
    assert np.array_equal(c == 0, f["lay_c"] == 0)
    dc = float(np.abs(c - f["lay_c"]).max())
    print("unrounded root shift", dc)
    assert dc <= TOL_C, dc
    for key in ("dc_dvs", "dc_dvp", "dc_drho"):
        err = rel_to_max(o[key].cpu().numpy(), f["lay_" + key], axis=0)
        print(key, err)
        assert err <= TOL, (key, err)
    g = dk.layered_vjp(b, o["c_unrounded"], torch.as_tensor(f["vjp_grad_c"], device="cuda"))
    for name, got in zip(("vjp_vs", "vjp_vp", "vjp_rho"), g[:3]):
        want = f[name]
        err = float(np.abs(got.cpu().numpy() - want).max() / np.abs(want).max())
        print(name, err)
        assert err <= TOL, (name, err)
