"""GPU tests of the period azimuthal-anisotropy map (dazi_dk_azimuth_map_dev /
dazi_dk_azimuth_map, Fortran FwdAzimuthalAniMap_hip) against the oracle, which
reproduces the reference's own output line for line.

Inputs: the GPU stage's Lsen_Gsc and phase velocities on a 32x32-column grid,
random Gc, Gs layer models.  Bars: cos, sin, amp, amp_rela bit-exact; angle
within 2 units in the last place (the reference's libm atan2f is not
correctly rounded; the device rounds a float64 atan2).
"""
import numpy as np
import pytest

from oracle import port as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    import dazi_current_release_amd as m
    from dazi_current_release_amd import configs
    W = configs.node_workload("plain_1024")
    nz = W["vel"].shape[-1]
    nx, ny = W["nx"], W["ny"]
    vel = torch.as_tensor(np.ascontiguousarray(W["vel"].reshape(-1, nz, order="F").T),
                          device="cuda")
    o = m.stage_dev(vel, torch.as_tensor(W["depz"], device="cuda"),
                    torch.as_tensor(W["periods"], device="cuda"), W["minthk"], iso=False, ti=True)
    rng = np.random.default_rng(21)
    nvx, nvz = nx - 2, ny - 2
    gcf = (rng.standard_normal((nz - 1, nvz, nvx)) * 0.03).astype(np.float32)
    gsf = (rng.standard_normal((nz - 1, nvz, nvx)) * 0.03).astype(np.float32)
    jj, ii = np.meshgrid(np.arange(1, nvz + 1), np.arange(1, nvx + 1), indexing="ij")
    post = (jj * nx + ii).ravel()
    tRcV = np.ascontiguousarray(o["pv"].cpu().numpy()[:, post])
    return m, W, nx, ny, nz, o, gcf, gsf, tRcV


def ulps(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def check(m, want):
    for k in ("cos", "sin", "amp", "amp_rela"):
        assert np.array_equal(m[k], want[k]), k
    assert ulps(m["angle"], want["angle"]).max() <= 2


def test_device_map_matches_the_oracle(setup):
    import torch
    from dazi_current_release_amd import azimuth
    m, W, nx, ny, nz, o, gcf, gsf, tRcV = setup
    T = lambda a: torch.as_tensor(a, device="cuda")  # noqa: E731
    g = azimuth.azimuth_map_dev(o["lsen"], T(gcf), T(gsf), T(tRcV), nx, ny)
    got = {k: v.cpu().numpy() for k, v in g.items()}
    want = P.azimuth_map(nx - 2, ny - 2, o["lsen"].cpu().numpy(), gcf, gsf, tRcV)
    check(got, want)
    print("angle ulps: max", ulps(got["angle"], want["angle"]).max(),
          "nonzero", int((ulps(got["angle"], want["angle"]) > 0).sum()), "of", got["angle"].size)
    h = azimuth.azimuth_map(o["lsen"].cpu().numpy(), gcf, gsf, tRcV, nx, ny)
    for k in h:
        assert np.array_equal(h[k], got[k]), k


def test_fortran_binding_writes_the_reference_lines(setup, tmp_path):
    """FwdAzimuthalAniMap_hip from the example Fortran host: its unit-42 lines
    against the oracle's lines (the angle column within its 2-ulp bar)."""
    import fortran_io as FIO
    from dazi_current_release_amd.azimuth import azm_lines
    m, W, nx, ny, nz, o, gcf, gsf, tRcV = setup
    vel = np.asfortranarray(W["vel"].reshape(nx, ny, nz, order="F"))
    fazin, fazout = tmp_path / "azm_in.bin", tmp_path / "azm.txt"
    with open(fazin, "wb") as fh:
        np.array([100.0, 30.0, 0.5, 0.25], np.float32).tofile(fh)
        np.transpose(gcf, (2, 1, 0)).ravel(order="F").tofile(fh)
        np.transpose(gsf, (2, 1, 0)).ravel(order="F").tofile(fh)
        tRcV.T.ravel(order="F").tofile(fh)
    r, fout = FIO.run(str(tmp_path), vel, W["depz"], W["periods"], float(W["minthk"]),
                      extra=["-", str(tmp_path / "g.bin"), str(fazin), str(fazout)])
    assert r.returncode == 0, r.stdout + r.stderr
    got = open(fazout).read().splitlines()
    want = azm_lines(P.azimuth_map(nx - 2, ny - 2, o["lsen"].cpu().numpy(), gcf, gsf, tRcV),
                     tRcV, W["periods"], nx, ny, 100.0, 30.0, 0.5, 0.25)
    assert len(got) == len(want)
    g = np.array([[float(l[10 * q:10 * q + 10]) for q in range(9)] for l in got])
    w = np.array([[float(l[10 * q:10 * q + 10]) for q in range(9)] for l in want])
    assert np.array_equal(np.delete(g, 4, axis=1), np.delete(w, 4, axis=1))
    # 2 ulp of a float32 angle below 256 degrees is 3.05e-5, plus the print rounding
    assert np.abs(g[:, 4] - w[:, 4]).max() <= 4e-5
