"""CPU tests of the oracle (the C restatement in oracle/surfdisp96_port.c).

The oracle is pinned by: the reference's golden phase-velocity map of the
Yunnan example, fixtures written by the reference's own Fortran (flang build),
the live reference library when it is built here, and closed forms / brute
force that a dropped term or a transposed operand would fail.
"""
import numpy as np
import pytest

from conftest import load_golden
from oracle import port as P


def test_brocher_and_gtsolh_match_reference_fixture():
    f = load_golden("ref_functions.npz")
    got = np.array([P.gtsolh(a, b) for a, b in zip(f["gt_a"], f["gt_b"])], np.float32)
    assert np.array_equal(got, f["gt_c"])


def test_dltar4_matches_reference_fixture():
    f = load_golden("ref_functions.npz")
    for row in f["dl"]:
        c, T, v = row[1], row[2], row[3]
        flat = row[4:].astype(np.float32).reshape(4, -1)
        got = P.dltar4(c, T, *flat)
        assert got == pytest.approx(v, rel=1e-12, abs=1e-300)


@pytest.mark.parametrize("name", ["ref_plain_1024.npz", "ref_plain_4096.npz",
                                  "ref_voronoi.npz"])
def test_column_fd_bit_exact_vs_reference_fixture(name):
    f = load_golden(name)
    n = 8 if "1024" in name else 4
    for j in range(n):
        pv, s1, s2, s3 = P.column_fd(f["vel"][j], f["periods"], f["depz"], float(f["minthk"]))
        assert np.array_equal(pv, f["pv"][j])
        assert np.array_equal(s1, f["sen_vs"][j])
        assert np.array_equal(s2, f["sen_vp"][j])
        assert np.array_equal(s3, f["sen_rho"][j])


def test_tight_layer_fd_density_invariant_and_bracket():
    """The layered FD oracle (tight roots, Richardson) against the exact
    identities of dispersion: sum rho dc/drho = 0 (c does not change when every
    density scales), and sum (alpha dc/dalpha + beta dc/dbeta) = c + T dc/dT
    (scaling every velocity by s scales c by s at period T/s), with dc/dT from
    tight roots at T(1 -/+ eps).  A dropped layer or a wrong step size fails."""
    from dazi_current_release_amd import configs
    W = configs.layered_workload("diff_4096", ncol=1)
    L, K = int(W["nl"][0]), int(W["nk"][0])
    mdl = [W[k][:L] for k in ("thk", "vp", "vs", "rho")]
    per = W["periods"][:K]
    c, (dvs, dvp, drho) = P.layer_kernels_fd(*mdl, per, nthreads=4)
    assert np.abs((mdl[3][None, :] * drho).sum(1) / c).max() < 1e-5
    eps = 1e-5
    cp, _ = P.forward_layered([L], [K], *mdl, per * (1 + eps), tol_rel=1e-13, round_f32=False)
    cm, _ = P.forward_layered([L], [K], *mdl, per * (1 - eps), tol_rel=1e-13, round_f32=False)
    rhs = c + per * (cp - cm) / (2 * eps * per)
    lhs = (mdl[2][None, :] * dvs + mdl[1][None, :] * dvp).sum(1)
    assert np.abs(lhs - rhs).max() / c.max() < 1e-5


def test_ragged_surfdisp96_bit_exact_vs_reference_fixture():
    f = load_golden("ref_ragged.npz")
    c, fail = P.forward_layered(f["nl"], f["nk"], f["thk"], f["vp"], f["vs"], f["rho"],
                                f["periods"], nthreads=4)
    assert np.array_equal(c, f["c"])


def test_yunnan_golden_phase_velocity_map():
    """All 51,840 values of period_phaseVMOD.dat within 1 LSB of f10.4."""
    g = load_golden("yunnan_golden.npz")
    vel = g["vel"]
    nx, ny, nz = vel.shape
    pv = P.depthkernel(vel, g["periods"], g["depz"], float(g["minthk"]), kernels=False,
                       nthreads=8)
    ii, jj = np.meshgrid(np.arange(1, nx - 1), np.arange(1, ny - 1), indexing="ij")
    post = (jj * nx + ii).T.reshape(-1)          # interior columns in file order
    c = pv[post]
    gold = g["c_gold_e4"] * 1e-4
    assert np.abs(c - gold).max() <= 1e-4 + 1e-9
    assert (np.abs(np.round(c, 4) - gold) > 1e-9).mean() < 0.002


@pytest.mark.skipif(not P.ref_available(), reason="reference library not built here")
def test_port_matches_live_reference_on_random_models():
    rng = np.random.default_rng(7)
    per = np.geomspace(1.5, 50, 12)
    for _ in range(40):
        L = int(rng.integers(3, 12))
        thk = rng.uniform(0.5, 15, L).astype(np.float32)
        vs = np.sort(rng.uniform(2.0, 4.7, L)).astype(np.float32)
        if rng.uniform() < 0.3:
            vs[1:3] *= 0.85                           # low-velocity zone
        vp, rho = P.brocher(vs)
        a = P.surfdisp96(thk, vp, vs, rho, per)
        b = P.surfdisp96(thk, vp, vs, rho, per, lib="ref")
        assert np.array_equal(a, b)


# ---------------------------------------------------------------- closed forms
def rayleigh_halfspace(a, b):
    """Rayleigh speed of a half-space: root of (2-k2)^2 = 4 sqrt(1-g2k2) sqrt(1-k2)."""
    lo, hi = 0.5 * b, b * (1 - 1e-15)
    g2 = (b / a) ** 2
    f = lambda c: (2 - (c / b) ** 2) ** 2 - 4 * np.sqrt(1 - g2 * (c / b) ** 2) * np.sqrt(1 - (c / b) ** 2)  # noqa
    for _ in range(200):
        m = 0.5 * (lo + hi)
        if f(lo) * f(m) <= 0:
            hi = m
        else:
            lo = m
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("vs", [2.0, 3.3, 4.5])
def test_uniform_halfspace_is_rayleigh_speed(vs):
    """Flat uniform medium: c is T-independent and equals the half-space Rayleigh speed."""
    L = 3
    thk = np.array([5, 5, 0], np.float32)
    b = np.full(L, vs, np.float32)
    a = (b * np.float32(1.8)).astype(np.float32)
    rho = np.full(L, 2.7, np.float32)
    c = P.surfdisp96(thk, a, b, rho, [2.0, 10.0, 40.0], iflsph=0, tol_rel=1e-13,
                     round_f32=False)
    cr = rayleigh_halfspace(float(a[0]), float(b[0]))
    assert np.allclose(c, cr, rtol=1e-9, atol=0)


def test_layer_over_halfspace_limits():
    """c -> c_R(top) as T -> 0 and c -> c_R(bottom) as T -> inf (flat earth)."""
    thk = np.array([3.0, 0], np.float32)
    b = np.array([3.0, 4.2], np.float32)
    a = (b * np.float32(1.75)).astype(np.float32)
    rho = np.array([2.6, 3.2], np.float32)
    c = P.surfdisp96(thk, a, b, rho, [0.05, 3000.0, 30000.0], iflsph=0, tol_rel=1e-13,
                     round_f32=False)
    assert c[0] == pytest.approx(rayleigh_halfspace(float(a[0]), float(b[0])), rel=1e-6)
    cb = rayleigh_halfspace(float(a[1]), float(b[1]))
    # first-order approach in k*h: ten times the period, a tenth of the deviation
    d1, d2 = cb - c[1], cb - c[2]
    assert 0 < d2 < 3e-5 * cb and 5 < d1 / d2 < 20


def test_root_is_first_sign_change_by_brute_force():
    """Dense bisection of the secular function on 2-3-layer models finds the same root."""
    rng = np.random.default_rng(3)
    for _ in range(6):
        L = int(rng.integers(2, 4))
        thk = rng.uniform(2, 10, L).astype(np.float32)
        b = np.sort(rng.uniform(2.5, 4.5, L)).astype(np.float32)
        a, rho = P.brocher(b)
        T = float(rng.uniform(3, 30))
        c = P.surfdisp96(thk, a, b, rho, [T], tol_rel=1e-13, round_f32=False)[0]
        _, flat = P.surfdisp96(thk, a, b, rho, [T], want_flat=True)
        grid = np.linspace(0.5 * flat[2].min(), c + 1e-3, 4000)
        s = np.sign([P.dltar4(x, T, *flat) for x in grid])
        j = int(np.argmax(s != s[0]))
        lo, hi = grid[j - 1], grid[j]
        for _ in range(80):
            m = 0.5 * (lo + hi)
            if np.sign(P.dltar4(m, T, *flat)) == s[0]:
                lo = m
            else:
                hi = m
        assert abs(0.5 * (lo + hi) - c) < 1e-9


def test_fd_kernel_density_scaling_invariant():
    """sum_nodes rho * dc/drho = 0: c is unchanged when all densities scale."""
    f = load_golden("ref_plain_1024.npz")
    vs = f["vel"][0]
    vp, rho = P.brocher(vs)
    pv, s1, s2, s3 = P.column_fd(vs, f["periods"], f["depz"], float(f["minthk"]),
                                 tol_rel=1e-13, round_f32=False)
    tot = (s3 * rho[None, :]).sum(axis=1)
    assert np.abs(tot).max() < 2e-4 * np.abs(s3 * rho[None, :]).max()


def test_failure_zero_fill_matches_reference_fixture():
    """Models on which the reference's root search fails part way: the port
    zero-fills from the same period with the same values before it."""
    f = load_golden("ref_ti_failures.npz")
    pv = P.depthkernel(f["vel"][:, None, :], f["periods"], f["depz"], float(f["minthk"]),
                       lib="port", kernels=False)
    assert np.array_equal(pv, f["pv"])
    assert np.array_equal(f["pv"] == 0, f["ti_pv"] == 0)
    # the reference's TI splice is NaN exactly where its root search failed
    assert np.array_equal(np.isnan(f["ti_lsen"]).any(axis=2), f["ti_pv"] == 0)


@pytest.mark.skipif(not P.ref_available(), reason="reference library not built")
def test_ti_at_reproduces_reference_columns(tmp_path, monkeypatch):
    """ti_at (the reference's tregn96 at given phase velocities, model and splice
    restated) equals the reference's own depthkernel_ti_column bit for bit when
    given that column's phase velocities; NaN where c = 0."""
    monkeypatch.chdir(tmp_path)  # the reference writes unit-66 messages to cwd
    f = load_golden("ref_ti_failures.npz")
    for j in range(f["vel"].shape[0]):
        ls = P.ti_at(f["vel"][j], f["periods"], f["ti_pv"][j].astype(np.float32), f["depz"],
                     float(f["minthk"]))
        ok = f["ti_pv"][j] > 0
        assert np.array_equal(ls[ok], f["ti_lsen"][j][ok])
        assert np.isnan(ls[~ok]).any(axis=1).all()
    g = load_golden("ref_plain_1024.npz")
    n = g["ti_pv"].shape[0]
    for j in range(n):
        ls = P.ti_at(g["vel"][j], g["periods"], g["ti_pv"][j].astype(np.float32), g["depz"],
                     float(g["minthk"]))
        assert np.array_equal(ls, g["ti_lsen"][j])
