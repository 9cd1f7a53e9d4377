"""The root search's secular function (csrc/dk_fdm.h, compiled for the host)
against the oracle's dltar4 (surfdisp96.f:767-865, Dunkin compound matrices).

The root search evaluates the Rayleigh determinant in each layer's potential
basis, which makes it dltar4's value times a positive factor.  These checks
pin that on random flattened models (gradients, low-velocity layers, slow
surface layers, thin and thick layers, periods 0.5-100 s, trial velocities
from 0.3 b_min to 1.1 b_max):
  * the sign agrees with dltar4's at every trial velocity (so every scan step
    of the reference's getsol takes the same decision);
  * the ratio to dltar4 is positive and finite, and smooth: the two
    functions' roots, bisected to 1e-12 km/s, agree to 1e-9 km/s;
  * a planted mistake (the interface's eta with the wrong sign) fails the
    sign check, so the comparison can fail.
Closed form: one half-space alone gives the Rayleigh function
(2 - c^2/b^2)^2 - 4 sqrt(1 - c^2/a^2) sqrt(1 - c^2/b^2) exactly.

The kernel pass and the VJP differentiate the same function through its
building blocks (forward tangent in c, adjoint sweep): at dltar4's own roots
their kernels -(dF/dm)/(dF/dc) in each layer's flattened a, b, rho and d
match central differences of dltar4 (relative step 1e-4, float32 model
values as the oracle takes them) to 2e-5 of the largest kernel, including
the half-space alone against the closed form's derivatives."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import port as P

LIB = os.path.join(ROOT, "dazi_current_release_amd", "libdazi_dk_hostcheck.so")


@pytest.fixture(scope="module")
def hc():
    if not os.path.exists(LIB):
        from dazi_current_release_amd import build as B
        B.build_hostcheck()
    lib = C.CDLL(LIB)
    f = lib.dk_host_secular
    f.restype = C.c_double
    fp = C.POINTER(C.c_float)
    f.argtypes = [C.c_double, C.c_double, C.c_int, fp, fp, fp, fp, C.POINTER(C.c_int)]

    def F(c, T, d, a, b, rho):
        ex = C.c_int(0)
        arr = [np.ascontiguousarray(x, np.float32) for x in (d, a, b, rho)]
        m = f(c, T, len(d), *[x.ctypes.data_as(fp) for x in arr], C.byref(ex))
        return np.ldexp(m, ex.value)
    return F


@pytest.fixture(scope="module")
def hk():
    lib = C.CDLL(LIB)
    f = lib.dk_host_kernels
    f.restype = C.c_double
    fp = C.POINTER(C.c_float)
    f.argtypes = [C.c_double, C.c_double, C.c_int, fp, fp, fp, fp, C.POINTER(C.c_double),
                  C.POINTER(C.c_int)]

    def K(c, T, d, a, b, rho):
        out = np.zeros(4 * len(d))
        ex = C.c_int(0)
        arr = [np.ascontiguousarray(x, np.float32) for x in (d, a, b, rho)]
        f(c, T, len(d), *[x.ctypes.data_as(fp) for x in arr],
          out.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ex))
        return out.reshape(len(d), 4)  # columns a, b, rho, d
    return K


def random_model(rng):
    n = int(rng.integers(2, 14))
    b = np.sort(rng.uniform(0.8, 4.7, n)).astype(np.float32)
    if rng.random() < 0.4:
        i = rng.integers(0, n - 1)
        b[i] *= rng.uniform(0.6, 0.95)
    if rng.random() < 0.2:
        b[0] = rng.uniform(0.3, 1.0)
    a = (b * rng.uniform(1.6, 2.2, n)).astype(np.float32)
    rho = (1.7 + 0.2 * a + rng.uniform(-0.3, 0.3, n)).astype(np.float32)
    d = rng.uniform(0.05, 20, n).astype(np.float32)
    d[-1] = 0
    T = float(np.exp(rng.uniform(np.log(0.5), np.log(100))))
    return d, a, b, rho, T


def bisect(f, lo, hi, flo):
    for _ in range(200):
        if hi - lo <= 1e-12:
            break
        mid = 0.5 * (lo + hi)
        fm = f(mid)
        if (fm < 0) == (flo < 0):
            lo, flo = mid, fm
        else:
            hi = mid
    return 0.5 * (lo + hi)


def test_halfspace_is_the_rayleigh_function(hc):
    a, b, rho = 6.0, 3.5, 2.8
    for c in (1.0, 2.5, 3.1, 3.2, 3.4):
        want = (2 - c * c / b ** 2) ** 2 - 4 * np.sqrt(1 - c * c / a ** 2) * np.sqrt(1 - c * c / b ** 2)
        got = hc(c, 10.0, [0.0], [a], [b], [rho])
        assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (c, got, want)


def test_signs_and_roots_match_dltar4(hc):
    rng = np.random.default_rng(20261204)
    nroots = 0
    for _ in range(120):
        d, a, b, rho, T = random_model(rng)
        cs = np.linspace(0.3 * float(b.min()), 1.1 * float(b.max()), 160)
        f1 = np.array([P.dltar4(c, T, d, a, b, rho) for c in cs])
        f2 = np.array([hc(c, T, d, a, b, rho) for c in cs])
        assert np.all(np.isfinite(f2))
        nz = f1 != 0
        assert np.array_equal(np.signbit(f1[nz]), np.signbit(f2[nz]))
        r = f2[nz] / f1[nz]
        assert np.all(r > 0)
        # roots of both in the first three sign changes
        chg = np.nonzero(np.signbit(f1[:-1]) != np.signbit(f1[1:]))[0][:3]
        for j in chg:
            r1 = bisect(lambda c: P.dltar4(c, T, d, a, b, rho), cs[j], cs[j + 1], f1[j])
            r2 = bisect(lambda c: hc(c, T, d, a, b, rho), cs[j], cs[j + 1], f2[j])
            assert abs(r1 - r2) <= 1e-9, (r1, r2)
            nroots += 1
    assert nroots > 100


def test_planted_sign_mistake_is_caught(hc):
    """The same comparison on a restatement with eta's sign flipped (numpy,
    the algebra of dk_fdm.h): its signs disagree with dltar4's."""
    def fdm(c, T, d, a, b, rho, s_eta):
        d, a, b, rho = (np.asarray(x, np.float64) for x in (d, a, b, rho))
        om = 2 * np.pi / T
        k, c2 = om / c, c * c
        mu = rho * b * b
        H = len(d) - 1
        ran, rbn = np.sqrt(abs(1 - c2 / a[H] ** 2)), np.sqrt(abs(1 - c2 / b[H] ** 2))
        e = [ran * rbn, 0.0, -1.0, -ran, rbn]

        def iface(m, e):
            eps, eta = rho[m - 1] / rho[m], s_eta * 2 * (mu[m - 1] - mu[m]) / (rho[m] * c2)
            A, B = eps - eta, 1 + eta
            y1, y2 = A * e[0] + (A - 1) * e[1], A * e[1] + (A - 1) * e[2]
            z1, z2 = eta * e[0] + B * e[1], eta * e[1] + B * e[2]
            return [A * y1 + (A - 1) * y2, eta * y1 + B * y2, eta * z1 + B * z2, eps * e[3], eps * e[4]]

        def hyp(X, kd):
            r = np.sqrt(abs(X))
            if X > 0:
                return np.cosh(r * kd), np.sinh(r * kd) / r, r * np.sinh(r * kd)
            return np.cos(r * kd), np.sin(r * kd) / r, -r * np.sin(r * kd)
        e = iface(H, e)
        for m in range(H - 1, -1, -1):
            Ca, Sa, Ra = hyp(1 - c2 / a[m] ** 2, k * d[m])
            Cb, Sb, Rb = hyp(1 - c2 / b[m] ** 2, k * d[m])
            t11, t12 = Ca * e[0] + Ra * e[4], Ca * e[3] + Ra * e[2]
            t21, t22 = Sa * e[0] + Ca * e[4], Sa * e[3] + Ca * e[2]
            e = [t11 * Cb - t12 * Rb, e[1], -t21 * Sb + t22 * Cb, -t11 * Sb + t12 * Cb,
                 t21 * Cb - t22 * Rb]
            if m > 0:
                e = iface(m, e)
        t = 2 - c2 / b[0] ** 2
        return -(4 * e[0] + 4 * t * e[1] + t * t * e[2])

    rng = np.random.default_rng(5)
    caught = 0
    for _ in range(20):
        d, a, b, rho, T = random_model(rng)
        # moderate phases: the unscaled restatement stays in range
        d = np.minimum(d, 3.0)
        T = max(T, 5.0)
        cs = np.linspace(0.5 * float(b.min()), 0.98 * float(b.max()), 60)
        f1 = np.array([P.dltar4(c, T, d, a, b, rho) for c in cs])
        good = np.array([fdm(c, T, d, a, b, rho, 1.0) for c in cs])
        bad = np.array([fdm(c, T, d, a, b, rho, -1.0) for c in cs])
        assert np.array_equal(np.signbit(f1), np.signbit(good))
        caught += not np.array_equal(np.signbit(f1), np.signbit(bad))
    assert caught >= 10


def test_kernels_match_dltar4_differences_at_its_roots(hk):
    rng = np.random.default_rng(3)
    F = P.dltar4
    checked = 0
    for _ in range(30):
        n = int(rng.integers(1, 9))
        b = np.sort(rng.uniform(1.0, 4.7, n)).astype(np.float32)
        if n > 2 and rng.random() < 0.4:
            b[1] *= 0.85
        a = (b * rng.uniform(1.65, 1.9, n)).astype(np.float32)
        rho = (1.7 + 0.2 * a).astype(np.float32)
        d = rng.uniform(0.5, 8, n).astype(np.float32)
        d[-1] = 0
        T = float(np.exp(rng.uniform(np.log(2), np.log(60))))
        cs = np.linspace(0.5 * float(b.min()), 0.999 * float(b.max()), 400)
        fv = np.array([F(c, T, d, a, b, rho) for c in cs])
        j = np.nonzero(np.signbit(fv[:-1]) != np.signbit(fv[1:]))[0]
        if len(j) == 0:
            continue
        c = bisect(lambda x: F(x, T, d, a, b, rho), cs[j[0]], cs[j[0] + 1], fv[j[0]])
        got = hk(c, T, d, a, b, rho)
        hc = 1e-6 * c
        dfc = (F(c + hc, T, d, a, b, rho) - F(c - hc, T, d, a, b, rho)) / (2 * hc)
        want = np.zeros((n, 4))
        model = [a, b, rho, d]
        for m in range(n):
            for p in range(4):
                if p == 3 and m == n - 1:
                    continue  # the half-space has no thickness
                x = float(model[p][m])
                up, dn = model[p].copy(), model[p].copy()
                up[m], dn[m] = np.float32(x * (1 + 1e-4)), np.float32(x * (1 - 1e-4))
                mp, mm = list(model), list(model)
                mp[p], mm[p] = up, dn
                dfm = (F(c, T, mp[3], mp[0], mp[1], mp[2]) - F(c, T, mm[3], mm[0], mm[1], mm[2])) \
                    / (float(up[m]) - float(dn[m]))
                want[m, p] = -dfm / dfc
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= 2e-5, (n, T, c, err)
        checked += 1
    assert checked >= 25


def test_root_search_elementary_functions_error_bounds():
    """The root search's exp(-p) and sin/cos of the layer phase (near-minimax
    polynomials on the reduced intervals) against long-double references:
    exp within 2e-12 relative on p in [0, 700], sin and cos within 2e-11
    absolute on p in [0, 2e4] (the phases a layer takes), and both exactly
    their Taylor limits at 0."""
    if not os.path.exists(LIB):
        from dazi_current_release_amd import build as B
        B.build_hostcheck()
    lib = C.CDLL(LIB)
    ex = lib.dk_host_exp_neg_rs
    ex.restype = C.c_double
    ex.argtypes = [C.c_double]
    sc = lib.dk_host_sincos_phase_rs
    sc.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    rng = np.random.default_rng(7)
    L = np.longdouble
    p = np.concatenate([np.linspace(0.0, 3.0, 3001), rng.uniform(0.0, 700.0, 20000)])
    got = np.array([ex(float(v)) for v in p])
    want = np.exp(-p.astype(L))
    assert float(np.max(np.abs((got - want) / want))) < 2e-12
    assert ex(0.0) == 1.0
    q = np.concatenate([np.linspace(0.0, 7.0, 7001), rng.uniform(0.0, 2e4, 20000)])
    s_, c_ = C.c_double(), C.c_double()
    gs, gc = [], []
    for v in q:
        sc(float(v), C.byref(s_), C.byref(c_))
        gs.append(s_.value)
        gc.append(c_.value)
    # long double sin/cos of a double argument: the reduction's own error is
    # part of what is bounded
    ql = q.astype(L)
    assert float(np.max(np.abs(np.array(gs) - np.sin(ql)))) < 2e-11
    assert float(np.max(np.abs(np.array(gc) - np.cos(ql)))) < 2e-11
    sc(0.0, C.byref(s_), C.byref(c_))
    assert s_.value == 0.0 and c_.value == 1.0
