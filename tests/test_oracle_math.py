"""Pins the oracle's set-up math: Gauss 2F1, natural cubic spline, growth
factor/rate and the linear P(k) table (oracle/orc_math.c, orc_field.c).

The reference's own math/hypergeom.c and math/cspline.c compile standalone;
oracle/Makefile builds them into oracle/_ref/libref_math.so from the
reference tree, and the oracle must agree with them bit for bit.  scipy and
closed forms (Einstein-de Sitter D = a, f = 1) pin the mathematics."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.special as sp
from scipy.interpolate import CubicSpline

from oracle import pyoracle as po

REF = os.path.join(os.path.dirname(po.__file__), "_ref", "libref_math.so")


def ref_lib():
    if not os.path.exists(REF):
        pytest.skip("oracle/_ref/libref_math.so not built (reference tree absent)")
    L = C.CDLL(REF)
    L.hyp2f1.argtypes = [C.c_double] * 5 + [C.c_int, C.POINTER(C.c_double)]
    L.cspline_ypp.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    L.cspline_eval.restype = C.c_double
    L.cspline_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_size_t]
    L.EZmock_errmsg.restype = C.c_char_p
    return L


def orc_hyp(a, b, c, z):
    r = C.c_double()
    e = po.lib("f64").orc_hyp2f1(a, b, c, z, 1e-12, 10000, C.byref(r))
    return e, r.value


CASES = []
for w in (-1.0, -0.9, -1.2):
    for Om in (0.31, 0.25, 0.5):
        for a in (1.0, 0.8, 2 / 3, 0.5, 0.25):
            iw = 1 / w
            c1, c2, c3 = 0.5 * (1 - iw), -iw / 3, 1 - 5 * iw / 6
            x = (1 - 1 / Om) * a ** (-3 * w)
            CASES.append((c1, c2, c3, x))
            CASES.append((c1 + 1, c2 + 1, c3 + 1, x))


@pytest.mark.parametrize("a,b,c,z", CASES[:24])
def test_hyp2f1_matches_reference_source(a, b, c, z):
    L = ref_lib()
    r = C.c_double()
    e1 = L.hyp2f1(a, b, c, z, 1e-12, 10000, C.byref(r))
    e2, v = orc_hyp(a, b, c, z)
    assert e1 == e2
    if e1 == 0:
        assert v == r.value


@pytest.mark.parametrize("a,b,c,z", CASES)
def test_hyp2f1_matches_scipy(a, b, c, z):
    e, v = orc_hyp(a, b, c, z)
    assert e == 0
    assert v == pytest.approx(sp.hyp2f1(a, b, c, z), rel=1e-10)


def spline_data(n=60, seed=0):
    rs = np.random.default_rng(seed)
    x = np.sort(rs.uniform(-3, 5, n))
    y = np.sin(x) + 0.1 * x ** 2
    return np.ascontiguousarray(x), np.ascontiguousarray(y)


def test_cspline_matches_reference_source():
    L = ref_lib()
    x, y = spline_data()
    a = np.zeros(2 * len(x))
    b = np.zeros(2 * len(x))
    L.cspline_ypp(x.ctypes.data, y.ctypes.data, len(x), a.ctypes.data)
    po.lib("f64").orc_cspline_ypp(x.ctypes.data, y.ctypes.data, len(x), b.ctypes.data)
    assert np.array_equal(a[:len(x)], b[:len(x)])
    for xv in np.linspace(x[0], x[-1] - 1e-9, 97):
        i = np.searchsorted(x, xv, side="right") - 1
        r1 = L.cspline_eval(x.ctypes.data, y.ctypes.data, a.ctypes.data, xv, int(i))
        r2 = po.lib("f64").orc_cspline_eval(x.ctypes.data, y.ctypes.data, b.ctypes.data, xv, int(i))
        assert r1 == r2


def test_cspline_matches_scipy_natural():
    x, y = spline_data(seed=3)
    ypp = np.zeros(2 * len(x))
    L = po.lib("f64")
    L.orc_cspline_ypp(x.ctypes.data, y.ctypes.data, len(x), ypp.ctypes.data)
    cs = CubicSpline(x, y, bc_type="natural")
    assert np.allclose(ypp[:len(x)], cs(x, 2), rtol=1e-8, atol=1e-10)
    xs = np.linspace(x[0], x[-1] - 1e-9, 200)
    for xv in xs:
        i = int(np.searchsorted(x, xv, side="right") - 1)
        v = L.orc_cspline_eval(x.ctypes.data, y.ctypes.data, ypp.ctypes.data, xv, i)
        assert v == pytest.approx(float(cs(xv)), rel=1e-10, abs=1e-12)


def test_cspline_reproduces_linear_data():
    x = np.linspace(0, 1, 11)
    y = 3 * x - 2
    ypp = np.zeros(22)
    L = po.lib("f64")
    L.orc_cspline_ypp(x.ctypes.data, y.ctypes.data, 11, ypp.ctypes.data)
    assert np.allclose(ypp[:11], 0, atol=1e-12)


def test_growth_einstein_de_sitter():
    ez = po.EZmockOracle(Lbox=100., Ngrid=8, rng=0, seed=1, prec="f64")
    for z in (0.0, 0.5, 2.0):
        # Omega_m = 1 (w irrelevant): D = a / a_pk, f = 1, H = 100 a^-1.5
        ez.eval_growth_params(z, 1.0)
        info = ez.info()
        a = 1 / (1 + z)
        assert info["growth2"] == pytest.approx(a * a, rel=1e-12)
        assert info["vfac"] == pytest.approx(100 * a ** -1.5 * a, rel=1e-12)


def test_growth_lcdm_against_integral():
    """D(a) of flat LCDM from the growth integral D ~ H(a) int da/(aH)^3."""
    from scipy.integrate import quad
    Om = 0.31
    ez = po.EZmockOracle(Lbox=100., Ngrid=8, rng=0, seed=1, prec="f64")
    ez.eval_growth_params(0.5, Om)
    g2 = ez.info()["growth2"]

    def D(a):
        E = lambda x: np.sqrt(Om * x ** -3 + 1 - Om)
        return E(a) * quad(lambda x: 1 / (x * E(x)) ** 3, 0, a, epsabs=0, epsrel=1e-12)[0]
    assert g2 == pytest.approx((D(1 / 1.5) / D(1.0)) ** 2, rel=1e-7)
    assert abs(g2 - 0.593395) < 1e-6
    assert abs(ez.info()["vfac"] - 66.5111) < 1e-4


def test_linear_pk_spline_and_trim(pk_table):
    k, P = pk_table
    ez = po.EZmockOracle(Lbox=1000., Ngrid=64, rng=0, seed=1, prec="f64")
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    g2 = ez.info()["growth2"]
    cs = CubicSpline(k, P * g2, bc_type="natural")
    kf = 2 * np.pi / 1000.
    for q in (1, 2, 3, 10, 100, 1000, 3 * 32 ** 2):
        kk = kf * np.sqrt(q)
        # the trimmed spline differs from the full one only by the boundary
        # conditions three bins away: agreement to 1e-4 away from the edges
        assert ez.pk_interp(kk) == pytest.approx(float(cs(kk)), rel=2e-3)
    # clamping outside the table
    assert ez.pk_interp(1e-9) == ez.pk_interp(1e-8)


def test_linear_pk_errors(pk_table):
    """Each error leaves the instance's error code set (the reference's
    sticky `int *err`, which its Python class never clears), so every case
    uses a fresh instance."""
    k, P = pk_table

    def fresh(set_cosmo=True):
        ez = po.EZmockOracle(Lbox=1000., Ngrid=64, rng=0, seed=1, prec="f64")
        if set_cosmo:
            ez.set_growth_params(1.0, 50.0)
        return ez

    cases = [
        (dict(set_cosmo=False), (k, P), 4),        # ARG_COSMO
        ({}, (k, np.where(np.arange(len(P)) == 5, -1.0, P)), 22),   # PK_NONPOS
        ({}, (k[::-1].copy(), P), 23),            # PK_NONASC
        ({}, (k[:100], P[:100]), 24),             # PK_RANGE
        ({}, (k, np.where(np.arange(len(P)) == 7, np.inf, P)), 21),  # PK_FINITE
        ({}, (k[:3], P[:3]), 13),                  # ARG_NBIN
    ]
    for kw, (kk, pp), code in cases:
        ez = fresh(**kw)
        with pytest.raises(po.OracleError) as e:
            ez.setup_linear_pk(kk, pp)
        assert e.value.code == code
        # sticky: the next call fails with the same code
        with pytest.raises(po.OracleError) as e2:
            ez.set_growth_params(1.0, 50.0)
        assert e2.value.code == code


def test_errmsg_matches_reference_source():
    L = ref_lib()
    Lo = po.lib("f32")
    for code in range(-2, 45):
        assert L.EZmock_errmsg(code) == Lo.orc_errmsg(code)
