"""CPU tests of the product library's host side (no GPU compute calls).

 - both flavours load and export every function declared in include/*.h;
 - the error texts equal the reference's own errmsg.c (compiled into
   oracle/_ref) and the oracle's;
 - argument validation codes of EZmock_init / set_cosmology /
   setup_linear_pk / populate_tracer (pre-GPU checks) match the reference
   functions (via the oracle, which restates them);
 - growth2, vfac and the trimmed P(k) spline table are bit-identical to the
   oracle's (same algorithms, same operation order);
 - EZmock_write_ascii writes the reference's header and row formats."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ezmock_amd import EZmock, _lib
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("EZmock.h", "EZmock_ext.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(EZmock_\w+)\s*\(", txt))
    return names


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_library_exports_every_declared_symbol(prec):
    L = C.CDLL(_lib.lib_path(prec))
    decl = declared_functions()
    assert len(decl) >= 8 + 18
    for name in decl:
        assert hasattr(L, name), name
    assert set(_lib.API_SYMBOLS + _lib.EXT_SYMBOLS) == decl
    assert _lib.load(prec).EZmock_ext_real_size() == (4 if prec == "f32" else 8)


def test_errmsg_matches_reference_and_oracle():
    L = _lib.load("f32")
    ref = os.path.join(ROOT, "oracle", "_ref", "libref_math.so")
    R = C.CDLL(ref) if os.path.exists(ref) else None
    if R is not None:
        R.EZmock_errmsg.restype = C.c_char_p
    for code in range(-3, 46):
        m = L.EZmock_errmsg(code)
        assert m == po.lib("f32").orc_errmsg(code)
        if R is not None:
            assert m == R.EZmock_errmsg(code)


@pytest.mark.parametrize("kw,code", [
    (dict(Lbox=0., Ngrid=64), 7), (dict(Lbox=100., Ngrid=0), 8),
    (dict(Lbox=100., Ngrid=70000), 8), (dict(Lbox=100., Ngrid=64, rng=2), 9),
    (dict(Lbox=100., Ngrid=64, seed=0), 10), (dict(Lbox=100., Ngrid=64, nthread=-1), 11),
])
def test_init_validation(kw, code):
    with pytest.raises(RuntimeError) as e:
        EZmock(**kw)
    assert str(e.value) == _lib.load("f32").EZmock_errmsg(code).decode()
    with pytest.raises(po.OracleError) as e2:
        po.EZmockOracle(**kw)
    assert e2.value.code == code


def test_required_keywords():
    with pytest.raises(KeyError):
        EZmock(Ngrid=64)
    with pytest.raises(KeyError):
        EZmock(Lbox=100.)


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("z,Om,w", [(0.5, 0.31, -1.0), (1.0, 0.25, -0.9), (0.0, 0.3, -1.1),
                                    (2.3, 0.5, -1.0)])
def test_growth_bitwise_equal_to_oracle(prec, z, Om, w):
    g = EZmock(Lbox=500., Ngrid=32, rng=0, seed=1, precision=prec)
    o = po.EZmockOracle(Lbox=500., Ngrid=32, rng=0, seed=1, prec=prec)
    g.eval_growth_params(z, Om, w=w)
    o.eval_growth_params(z, Om, w=w)
    assert g.info()["growth2"] == o.info()["growth2"]
    assert g.info()["vfac"] == o.info()["vfac"]


def test_cosmology_errors():
    for args, code in [((0.5, 0.0), 30), ((0.5, 1.5), 30), ((-0.1, 0.3), 30)]:
        g = EZmock(Lbox=500., Ngrid=32, rng=0)
        with pytest.raises(RuntimeError) as e:
            g.eval_growth_params(*args)
        assert str(e.value) == "invalid cosmological parameters"
    g = EZmock(Lbox=500., Ngrid=32, rng=0)
    with pytest.raises(RuntimeError):
        g.eval_growth_params(0.5, 0.3, w=-0.2)
    g = EZmock(Lbox=500., Ngrid=32, rng=0)
    with pytest.raises(RuntimeError):
        g.set_growth_params(0., 50.)


@pytest.mark.parametrize("logint,bao", [(False, 0.0), (True, 0.0), (False, -2.0), (True, 3.0)])
def test_pk_table_bitwise_equal_to_oracle(pk_table, logint, bao):
    k, P = pk_table
    Pnw = P * (1 + 0.05 * np.sin(k * 100))
    g = EZmock(Lbox=1000., Ngrid=64, rng=0, precision="f64")
    o = po.EZmockOracle(Lbox=1000., Ngrid=64, rng=0, prec="f64")
    for e in (g, o):
        e.eval_growth_params(0.5, 0.31)
        e.setup_linear_pk(k, P, Pnw=Pnw, BAO_enhance=bao, interp_log=logint)
    kg, Pg, P2g = g.pk_table()
    n = o.L.orc_pk_nbin(o._ez)
    ko, Po, P2o = np.zeros(n), np.zeros(n), np.zeros(n)
    o.L.orc_pk_table(o._ez, ko.ctypes.data, Po.ctypes.data, P2o.ctypes.data)
    assert np.array_equal(kg, ko) and np.array_equal(Pg, Po) and np.array_equal(P2g, P2o)


def test_pk_setup_errors_match_oracle(pk_table):
    k, P = pk_table
    bad_inputs = [
        (k, np.where(np.arange(len(P)) == 5, -1.0, P), {}),
        (k[::-1].copy(), P, {}),
        (k[:100], P[:100], {}),
        (k, np.where(np.arange(len(P)) == 7, np.nan, P), {}),
        (k[:3], P[:3], {}),
        (k, P, dict(Pnw=-P, BAO_enhance=1.0)),
        (k, P, dict(Pnw=P * 1e30, BAO_enhance=50.0)),
    ]
    for kk, pp, kw in bad_inputs:
        g = EZmock(Lbox=1000., Ngrid=64, rng=0)
        o = po.EZmockOracle(Lbox=1000., Ngrid=64, rng=0)
        g.set_growth_params(1.0, 50.)
        o.set_growth_params(1.0, 50.)
        with pytest.raises(RuntimeError) as e1:
            g.setup_linear_pk(kk, pp, **kw)
        with pytest.raises(po.OracleError) as e2:
            o.setup_linear_pk(kk, pp, **kw)
        assert str(e1.value) == str(e2.value)


def test_populate_validation_before_gpu(pk_table):
    g = EZmock(Lbox=500., Ngrid=32, rng=0)
    with pytest.raises(RuntimeError) as e:
        g.populate_tracer(0.5, 10., 0.3, 100., 1000)
    assert str(e.value) == "EZmock density fields uninitialized"
    g = EZmock(Lbox=500., Ngrid=32, rng=0)
    with pytest.raises(RuntimeError) as e:
        g.populate_tracer(0.5, 10., 0.3, 100., 100)
    assert str(e.value) == "too few expected number of tracers"
    g = EZmock(Lbox=500., Ngrid=32, rng=0)
    with pytest.raises(RuntimeError) as e:
        g.populate_tracer(0.5, 10., 0.3, 100., 32 ** 3)
    assert str(e.value) == "too many expected number of tracers"
    g = EZmock(Lbox=500., Ngrid=32, rng=0)
    with pytest.raises(RuntimeError) as e:
        g.create_dens_field_from_ic()
    assert str(e.value) == "EZmock power spectrum uninitialized"


def test_pool_admission_checks_before_gpu():
    """The request pool's checks that run before any device work: slot count,
    seed 0 and requests over the memory budget are refused at submit; options
    are accepted only while nothing has been admitted (every instance is
    created with the same options)."""
    from ezmock_amd.pool import MockPool
    L = _lib.load("f32")
    k = np.logspace(-3, 1, 64)
    with pytest.raises(RuntimeError) as e:
        MockPool(k, k ** -1.5, 0.5, 0.31, nslot=0)
    assert str(e.value) == L.EZmock_errmsg(11).decode()          # ARG_THREAD
    assert L.EZmock_ext_pool_option(None, 8, 1) == 12             # ARG_NULL
    pool = MockPool(k, k ** -1.5, 0.5, 0.31, nslot=2, mem_budget=1e9, options={8: 1, 9: 0})
    try:
        with pytest.raises(RuntimeError) as e:
            pool.submit(64, 500., 0, (0.5, 10., 0.3, 150.), 5000)
        assert str(e.value) == L.EZmock_errmsg(10).decode()      # ARG_SEED
        assert pool.footprint(2048) > 1e9
        with pytest.raises(RuntimeError) as e:
            pool.submit(2048, 2500., 1, (0.5, 10., 0.3, 150.), 5000)
        assert str(e.value) == L.EZmock_errmsg(1).decode()       # MEMORY
        # refused requests admit nothing: options are still open
        assert L.EZmock_ext_pool_option(pool._p, 7, 1) == 0
    finally:
        pool.close()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("rsd", [0.0, 0.011384])
def test_write_ascii_formats(tmp_path, prec, rsd):
    rt = np.float32 if prec == "f32" else np.float64
    L = _lib.load(prec)
    rs = np.random.default_rng(0)
    n = 1000
    Lbox = 500.
    x, y, z = [rs.uniform(1, Lbox, n).astype(rt) for _ in range(3)]
    z[0] = Lbox - 0.001
    v = [rs.normal(0, 300, n).astype(rt) for _ in range(3)]
    v[2][0] = 500.
    ez = L.EZmock_init(Lbox, 32, 0, 42, 0, C.byref(C.c_int(0)))
    err = C.c_int(0)
    fname = str(tmp_path / "cat.dat")
    L.EZmock_write_ascii(ez, x.ctypes.data, y.ctypes.data, z.ctypes.data, v[0].ctypes.data,
                         v[1].ctypes.data, v[2].ctypes.data, n, rsd, True,
                         fname.encode(), C.byref(err))
    assert err.value == 0
    lines = open(fname).read().splitlines()
    assert lines[0].startswith("# BOX_SIZE=500 , NUM_GRID=32 , FIX_AMPLITUDE=F , INVERT_PHASE=F")
    assert lines[1] == "# GROWTH_PK=inf , VELOCITY_FAC=inf"
    assert lines[3] == "# RAND_GENERATOR=MRG32K3A , RAND_SEED=42"
    fmt = "%.6g" if prec == "f32" else "%.10g"
    body = [l for l in lines if not l.startswith("#")]
    assert len(body) == n
    o = po.EZmockOracle(Lbox=Lbox, Ngrid=32, rng=0, seed=42, prec=prec)
    for i in (0, 1, 500, n - 1):
        if rsd > 0:
            zr = o.rsd_z([z[i]], [v[2][i]], rsd)[0]
            assert body[i] == " ".join(fmt % a for a in (x[i], y[i], zr))
        else:
            assert body[i] == " ".join(fmt % a for a in (x[i], y[i], z[i], v[0][i], v[1][i], v[2][i]))
    # rsd_fac > Lbox is rejected; a first coordinate equal to 0 is rejected
    err = C.c_int(0)
    L.EZmock_write_ascii(ez, x.ctypes.data, y.ctypes.data, z.ctypes.data, v[0].ctypes.data,
                         v[1].ctypes.data, v[2].ctypes.data, n, Lbox + 1, True,
                         fname.encode(), C.byref(err))
    assert err.value == 19
    x[0] = 0
    err = C.c_int(0)
    L.EZmock_write_ascii(ez, x.ctypes.data, y.ctypes.data, z.ctypes.data, v[0].ctypes.data,
                         v[1].ctypes.data, v[2].ctypes.data, n, rsd, True,
                         fname.encode(), C.byref(err))
    assert err.value == 12
    L.EZmock_destroy(ez)
