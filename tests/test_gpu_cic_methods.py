"""CIC painting by particle binning into fixed list ranges (default), by
binning with exact ranges (EZmock_ext_set_option 3 = 1) and by tile scan
(option 3 = 2): all against the oracle, and against each other, on generated fields,
on random displacement fields of large range (forced scan, and the ranges
the scan declines), and on non-finite input (the scan declines; binning
paints).  Tolerances: FP64 density within 1e-10 relative of the oracle;
FP32 density rms within 1e-5; between the methods, FP64 densities within
1e-13 relative and FP32 densities identical in > 99.9% of the cells; tracer
positions bit-exact whenever the tracer cells agree.""" 
import numpy as np
import pytest

from ezmock_amd import EZmock
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def run_method(method, Ng, L, z, prec, pk, disp=None, params=(0.5, 10., 0.3, 150.), nexp=None,
               seed=3):
    k, P = pk
    g = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=seed, precision=prec)
    g.cic_method(method)
    g.profile(True)
    g.eval_growth_params(z, 0.31)
    if disp is None:
        g.setup_linear_pk(k, P)
        g.create_dens_field_from_ic()
    else:
        g.create_dens_field_from_disp(*disp)
    names = set(g.kernel_times())
    cat = g.populate_tracer(*params, nexp or max(200, int(0.02 * Ng ** 3)))
    return g, names, cat


def oracle(Ng, L, z, prec, pk, disp=None, params=(0.5, 10., 0.3, 150.), nexp=None, seed=3):
    k, P = pk
    o = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=seed, nthread=4, prec=prec)
    o.eval_growth_params(z, 0.31)
    if disp is None:
        o.setup_linear_pk(k, P)
        o.create_dens_field_from_ic()
    else:
        o.create_dens_field_from_disp(*disp)
    cat = o.populate_tracer(*params, nexp or max(200, int(0.02 * Ng ** 3)))
    return o, cat


def used_scan(names):
    return any(n.startswith("cic_scan") for n in names) and not any(
        n.startswith("cic_tiles") for n in names)


@pytest.mark.parametrize("Ng,L,z,prec", [
    (32, 256., 0.5, "f64"), (48, 96., 0.0, "f64"), (64, 128., 0.0, "f32"), (24, 48., 0.0, "f32"),
    (12, 60., 0.5, "f64"),
])
def test_methods_against_oracle(gpu_available, pk_table, Ng, L, z, prec):
    o, co = oracle(Ng, L, z, prec, pk_table)
    ro = o.field("rho")
    res = {}
    for method in (0, 1, 2):
        g, names, cg = run_method(method, Ng, L, z, prec, pk_table)
        assert used_scan(names) == (method == 2)
        rg = g.field("rho")
        if prec == "f64":
            assert np.abs(rg - ro).max() <= 1e-10 * np.abs(ro).max()
        else:
            d = rg.astype(np.float64) - ro
            assert np.sqrt(np.mean(d ** 2) / np.mean(ro.astype(np.float64) ** 2)) < 1e-5
        res[method] = (rg, g.field("nt"), cg)
    for method in (1, 2):
        rg, nt, cg = res[method]
        rb, ntb, cb = res[0]
        # the same sums in another order: FP64 to rounding, FP32 (double
        # sums rounded once) almost always identical
        if prec == "f64":
            assert np.abs(rg - rb).max() <= 1e-13 * np.abs(rb).max()
        else:
            assert np.mean(rg == rb) > 0.999
        if np.array_equal(nt, ntb):
            for a, b in zip(cg[:3], cb[:3]):
                assert np.array_equal(a, b)
            tol = 1e-9 if prec == "f64" else 2e-5
            for a, b in zip(cg[3:], cb[3:]):
                assert np.abs(a.astype(float) - b).max() <= tol * max(1.0, np.abs(b).max())
    # against the oracle catalogue when the tracer cells agree
    if np.array_equal(res[0][1], o.field("nt")):
        for a, b in zip(res[0][2][:3], co[:3]):
            assert np.array_equal(a, b)
    o.close()


@pytest.mark.parametrize("Ng,amp,expect_scan", [(32, 5.0, True), (48, 10.0, True), (32, 14.5, False)])
def test_random_displacements(gpu_available, pk_table, Ng, amp, expect_scan):
    """Uniform random displacements of +-amp cells (no spatial coherence:
    the widest candidate boxes), FP64, forced scan against binning and the
    oracle; ranges near half the box are declined by the scan."""
    L = float(Ng)                                   # cell size 1
    rs = np.random.default_rng(int(amp * 10) + Ng)
    disp = [rs.uniform(-amp, amp, Ng ** 3) for _ in range(3)]
    o, co = oracle(Ng, L, 0.5, "f64", pk_table, disp=disp, nexp=200)
    ro = o.field("rho")
    out = {}
    for method in (2, 1):
        g, names, cg = run_method(method, Ng, L, 0.5, "f64", pk_table, disp=disp, nexp=200)
        assert used_scan(names) == (expect_scan and method == 2)
        rg = g.field("rho")
        assert np.abs(rg - ro).max() <= 1e-10 * np.abs(ro).max()
        out[method] = (g.field("nt"), cg)
    if np.array_equal(out[2][0], out[1][0]):
        for a, b in zip(out[2][1][:3], out[1][1][:3]):
            assert np.array_equal(a, b)
    o.close()


def test_non_finite_displacement_takes_binning(gpu_available, pk_table):
    Ng = 16
    rs = np.random.default_rng(11)
    disp = [rs.uniform(-1, 1, Ng ** 3) for _ in range(3)]
    disp[1][123] = np.inf
    rho = {}
    for method in (2, 1):
        k, P = pk_table
        g = EZmock(Lbox=float(Ng), Ngrid=Ng, rng=0, seed=1, precision="f64")
        g.cic_method(method)
        g.profile(True)
        g.eval_growth_params(0.5, 0.31)
        g.create_dens_field_from_disp(*disp)
        names = set(g.kernel_times())
        assert any(n.startswith("cic_tiles") for n in names)
        rho[method] = g.field("rho")
    assert np.allclose(rho[1], rho[2], rtol=1e-13, atol=0, equal_nan=True)


def test_cic_method_option_is_validated(gpu_available):
    g = EZmock(Lbox=100., Ngrid=16, rng=0, seed=1, precision="f32")
    with pytest.raises(ValueError):
        g.cic_method(4)



def test_fixed_lists_overflow_falls_back(gpu_available, pk_table):
    """Displacements that pull every particle into a few cells overflow the
    fixed list ranges of their tiles: the painting is redone with exact
    ranges (a count pass appears) and matches the oracle."""
    Ng = 32
    L = float(Ng)
    q = np.arange(Ng, dtype=np.float64)
    i, j, k = np.meshgrid(q, q, q, indexing="ij")
    disp = [(-0.9 * (a - Ng / 2)).ravel() for a in (i, j, k)]
    o = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=3, nthread=4, prec="f64")
    o.eval_growth_params(0.5, 0.31)
    o.create_dens_field_from_disp(*disp)
    g = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=3, precision="f64")
    g.profile(True)
    g.eval_growth_params(0.5, 0.31)
    g.create_dens_field_from_disp(*disp)
    names = set(g.kernel_times())
    assert "cic_bin_count" in names
    ro = o.field("rho")
    assert np.abs(g.field("rho") - ro).max() <= 1e-10 * np.abs(ro).max()
    o.close()
