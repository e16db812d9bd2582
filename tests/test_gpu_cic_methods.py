"""CIC painting by particle binning into fixed list ranges (default), by
binning with exact ranges (EZmock_ext_set_option 3 = 1) and by tile scan
(option 3 = 2): all against the oracle, and against each other, on generated fields,
on random displacement fields of large range (forced scan, and the ranges
the scan declines), and on non-finite input (the scan declines; binning
paints).  Tolerances: FP64 density within 1e-10 relative of the oracle;
FP32 density rms within 1e-5; between the methods, FP64 densities within
1e-13 relative and FP32 densities identical in > 99.9% of the cells;
>= 99.99% identical tracers (positions bit-equal, velocities to rounding,
matched by host cell and rank within the cell) between the methods, and
against the oracle in FP64; in FP32, every tracer before the first cell
whose count differs from the oracle's is identical.""" 
import numpy as np
import pytest

import pincheck as pc
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
        tol = 1e-9 if prec == "f64" else 2e-5
        assert pc.identical_tracer_fraction(nt, cg, ntb, cb, tol) >= 0.9999
    # against the oracle catalogue: FP64 tracer for tracer; FP32 up to the
    # first cell whose count differs (the oracle's float sums round some
    # densities differently, which can move a cell across a PDF bin edge)
    nto = o.field("nt")
    if prec == "f64":
        assert pc.identical_tracer_fraction(res[0][1], res[0][2], nto, co, 1e-9) >= 0.9999
    else:
        n = pc.identical_before_first_difference(res[0][1], res[0][2], nto, co, 2e-5)
        print(f"FP32 vs oracle: tracers per cell identical in {np.mean(res[0][1] == nto):.6f} "
              f"of the cells; {n} of {len(co[0])} tracers before the first differing cell")
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
    assert pc.identical_tracer_fraction(out[2][0], out[2][1], out[1][0], out[1][1],
                                        1e-9) >= 0.9999
    assert pc.identical_tracer_fraction(out[1][0], out[1][1], o.field("nt"), co,
                                        1e-9) >= 0.9999
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


def tile_list_counts(disp, Ng, L, T=(8, 8, 16)):
    """Entries of each tile's particle list in the binning: a particle is
    listed in the tile of its (wrapped) base cell and, along each axis where
    that cell is the last of its tile, in the next tile too."""
    q = np.arange(Ng)
    lat = np.meshgrid(q, q, q, indexing="ij")
    base = [np.mod(np.floor(d.reshape(Ng, Ng, Ng) * (Ng / L) + a), Ng).astype(np.int64).ravel()
            for d, a in zip(disp, lat)]
    nt = [Ng // t for t in T]
    cnt = np.zeros(nt, np.int64)
    spans = []
    for b, t, n in zip(base, T, nt):
        lo = b // t
        hi = np.where((b % t == t - 1) & (n > 1), (lo + 1) % n, -1)
        spans.append((lo, hi))
    for sx in (0, 1):
        for sy in (0, 1):
            for sz in (0, 1):
                ix = spans[0][sx]
                iy = spans[1][sy]
                iz = spans[2][sz]
                ok = (ix >= 0) & (iy >= 0) & (iz >= 0)
                np.add.at(cnt, (ix[ok], iy[ok], iz[ok]), 1)
    return cnt


def compressed(Ng, a, b):
    """Lattice compressed by a along x and by b along y (cell size 1)."""
    q = np.arange(Ng, dtype=np.float64)
    i, j, _ = np.meshgrid(q, q, q, indexing="ij")
    return [(i / a - i).ravel(), (j / b - j).ravel(), np.zeros(Ng ** 3)]


@pytest.mark.parametrize("limit", [False, True])
def test_fixed_list_resize_and_second_overflow(gpu_available, pk_table, limit):
    """One instance, four fields (32^3, FP64, 8x8x16 tiles of 1024 cells):
      A  fullest tile 4.03 entries per cell: overflows the default ranges
         (4 per cell), is repainted with exact ranges (count pass), and the
         ranges of later fields become 1.25 x 4131 -> 5184 entries;
      B  fullest tile 4.18 per cell: fits the resized ranges (no count pass,
         although it would have overflowed the default ones);
      C  fullest tile 5.84 per cell: overflows the resized ranges; 1.25 x
         that is 7.3 per cell, above 6: with the device's free memory the
         ranges grow to it (up to 16 per cell), while with option 5 limiting
         the lists to 24 bytes per cell the instance keeps exact ranges;
      D  the field of B again: fixed ranges, or exact ranges (count pass)
         under the limit.
    Every density matches the oracle's, and the catalogues of B and D
    match the oracle's tracer for tracer."""
    Ng, L = 32, 32.
    fields = {"A": compressed(Ng, 3.27, 1.0), "B": compressed(Ng, 3.4, 1.0),
              "C": compressed(Ng, 4.0, 1.2)}
    fields["D"] = fields["B"]
    mx = {n: tile_list_counts(d, Ng, L).max() for n, d in fields.items()}
    assert 4096 < mx["A"] <= 4915                    # overflows; resize <= 6 / cell
    cap = (mx["A"] + mx["A"] // 4 + 63) // 64 * 64
    assert 4096 < mx["B"] <= cap < mx["C"]
    assert 6 * 1024 < mx["C"] + mx["C"] // 4 <= 16 * 1024
    g = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=3, precision="f64")
    if limit:
        assert g.set_option(5, 24 * Ng ** 3) == 0
    g.profile(True)
    g.eval_growth_params(0.5, 0.31)
    count_pass = {}
    for name in "ABCD":
        g.reset_kernel_times()
        g.create_dens_field_from_disp(*fields[name], deepcopy=True)
        count_pass[name] = "cic_bin_count" in g.kernel_times()
        o = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=3, nthread=4, prec="f64")
        o.eval_growth_params(0.5, 0.31)
        o.create_dens_field_from_disp(*fields[name])
        ro = o.field("rho")
        assert np.abs(g.field("rho") - ro).max() <= 1e-10 * np.abs(ro).max(), name
        if name in "BD":
            cg = g.populate_tracer(0.5, 10., 0.3, 150., 2000)
            co = o.populate_tracer(0.5, 10., 0.3, 150., 2000)
            assert pc.identical_tracer_fraction(g.field("nt"), cg, o.field("nt"), co,
                                                1e-9) >= 0.9999, name
        o.close()
    assert count_pass == {"A": True, "B": False, "C": True, "D": limit}


def test_fixed_list_memory_limit_takes_exact_ranges(gpu_available, pk_table):
    """Option 5 limits the bytes of the fixed list ranges: below what they
    need, the painting uses exact ranges (count pass) and gives the same
    density as the default."""
    k, P = pk_table
    rho = {}
    for limit in (0, 4096):
        g = EZmock(Lbox=256., Ngrid=32, rng=0, seed=3, precision="f64")
        assert g.set_option(5, limit) == 0
        g.profile(True)
        g.eval_growth_params(0.5, 0.31)
        g.setup_linear_pk(k, P)
        g.create_dens_field_from_ic()
        assert ("cic_bin_count" in g.kernel_times()) == (limit > 0)
        rho[limit] = g.field("rho")
    assert np.abs(rho[0] - rho[4096]).max() <= 1e-13 * np.abs(rho[0]).max()
