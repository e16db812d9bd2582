"""GPU parity tests: the HIP engine (through its C ABI) against the CPU
oracle on the same seeded inputs, and against numpy / closed forms.

Tolerances (stated per check): integer, index and stream work is bit-exact;
spectra are bit-exact in FP32 and within a few ulp in FP64 (device vs glibc
transcendentals); FFT outputs within FP rounding; the FP64 density field
within 1e-10 relative, >= 99.99% identical tracer cells and >= 99.99%
identical tracers; FP32 tracers identical up to the first cell whose count
differs from the oracle's."""
import os

import numpy as np
import pytest

import pincheck as pc
from ezmock_amd import EZmock, uniform, fft_c2r, fft_r2c, pdf_walk
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz"))


def test_uniform_draws_bit_exact(gpu_available):
    for seed in (1, 42, 2 ** 40 + 3):
        for start in (0, 1, 997, 5 * 512 ** 3 + 12345, 8 * 1536 ** 3 + 6):
            g = uniform(seed, start, 3000)
            o = po.Rng(0, seed, start, prec="f64").double(3000)
            assert np.array_equal(g, o), (seed, start)


@pytest.mark.parametrize("seed", [1, 42, 5489, 2 ** 40 + 3])
def test_uniform_draws_mt19937_bit_exact(gpu_available, seed):
    """MT19937 streams on the GPU (jump-ahead by x^e mod phi, segment
    generation) against the oracle's sequential generator: get_double and
    get_double_pos (two words each), at offsets below and across segments
    of 2^18 words."""
    for start in (0, 1, 311, 312, 997, 131071, 131072, 300001):
        o = po.Rng(1, seed, start, prec="f64")
        ref = o.double(700)
        assert np.array_equal(uniform(seed, start, 700, rng=1), ref), (seed, start)
        o = po.Rng(1, seed, start, prec="f64")
        ref = o.double_pos(700)
        assert np.array_equal(uniform(seed, start, 700, rng=1, pos=True), ref), (seed, start)


@pytest.mark.parametrize("n", [2_200_000, 34_000_000])
def test_mt19937_long_stream_segments(gpu_available, n):
    """Streams long enough for two and three levels of segment windows
    (4.4 M words: 17 segments of 2^18 words; 68 M words: 260 segments):
    spot checks against the oracle at segment and level boundaries."""
    g = uniform(7, 0, n, rng=1)
    S = 131072                                      # doubles per segment
    for d in (0, S - 1, S, S + 1, 2 * S, 16 * S - 1, 16 * S, 16 * S + 5, 67 * S + 3,
              256 * S - 1, 256 * S, 259 * S + 7, n - 3):
        if d >= n:
            continue
        ref = po.Rng(1, 7, d, prec="f64").double(3)
        assert np.array_equal(g[d:d + 3], ref[:len(g[d:d + 3])]), d


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("Ng", [8, 12, 16, 24, 32, 48, 64, 96])
def test_fft_against_numpy(gpu_available, Ng, prec):
    rs = np.random.default_rng(Ng)
    H = rs.standard_normal((Ng, Ng, Ng // 2 + 1)) + 1j * rs.standard_normal((Ng, Ng, Ng // 2 + 1))
    ref = np.fft.irfftn(H, s=(Ng, Ng, Ng), axes=(0, 1, 2)) * Ng ** 3
    tol = 1e-13 if prec == "f64" else 3e-6
    assert np.abs(fft_c2r(H, prec) - ref).max() <= tol * np.abs(ref).max()
    x = rs.standard_normal((Ng, Ng, Ng))
    refk = np.fft.rfftn(x, axes=(0, 1, 2))
    assert np.abs(fft_r2c(x, prec) - refk).max() <= tol * np.abs(refk).max()


@pytest.mark.parametrize("Ng", [8, 12, 16])
def test_fft_golden_vectors(gpu_available, Ng):
    out = fft_c2r(GOLD[f"c2r_in_{Ng}"], "f64")
    ref = GOLD[f"c2r_out_{Ng}"]
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()
    out = fft_r2c(GOLD[f"r2c_in_{Ng}"], "f64")
    ref = GOLD[f"r2c_out_{Ng}"]
    assert np.abs(out - ref).max() <= 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_pdf_walk_with_planted_ties(gpu_available, prec, seed):
    rs = np.random.default_rng(seed)
    n = 300000
    rhot = rs.exponential(2.0, n)
    rhot[rs.random(n) < 0.4] = 0
    rhot = rhot.astype(po.rtype(prec))
    e, info = po.eval_pdf(25000, 0.3 + 0.05 * seed, n, prec)
    order = np.argsort(-rhot.astype(np.float64), kind="stable")
    for c in np.cumsum(info["pdf"])[:-1]:
        rhot[order[c:c + 2 + seed]] = rhot[order[c - 1]]
    e1, nt1, p1, nc1, ntot1 = po.pdf_walk(rhot, info["ncell"], info["pdf"], prec)
    e2, nt2, p2, nc2, ntot2 = pdf_walk(rhot, info["ncell"], info["pdf"], prec)
    assert e1 == 0 and e2 == 0
    assert np.array_equal(nt1, nt2)
    assert np.array_equal(p1, p2) and (nc1, ntot1) == (nc2, ntot2)
    assert nc1 > info["ncell"]


def run_pair(Ng, L, prec, seed=1, fixamp=False, iphase=False, logpk=False, params=None,
             nexp=None, pk=None, nthread=1, rng=0):
    k, P = pk
    g = EZmock(Lbox=L, Ngrid=Ng, rng=rng, seed=seed, precision=prec)
    o = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=rng, seed=seed, nthread=nthread, prec=prec,
                        keep_spec=True)
    g.keep_spectrum(True)
    for e in (g, o):
        e.eval_growth_params(0.5, 0.31)
        e.setup_linear_pk(k, P, interp_log=logpk)
        e.create_dens_field_from_ic(fixamp=fixamp, iphase=iphase)
    params = params or (0.5, 10., 0.3, 150.)
    nexp = nexp or int(0.02 * Ng ** 3)
    cg = g.populate_tracer(*params, nexp)
    co = o.populate_tracer(*params, nexp)
    return g, o, cg, co


def rel_rms(a, b):
    a, b = a.astype(np.float64), b.astype(np.float64)
    return np.sqrt(np.mean((a - b) ** 2) / np.mean(b ** 2))


@pytest.mark.parametrize("Ng,prec,fixamp,iphase,logpk,rng", [
    (32, "f64", False, False, False, 0), (32, "f32", False, False, False, 0),
    (48, "f64", True, False, False, 0), (48, "f32", False, True, False, 0),
    (64, "f64", False, False, True, 0), (64, "f32", True, True, True, 0),
    (32, "f64", False, False, False, 1), (48, "f32", False, True, False, 1),
    (64, "f64", True, False, True, 1),
])
def test_pipeline_parity_small(gpu_available, pk_table, Ng, prec, fixamp, iphase, logpk, rng):
    """Whole mock, GPU against the oracle, MRG32k3a (rng 0) and MT19937
    (rng 1) streams."""
    g, o, cg, co = run_pair(Ng, 8.0 * Ng, prec, fixamp=fixamp, iphase=iphase, logpk=logpk,
                            pk=pk_table, rng=rng, seed=1 if rng == 0 else 4357)
    sg, so = g.field("spec"), o.field("spec")
    if prec == "f32":
        # float spectra: the double->float rounding absorbs ulp-level device
        # vs glibc differences except on rare rounding ties
        assert np.mean(sg == so) > 0.9999
        assert np.abs(sg.astype(float) - so).max() <= 2e-7 * np.abs(so).max()
    else:
        assert np.abs(sg - so).max() <= 4e-16 * np.abs(so).max()
    for f in ("psi_x", "psi_y", "psi_z"):
        assert rel_rms(g.field(f), o.field(f)) < (1e-13 if prec == "f64" else 1e-6)
    rg, ro = g.field("rho"), o.field("rho")
    if prec == "f64":
        assert np.abs(rg - ro).max() <= 1e-10 * np.abs(ro).max()
    else:
        assert rel_rms(rg, ro) < 1e-5
    pg, pO = g.pdf_info(), o.pdf_info()
    assert pg["nmax"] == pO["nmax"]
    ntg, nto = g.field("nt"), o.field("nt")
    assert np.mean(ntg == nto) >= 0.9999
    assert abs(pg["ntot"] - pO["ntot"]) <= 1e-4 * pO["ntot"]
    # identical tracers (positions bit-equal, velocities within vt of the
    # largest speed), matched by host cell and rank within the cell; in FP32
    # up to the first cell whose count differs (float sums of the oracle)
    vt = 1e-9 if prec == "f64" else 2e-5
    frac = pc.identical_tracer_fraction(ntg, cg, nto, co, vt)
    print(f"identical tracers {frac:.6f} of {len(co[0])}")
    if prec == "f64":
        assert frac >= 0.9999
    else:
        pc.identical_before_first_difference(ntg, cg, nto, co, vt)


def test_white_noise_and_displacement_inputs(gpu_available, pk_table):
    k, P = pk_table
    Ng, L = 32, 256.
    rs = np.random.default_rng(7)
    delta = rs.standard_normal(Ng ** 3)
    g = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=3, precision="f64")
    o = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=3, nthread=1, prec="f64")
    for e in (g, o):
        e.eval_growth_params(0.5, 0.31)
        e.setup_linear_pk(k, P)
        e.create_dens_field_from_wn(delta)
    for f in ("psi_x", "psi_y", "psi_z", "rho"):
        assert rel_rms(g.field(f), o.field(f)) < 1e-12, f
    # displacement input, deep copy and by reference
    psi = [o.field(f) for f in ("psi_x", "psi_y", "psi_z")]
    for deep in (True, False):
        g2 = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=3, precision="f64")
        g2.eval_growth_params(0.5, 0.31)
        g2.create_dens_field_from_disp(*psi, deepcopy=deep)
        assert np.abs(g2.field("rho") - o.field("rho")).max() < 1e-10
        cg = g2.populate_tracer(0.4, 8., 0.3, 100., 600)
        co = o.populate_tracer(0.4, 8., 0.3, 100., 600)
        assert np.array_equal(cg[0], co[0])


def test_repeated_populate_and_determinism(gpu_available, pk_table):
    k, P = pk_table
    g = EZmock(Lbox=320., Ngrid=32, rng=0, seed=9, precision="f64")
    g.eval_growth_params(0.5, 0.31)
    g.setup_linear_pk(k, P)
    g.create_dens_field_from_ic()
    a = g.populate_tracer(0.5, 10., 0.3, 150., 700)
    b = g.populate_tracer(0.7, 6., 0.25, 50., 900)
    c = g.populate_tracer(0.5, 10., 0.3, 150., 700)
    for x, y in zip(a, c):
        assert np.array_equal(x, y)
    assert len(b[0]) != len(a[0]) or not np.array_equal(a[0], b[0])


def test_mt19937_instance_reused_across_seeds_and_sizes(gpu_available, pk_table):
    """MT19937 on one instance across seeds and tracer counts (the jump
    windows of each stream use are kept per seed and first draw, and built
    ahead of use on a second stream): catalogues equal to fresh instances'.
    50000 tracers need two 2^18-word segments of position and velocity
    draws, 3000 one, so the windows are rebuilt in both directions."""
    k, P = pk_table

    def made(seed=None):
        e = EZmock(Lbox=640., Ngrid=64, rng=1, seed=seed or 11, precision="f32")
        e.eval_growth_params(0.5, 0.31)
        e.setup_linear_pk(k, P)
        return e

    g = made()
    for seed, ns in ((11, (3000, 50000, 3000)), (12, (50000,)), (11, (50000,))):
        g.set_seed(seed)
        g.create_dens_field_from_ic()
        for n in ns:
            a = g.populate_tracer(0.5, 8., 0.3, 150., n)
            f = made(seed)
            f.create_dens_field_from_ic()
            b = f.populate_tracer(0.5, 8., 0.3, 150., n)
            assert len(a[0]) == len(b[0]) > 0.9 * n
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (seed, n)


def test_mt19937_reused_instance_mock512(gpu_available, pk_table):
    """The same at 512^3 (Lbox 2500 Mpc/h, 3,125,000 tracers), where the
    spectrum and bias draws need three-level jump trees built on two
    streams at once: seed 42, then 77, then 42 again on one instance, each
    equal to a fresh instance's noise field and catalogue."""
    k, P = pk_table

    def mock(e):
        e.create_dens_field_from_ic()
        e.populate_tracer_device(0.8, 8., 0.25, 200., 3125000)
        return e.field("spec"), e.device_catalog()

    def made(seed):
        e = EZmock(Lbox=2500., Ngrid=512, rng=1, seed=seed, precision="f32")
        e.eval_growth_params(0.5, 0.31)
        e.setup_linear_pk(k, P)
        e.keep_spectrum(True)
        return e

    g = made(42)
    fresh = {}
    for seed in (42, 77, 42):
        g.set_seed(seed)
        sg, cg = mock(g)
        if seed not in fresh:
            f = made(seed)
            fresh[seed] = mock(f)
            del f
        sf, cf = fresh[seed]
        assert np.array_equal(sg, sf), seed
        assert len(cg[0]) == len(cf[0]), seed
        for x, y in zip(cg, cf):
            assert np.array_equal(x, y), seed


def test_device_rsd_matches_reference_formula(gpu_available, pk_table):
    k, P = pk_table
    g = EZmock(Lbox=320., Ngrid=32, rng=0, seed=9, precision="f32")
    g.eval_growth_params(0.5, 0.31)
    g.setup_linear_pk(k, P)
    g.create_dens_field_from_ic()
    g.populate_tracer_device(0.5, 10., 0.3, 150., 700)
    x, y, z, vx, vy, vz = g.device_catalog()
    fac = 0.011384
    zr = g.rsd_z(fac)
    ref = z.astype(np.float64) + fac * vz.astype(np.float64)
    ref = np.where(ref >= 320., ref - 320., ref)
    ref = np.where(ref < 0, ref + 320., ref)
    assert np.array_equal(zr, ref)


def test_unsupported_configurations_fail_loudly(gpu_available, pk_table):
    k, P = pk_table
    g = EZmock(Lbox=320., Ngrid=40, rng=0)
    g.eval_growth_params(0.5, 0.31)
    g.setup_linear_pk(k, P)
    with pytest.raises(RuntimeError, match="invalid grid size"):
        g.create_dens_field_from_ic()


def test_mock256_fp64_against_oracle(gpu_available, pk_table):
    """Plain EZmock single mock, FP64, as the reference CPU path runs it:
    256^3, Lbox 1000, z 0.5, MRG32k3a seed 1, (0.5, 10, 0.3, 150),
    3e-4 (h/Mpc)^3 -> 300000 expected tracers."""
    g, o, cg, co = run_pair(256, 1000., "f64", seed=1, params=(0.5, 10., 0.3, 150.),
                            nexp=300000, pk=pk_table, nthread=os.cpu_count() or 1)
    rg, ro = g.field("rho"), o.field("rho")
    assert np.abs(rg - ro).max() <= 1e-10 * np.abs(ro).max()
    ntg, nto = g.field("nt"), o.field("nt")
    assert np.mean(ntg == nto) >= 0.9999
    pg, pO = g.pdf_info(), o.pdf_info()
    assert abs(pg["ntot"] - pO["ntot"]) <= 1e-4 * pO["ntot"]
    frac = pc.identical_tracer_fraction(ntg, cg, nto, co, 1e-9)
    print(f"256^3 FP64 mock: identical tracers {frac:.6f} of {len(co[0])}, ntot {pg['ntot']} "
          f"vs {pO['ntot']}")
    assert frac >= 0.9999


def test_fft_component_streams_same_fields(gpu_available, pk_table):
    """The three displacement components' last FFT passes on three streams
    (default) or on one stream (option 4 = 0): identical displacements."""
    k, P = pk_table
    out = {}
    for fs in (1, 0):
        g = EZmock(Lbox=384., Ngrid=64, rng=0, seed=5, precision="f32")
        g.set_option(4, fs)
        g.eval_growth_params(0.5, 0.31)
        g.setup_linear_pk(k, P)
        g.create_dens_field_from_ic()
        out[fs] = [g.field(f) for f in ("psi_x", "psi_y", "psi_z")]
    for a, b in zip(out[1], out[0]):
        assert np.array_equal(a, b)
