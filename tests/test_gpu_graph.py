"""Populate with the selection on the device, and whole mocks replayed from
captured HIP graphs (EZmock_ext_mock), against the host-driven path and the
plain calls: catalogues, PDFs and errors identical (bit-exact: the same
kernels draw the same streams; only where the counts live differs).

The host-driven populate (option 7 = 0) reads the valid-cell count, the
density histogram and the PDF walk's results back to the host between its
stages, as it has been pinned against the oracle; the device selection keeps
them on the GPU.  Option 7 = 2 gives the device selection capacities of a
fraction of the counts, so that its redo by the host-driven path runs.
A captured mock is replayed several times, with other instances' work and
synchronisations between the replays, and every replay is checked against
the plain calls."""
import numpy as np
import pytest

from ezmock_amd import EZmock

pytestmark = pytest.mark.gpu

PARAMS = [(0.5, 10., 0.3, 150.), (0.8, 8., 0.25, 200.), (0.0, 6., 0.4, 0.),
          (1.1, 15., 0.15, 300.)]


def made(pk, Ng, L, seed, prec, opt=None):
    e = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=seed, precision=prec)
    for k, v in (opt or {}).items():
        e.set_option(k, v)
    e.eval_growth_params(0.5, 0.31)
    e.setup_linear_pk(*pk)
    return e


def same_catalogue(a, b, tag, vtol=0.0):
    """Counts and positions bit-equal; velocities bit-equal, or within vtol
    of the largest speed (FP64 fields painted in two processes of the
    catalogue: double atomic sums in LDS round in the order they land)."""
    assert len(a[0]) == len(b[0]), tag
    for i, (x, y) in enumerate(zip(a, b)):
        if i < 3 or vtol == 0:
            assert np.array_equal(x, y), (tag, i)
        else:
            assert np.abs(x - y).max() <= vtol * np.abs(y).max(), (tag, i)


def same_pdf(a, b, tag):
    pa, pb = a.pdf_info(), b.pdf_info()
    for k in ("nmax", "ncell", "ntot", "ncell0", "ntot0"):
        assert pa[k] == pb[k], (tag, k)
    assert np.array_equal(pa["pdf"], pb["pdf"]), tag


@pytest.mark.parametrize("Ng,L,prec", [(32, 320., "f64"), (48, 400., "f32"),
                                       (64, 640., "f32"), (128, 1000., "f32")])
def test_device_selection_matches_host_path(gpu_available, pk_table, Ng, L, prec):
    """One field, populated by both paths in turn (option 7 switched
    between the calls): identical catalogues and PDFs."""
    e = made(pk_table, Ng, L, 3, prec)
    e.create_dens_field_from_ic()
    nexp = Ng ** 3 // 40
    for i, p in enumerate(PARAMS):
        tag = (Ng, prec, p)
        out = {}
        for opt in (1, 0):
            e.set_option(7, opt)
            n = e.populate_tracer_device(*p, nexp + 97 * i)
            out[opt] = (n, e.device_catalog(), e.pdf_info(),
                        e.populate_tracer(*p, nexp + 97 * i))
        assert out[1][0] == out[0][0], tag
        same_catalogue(out[1][1], out[0][1], tag)
        # host arrays (the device selection synchronises before the
        # velocities, whose staging overlaps the positions' copies)
        same_catalogue(out[1][3], out[0][3], tag)
        for k in ("nmax", "ncell", "ntot", "ncell0", "ntot0"):
            assert out[1][2][k] == out[0][2][k], (tag, k)
        assert np.array_equal(out[1][2]["pdf"], out[0][2]["pdf"]), tag
    c = e.path_counts()
    assert c["devsel"] == 2 * len(PARAMS) and c["host"] == 2 * len(PARAMS) and c["retry"] == 0


def test_device_selection_capacity_redo(gpu_available, pk_table):
    """Capacities short of the counts (option 7 = 2 on a new instance): the
    populate is redone host-driven, with the catalogue the host-driven path
    gives on the same field; later calls fit the grown buffers."""
    for dev in (True, False):
        r = made(pk_table, 64, 640., 5, "f32", {7: 2})
        r.create_dens_field_from_ic()
        for p in PARAMS[:2]:
            out = {}
            for opt in (2, 0):
                r.set_option(7, opt)
                if dev:
                    r.populate_tracer_device(*p, 7000)
                    out[opt] = (r.device_catalog(), r.pdf_info())
                else:
                    out[opt] = (r.populate_tracer(*p, 7000), r.pdf_info())
            same_catalogue(out[2][0], out[0][0], (dev, p))
            assert out[2][1]["ntot"] == out[0][1]["ntot"]
            assert np.array_equal(out[2][1]["pdf"], out[0][1]["pdf"])
        c = r.path_counts()
        assert c["retry"] >= 1 and c["devsel"] == 2, c


def test_device_selection_errors(gpu_available, pk_table):
    """Too few valid cells for the PDF: the same error from both paths."""
    for opt in (1, 0):
        e = made(pk_table, 32, 320., 2, "f64", {7: opt})
        e.create_dens_field_from_ic()
        with pytest.raises(RuntimeError) as ex:
            e.populate_tracer_device(50., 10., 0.3, 100., 9000)
        assert "density" in str(ex.value).lower() or "cell" in str(ex.value).lower()
        # the instance stays usable once the (sticky) error code is cleared
        e._err.value = 0
        assert e.populate_tracer_device(0.5, 10., 0.3, 100., 900) > 0


@pytest.mark.parametrize("Ng,L,prec", [(32, 320., "f64"), (64, 640., "f32"),
                                       (128, 1000., "f32"), (256, 1000., "f32")])
def test_mock_graph_replays_match_plain_calls(gpu_available, pk_table, Ng, L, prec):
    """A sweep of seeds and bias parameters on one instance (graphs captured
    after the first mock, then replayed) against an instance that makes the
    plain calls every time (option 8 = 0)."""
    g = made(pk_table, Ng, L, 11, prec)
    p = made(pk_table, Ng, L, 11, prec, {8: 0})
    nexp = Ng ** 3 // 50
    sweep = [(11, PARAMS[0], nexp), (12, PARAMS[1], nexp), (13, PARAMS[2], nexp + 1000),
             (11, PARAMS[0], nexp), (14, PARAMS[3], nexp // 2)]
    for i, (seed, par, n) in enumerate(sweep):
        tag = (Ng, prec, i, seed)
        for e in (g, p):
            e.set_seed(seed)
        a = g.mock(*par, n)
        # plain calls, and some other work on the device between replays
        p.create_dens_field_from_ic()
        b = p.populate_tracer_device(*par, n)
        assert a == b, tag
        vt = 1e-9 if prec == "f64" else 0.0
        same_catalogue(g.device_catalog(), p.device_catalog(), tag, vt)
        same_pdf(g, p, tag)
        assert np.array_equal(g.field("psi_x"), p.field("psi_x")), tag
        rg, rp = g.field("rho"), p.field("rho")
        assert np.abs(rg - rp).max() <= (1e-12 * np.abs(rp).max() if vt else 0), tag
        p.sync()
    c = g.path_counts()
    assert c["capture"] >= 1 and c["replay"] >= len(sweep) - 2, c
    assert p.path_counts()["replay"] == 0


def test_mock_graph_populate_after_replay(gpu_available, pk_table):
    """After a replayed mock the instance holds an ordinary field: more
    populate calls on it (the calibration loop), a field from displacements
    in between, and a replay again."""
    g = made(pk_table, 64, 640., 21, "f32")
    p = made(pk_table, 64, 640., 21, "f32", {8: 0})
    for seed in (21, 22):
        g.set_seed(seed)
        p.set_seed(seed)
        g.mock(*PARAMS[1], 5000)
        p.create_dens_field_from_ic()
        p.populate_tracer_device(*PARAMS[1], 5000)
    for par in PARAMS:
        same_catalogue(g.populate_tracer(*par, 6000), p.populate_tracer(*par, 6000), par)
    psi = [p.field(f) for f in ("psi_x", "psi_y", "psi_z")]
    for e in (g, p):
        e.create_dens_field_from_disp(*psi, deepcopy=True)
    same_catalogue(g.populate_tracer(*PARAMS[0], 4000), p.populate_tracer(*PARAMS[0], 4000), 0)
    for e in (g, p):
        e.set_seed(23)
    a = g.mock(*PARAMS[2], 5000)
    p.create_dens_field_from_ic()
    assert a == p.populate_tracer_device(*PARAMS[2], 5000)
    same_catalogue(g.device_catalog(), p.device_catalog(), "after disp")
    assert g.path_counts()["replay"] >= 2


@pytest.mark.parametrize("Ng,L,prec,fixamp,iphase", [
    (32, 320., "f64", False, False), (48, 400., "f64", True, True),
    (64, 640., "f32", False, False), (64, 640., "f32", True, False),
    (96, 600., "f32", False, True), (256, 1000., "f32", False, False)])
def test_spectrum_generated_in_first_fft_pass(gpu_available, pk_table, Ng, L, prec, fixamp,
                                              iphase):
    """The spectrum generated inside the first FFT pass (option 9, default;
    the kz = 0 and Nyquist planes by their own kernel) against the
    spectrum kernel whose output the pass reads (option 9 = 0): identical
    displacement fields, for two seeds (at 256^3 the first field of the
    grid size times the FFT schedules on the separate kernel's spectrum)."""
    f = made(pk_table, Ng, L, 5, prec)
    s = made(pk_table, Ng, L, 5, prec, {9: 0})
    for seed in (5, 6):
        for e in (f, s):
            e.set_seed(seed)
            e.create_dens_field_from_ic(fixamp=fixamp, iphase=iphase)
        for c in ("psi_x", "psi_y", "psi_z"):
            assert np.array_equal(f.field(c), s.field(c)), (seed, c)
        rf, rs = f.field("rho"), s.field("rho")
        assert np.abs(rf - rs).max() <= (1e-12 * np.abs(rs).max() if prec == "f64" else 0)
    # and the whole mock, replayed from graphs, with the separate kernel
    n = Ng ** 3 // 50
    for e in (f, s):
        e.set_seed(9)
        e.mock(0.5, 10., 0.3, 150., n)
        e.set_seed(10)
        e.mock(0.5, 10., 0.3, 150., n)
    assert f.path_counts()["replay"] >= 1
    assert np.array_equal(f.field("psi_y"), s.field("psi_y"))
