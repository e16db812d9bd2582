"""Pins the oracle's tracer stages (oracle/orc_tracer.c).

 - eval_pdf against closed-form values computed independently for the
   benchmark workloads;
 - the PDF walk (literal restatement of pop_tracer.c:413-455) against the
   rank formulation the GPU engine uses (sort densities descending; bin p
   takes num[p] ranks and every further cell equal to its lowest value), on
   arrays with planted ties at the bin boundaries;
 - thread-count independence of the bias pairing and of the stream offsets;
 - the reference's add_vel_scatter thread split (pop_tracer.c:1058-1067)
   always starts threads on even tracer indices, so the pair formulation
   (pair q at draws 8N + 6q) is exact;
 - tracers lie inside their host cells, per-cell counts equal nt, and the
   velocity scatter has the requested dispersion."""
import numpy as np
import pytest

from oracle import pyoracle as po


@pytest.mark.parametrize("nexp,pbase,prec,expect", [
    (300000, 0.3, "f64", (11, 210003, 300003)),
    (3125000, 0.25, "f32", (11, 2343755, 3125006)),
    (13500000, 0.3, "f64", (14, 9450004, 13500004)),
])
def test_eval_pdf_closed_form(nexp, pbase, prec, expect):
    e, info = po.eval_pdf(nexp, pbase, 10 ** 10, prec)
    assert e == 0
    assert (info["nmax"], info["ncell"], info["ntot"]) == expect
    # definition: A b^(n+1) < 1/2 at the final n, num[i] = round(A b^(n-i))
    n = info["nmax"]
    b = float(np.float32(pbase)) if prec == "f32" else pbase
    A = nexp / sum(b ** m * m for m in range(1, n + 1))
    assert A * b ** (n + 1) < 0.5
    assert list(info["pdf"]) == [int(np.round(A * b ** (n - i))) for i in range(n)]


def test_eval_pdf_errors():
    assert po.eval_pdf(10 ** 6, 0.9999, 10 ** 9, "f64")[0] == 37   # RHO_PBASE
    assert po.eval_pdf(10 ** 6, 0.3, 1000, "f64")[0] == 33         # PAR_PBASE


def walk_by_rank(rhot, ncell, pdf):
    """Rank formulation of the reference walk (what the GPU engine does)."""
    r = np.asarray(rhot, dtype=np.float64)
    hist = np.zeros(64, dtype=np.int64)
    nz = r[r != 0]
    e = np.array([np.frexp(v)[1] - 1 for v in nz])
    np.add.at(hist, np.clip(5 - e, 0, 63), 1)
    c = np.cumsum(hist)
    imax = int(np.searchsorted(c, ncell))
    thres = 2.0 ** (5 - imax) if imax < 63 else np.finfo(rhot.dtype).tiny
    cand = np.sort(rhot[rhot >= thres].astype(np.float64))[::-1]
    nt = np.zeros(len(r), dtype=np.int16)
    nmax = len(pdf)
    s, out, bounds = 0, [], []
    for p in range(nmax):
        if pdf[p] == 0:
            out.append(0)
            bounds.append(bounds[-1] if bounds else np.inf)
            continue
        e_ = s + pdf[p]
        assert e_ <= len(cand)
        v = cand[e_ - 1]
        while e_ < len(cand) and cand[e_] == v:
            e_ += 1
        out.append(e_ - s)
        bounds.append(v)
        s = e_
    for p in range(nmax - 1, -1, -1):
        hi = bounds[p - 1] if p > 0 else np.inf
        sel = (r >= bounds[p]) & (r < hi) & (r >= thres)
        nt[sel] = nmax - p
    return nt, np.array(out), s, int(sum(o * (nmax - p) for p, o in enumerate(out)))


@pytest.mark.parametrize("prec", ["f32", "f64"])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_walk_rank_formulation_matches_literal_walk(prec, seed):
    rs = np.random.default_rng(seed)
    n = 60000
    rhot = rs.exponential(2.0, n)
    rhot[rs.random(n) < 0.45] = 0
    rt = po.rtype(prec)
    rhot = rhot.astype(rt)
    e, info = po.eval_pdf(6000, 0.3, n, prec)
    # plant ties exactly at each bin boundary of the descending order
    order = np.argsort(-rhot.astype(np.float64), kind="stable")
    cum = np.cumsum(info["pdf"])
    for c in cum[:-1]:
        v = rhot[order[c - 1]]
        rhot[order[c:c + 1 + seed]] = v
    e1, nt1, p1, nc1, ntot1 = po.pdf_walk(rhot, info["ncell"], info["pdf"], prec)
    assert e1 == 0
    nt2, p2, nc2, ntot2 = walk_by_rank(rhot, info["ncell"], info["pdf"])
    assert np.array_equal(nt1, nt2)
    assert np.array_equal(p1, p2)
    assert (nc1, ntot1) == (nc2, ntot2)
    assert nc1 > info["ncell"]           # the planted ties were absorbed


def make(Ng=24, L=300., prec="f64", nthread=1, seed=5):
    from ezmock_amd import inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=seed, nthread=nthread, prec=prec)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    return ez


def exact_field(Ng, L, prec, nthread, seed=11):
    """Half-cell displacements: every CIC weight is 0, 1/2 or 1, so the
    density is exact whatever the summation order (threads)."""
    rs = np.random.default_rng(seed)
    gs = L / Ng
    dx = 0.5 * gs * rs.integers(0, 2, Ng ** 3)
    dy = 0.5 * gs * rs.integers(0, 2, Ng ** 3)
    dz = np.zeros(Ng ** 3)
    ez = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=5, nthread=nthread, prec=prec)
    ez.set_growth_params(0.6, 60.)
    ez.create_dens_field_from_disp(dx, dy, dz, deepcopy=True)
    return ez


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_thread_count_independent_catalogue(prec):
    a, b = exact_field(24, 240., prec, 1), exact_field(24, 240., prec, 7)
    assert np.array_equal(a.field("rho"), b.field("rho"))
    ca = a.populate_tracer(0.6, 8., 0.3, 100., 900)
    cb = b.populate_tracer(0.6, 8., 0.3, 100., 900)
    assert np.array_equal(a.field("rhot"), b.field("rhot"))
    assert np.array_equal(a.field("nt"), b.field("nt"))
    for x, y in zip(ca, cb):
        assert np.array_equal(x, y)


def test_reference_scatter_split_starts_on_even_tracers():
    for n in list(range(1, 80)) + [999, 1000, 123457]:
        for nthread in range(1, 19):
            pnum, rem = n // nthread, n % nthread
            starts, end = [], 0
            for tid in range(nthread):
                if tid < rem:
                    pcnt = pnum + (tid & 1) + ((pnum ^ tid) & 1)
                    istart = (pnum + 1) * tid - (tid & (~pnum) & 1)
                else:
                    pcnt = pnum + (pnum & 1) * ((((tid ^ rem) & 1) << 1) - 1)
                    istart = pnum * tid + rem - (rem & (~pnum) & 1) - ((tid ^ rem) & pnum & 1)
                if tid == nthread - 1:
                    pcnt = n - istart
                assert istart == end, (n, nthread, tid)
                end = istart + pcnt
                if pcnt > 0:
                    assert istart % 2 == 0, (n, nthread, tid)
            assert end == n


def test_tracers_inside_host_cells_and_counts():
    Ng, L = 24, 300.
    ez = make(Ng, L)
    x, y, z, vx, vy, vz = ez.populate_tracer(0.4, 8., 0.3, 100., 600)
    nt = ez.field("nt").reshape(Ng, Ng, Ng)
    info = ez.pdf_info()
    assert len(x) == info["ntot"] == nt.sum()
    gs = L / Ng
    for arr in (x, y, z):
        assert arr.min() >= 0 and arr.max() <= L
    # each tracer lies within half a cell of the centre of some tracer cell
    ci = np.floor(np.mod(np.stack([x, y, z], 1) / gs + 0.5, Ng)).astype(int)
    cnt = np.zeros((Ng, Ng, Ng), int)
    np.add.at(cnt, tuple(ci.T), 1)
    assert np.array_equal(cnt, nt)
    # bins are rank ordered: every cell with more tracers is denser
    rt = ez.field("rhot").reshape(Ng, Ng, Ng)
    for m in range(1, nt.max()):
        lo = rt[nt == m]
        hi = rt[nt == m + 1]
        if len(lo) and len(hi):
            assert lo.max() <= hi.min()


def test_velocity_scatter_dispersion():
    ez = make(32, 400.)
    _, _, _, vx0, vy0, vz0 = ez.populate_tracer(0.3, 8., 0.3, 0., 3000)
    _, _, _, vx1, vy1, vz1 = ez.populate_tracer(0.3, 8., 0.3, 250., 3000)
    d = np.concatenate([vx1 - vx0, vy1 - vy0, vz1 - vz0])
    assert abs(d.std() / 250. - 1) < 0.03
    assert abs(d.mean()) < 10


def test_att_part_runs_and_counts():
    ez = make(16, 200.)
    out = ez.populate_tracer(0.4, 8., 0.3, 100., 300, att_part=True)
    info = ez.pdf_info()
    assert len(out[0]) == info["ntot"]
    assert ez.info()["iv"] > 0


def test_populate_errors():
    ez = make(16, 200.)
    for params, code in [((-1, 8, 0.3, 10), 31), ((0.5, 0, 0.3, 10), 32),
                         ((0.5, 8, 1.0, 10), 33), ((0.5, 8, 0.3, -1), 34)]:
        e2 = make(16, 200.)
        with pytest.raises(po.OracleError) as e:
            e2.populate_tracer(*params, 300)
        assert e.value.code == code
    e3 = make(16, 200.)
    with pytest.raises(po.OracleError) as e:
        e3.populate_tracer(0.5, 8, 0.3, 10, 100)
    assert e.value.code == 15                  # LOWEXP
    e4 = make(16, 200.)
    with pytest.raises(po.OracleError) as e:
        e4.populate_tracer(50., 8, 0.3, 10, 3000)
    assert e.value.code == 36                  # RHO_LESS
