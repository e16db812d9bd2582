"""Closed-form pins of the tracer stages, shared by the oracle tests (CPU)
and their mirrors on the GPU engine.

Each check builds a field whose tracer stage has a closed form or an
invariant that the engine under test does not compute itself, so a dropped
term, a swapped operand or a wrong stream offset fails it:

 * bias (pop_tracer.c:70-79, 245-289): with zero displacements rho == 1 in
   every cell, so every cell is valid and valid-cell rank == cell index.  The
   stochastic factor g' of each cell is recovered from rho_t by dividing out
   (1 - exp(-1/rho_exp)) and inverting g' = 1 + g (g >= 0) / exp(g) (g < 0).
   Cells (2p, 2p+1) must then be the two Box-Muller normals of draws
   2N + 2p (radius, -2 ln u) and 2N + 2p + 1 (angle, 2 pi u), with
   lambda = 10: g1^2 + g2^2 = lambda^2 (-2 ln u) and atan2(g2, g1) = 2 pi u'.
 * positions (pop_tracer.c:919, 946-985): tracer t of the cell-major
   catalogue sits at (i + u - 1/2) gs in its host cell, with u = draw
   5N + 3t (+1, +2 for y, z); recovered as x/gs - i + 1/2.
 * velocities (pop_tracer.c:593-753): a uniform displacement c gives
   v = c vfac for every tracer; displacements that depend on the z index
   only (psi_z compresses planes, psi_x shears them, psi_y = 0) reduce the
   mass-weighted CIC paint and the CIC gather to one dimension, evaluated
   here with numpy.
 * sigma_v scatter (pop_tracer.c:1071-1094): two populate calls that differ
   only in sigma_v; (v(s) - v(0)) / s of tracer pair q must be the
   Box-Muller pairs of draws 8N + 6q .. 8N + 6q + 5 in the order
   (vx_a, vy_a), (vz_a, vx_b), (vy_b, vz_b), checked by the same radius and
   angle invariants.

Draws come from the oracle's MRG32k3a stream, itself pinned against
L'Ecuyer's published constants and brute-force stepping
(tests/test_oracle_rng.py); for MRG32k3a get_double == get_double_pos."""
import numpy as np

from oracle import pyoracle as po

LAMBDA = 10.0                   # config.h:72, width of the bias Gaussian


def rtype(prec):
    return np.float32 if prec == "f32" else np.float64


def draws(seed, start, n):
    return po.Rng(0, seed, start, prec="f64").double(n)


def _circ(d):
    """Angle difference folded into [-pi, pi)."""
    return np.mod(d + np.pi, 2 * np.pi) - np.pi


def check_box_muller_pairs(g1, g2, u_r, u_t, scale, tol):
    """g1, g2: recovered normals (times scale); u_r, u_t: the radius and angle
    draws.  Radius and angle invariants, with an arc-length tolerance so that
    pairs near the origin do not amplify rounding."""
    R = np.hypot(g1, g2)
    Rexp = scale * np.sqrt(-2 * np.log(u_r))
    assert np.all(np.abs(R - Rexp) <= tol * (1 + Rexp)), np.abs(R - Rexp).max()
    dth = _circ(np.arctan2(g2, g1) - 2 * np.pi * u_t)
    assert np.all(np.abs(dth) * Rexp <= tol * (1 + Rexp)), (np.abs(dth) * Rexp).max()


def zero_field(ez, Ng):
    z = np.zeros(Ng ** 3)
    ez.set_growth_params(0.6, 60.)
    ez.create_dens_field_from_disp(z, z, z, deepcopy=True)


def check_bias(ez, Ng, seed, prec, rho_exp=6.0, nexp=500):
    """ez: a fresh instance (Lbox, Ngrid=Ng, MRG32k3a, seed) of the engine."""
    rt = rtype(prec)
    zero_field(ez, Ng)
    assert np.all(ez.field("rho") == 1)
    ez.populate_tracer(0.5, rho_exp, 0.3, 0., nexp)
    N = Ng ** 3
    rhot = ez.field("rhot").astype(np.float64)
    # iexp = -1 / rho_exp evaluated in `real` (pop_tracer.c:162)
    iexp = float(rt(-1) / rt(rho_exp))
    fac = 1 - np.exp(iexp)
    gp = rhot / fac
    g = np.where(gp >= 1, gp - 1, np.log(gp))
    u = draws(seed, 2 * N, N)
    tol = 1e-11 if prec == "f64" else 2e-5
    check_box_muller_pairs(g[0::2], g[1::2], u[0::2], u[1::2], LAMBDA, tol)
    return g


def check_positions(ez, cat, Ng, L, seed, prec):
    x, y, z = (np.asarray(a, dtype=np.float64) for a in cat[:3])
    nt = ez.field("nt").astype(np.int64)
    N = Ng ** 3
    cells = np.repeat(np.arange(N), nt)
    assert len(cells) == len(x)
    idx = np.unravel_index(cells, (Ng, Ng, Ng))
    u = draws(seed, 5 * N, 3 * len(x)).reshape(-1, 3)
    gs = float(rtype(prec)(L / Ng))
    tol = 1e-12 if prec == "f64" else 2.5e-7 * Ng
    for col, (a, i) in enumerate(zip((x, y, z), idx)):
        rec = a / gs - i + 0.5
        d = np.mod(rec - u[:, col] + Ng / 2, Ng) - Ng / 2
        assert np.abs(d).max() <= tol, (col, np.abs(d).max())


def check_sigma_v(ez, Ng, seed, prec, params=(0.3, 8., 0.3), nexp=600, sigma=150.):
    """ez holds a field; two populate calls that differ only in sigma_v."""
    c0 = ez.populate_tracer(*params, 0., nexp)
    c1 = ez.populate_tracer(*params, sigma, nexp)
    for a, b in zip(c0[:3], c1[:3]):
        assert np.array_equal(a, b)
    N = Ng ** 3
    n = len(c0[0])
    dv = [np.asarray(b, np.float64) - np.asarray(a, np.float64) for a, b in zip(c0[3:], c1[3:])]
    v0 = np.max([np.abs(np.asarray(a, np.float64)).max() for a in c0[3:]])
    tol = 1e-12 if prec == "f64" else 4e-7 * (v0 + 5 * sigma) / sigma
    npair = n // 2
    u = draws(seed, 8 * N, 6 * ((n + 1) // 2)).reshape(-1, 6)
    a, b = 2 * np.arange(npair), 2 * np.arange(npair) + 1
    vx, vy, vz = (d / sigma for d in dv)
    pairs = [(vx[a], vy[a]), (vz[a], vx[b]), (vy[b], vz[b])]
    for m, (g1, g2) in enumerate(pairs):
        check_box_muller_pairs(g1, g2, u[:npair, 2 * m], u[:npair, 2 * m + 1], 1.0, tol)
    if n % 2:                       # lone last tracer: (vx, vy) and vz
        t, q = n - 1, npair
        check_box_muller_pairs(vx[t:], vy[t:], u[q:, 0], u[q:, 1], 1.0, tol)
        r = np.sqrt(-2 * np.log(u[q, 2]))
        assert abs(vz[t] - r * np.cos(2 * np.pi * u[q, 3])) <= tol * (1 + r)
    return n


def check_velocity_uniform(ez, Ng, L, prec, c=(3.7, -1.25, 0.6)):
    rt = rtype(prec)
    N = Ng ** 3
    ez.set_growth_params(0.6, 60.)
    ez.create_dens_field_from_disp(*[np.full(N, ci) for ci in c], deepcopy=True)
    cat = ez.populate_tracer(0.5, 8., 0.3, 0., max(200, N // 20))
    vfac = ez.info()["vfac"]
    tol = 1e-12 if prec == "f64" else 3e-6
    for v, ci in zip(cat[3:], c):
        ref = float(rt(ci)) * vfac
        assert np.abs(np.asarray(v, np.float64) - ref).max() <= tol * abs(ref)


def zshear_fields(Ng, L, prec):
    """psi_z = 0.3 cells sin(2 pi k / Ng) (planes compressed, no crossing),
    psi_x = 0.45 cells cos(2 pi k / Ng + 0.7) (planes sheared), psi_y = 0."""
    rt = rtype(prec)
    gs = L / Ng
    k = np.arange(Ng)
    fz = (0.3 * gs * np.sin(2 * np.pi * k / Ng)).astype(rt)
    fx = (0.45 * gs * np.cos(2 * np.pi * k / Ng + 0.7)).astype(rt)
    sh = (Ng, Ng, Ng)
    px = np.broadcast_to(fx, sh).reshape(-1).copy()
    pz = np.broadcast_to(fz, sh).reshape(-1).copy()
    return px, np.zeros(Ng ** 3, rt), pz, fx, fz


def paint_1d(Ng, L, prec, fz, vals):
    """One-dimensional CIC of the planes k -> k + psi_z(k) / gs, carrying
    `vals` (density: ones), floor and single wrap as dens_field.c:233-310."""
    rt = rtype(prec)
    igs = rt(Ng / L)
    out = np.zeros(Ng)
    for k in range(Ng):
        zz = rt(fz[k] * igs + rt(k))
        z0 = int(np.floor(zz))
        w1 = float(zz - rt(z0))
        w0 = 1 - w1
        z0 %= Ng
        out[z0] += w0 * vals[k]
        out[(z0 + 1) % Ng] += w1 * vals[k]
    return out


def check_velocity_zshear(ez, Ng, L, prec):
    rt = rtype(prec)
    px, py, pz, fx, fz = zshear_fields(Ng, L, prec)
    ez.set_growth_params(0.6, 60.)
    ez.create_dens_field_from_disp(px, py, pz, deepcopy=True)
    rho1 = paint_1d(Ng, L, prec, fz, np.ones(Ng))
    rho = ez.field("rho").reshape(Ng, Ng, Ng).astype(np.float64)
    tol_rho = 1e-12 if prec == "f64" else 4e-7 * Ng
    assert np.abs(rho - rho1[None, None, :]).max() <= tol_rho
    ux = paint_1d(Ng, L, prec, fz, fx.astype(np.float64)) / rho1
    uz = paint_1d(Ng, L, prec, fz, fz.astype(np.float64)) / rho1
    cat = ez.populate_tracer(0.3, 8., 0.3, 0., max(300, Ng ** 3 // 20))
    vfac = ez.info()["vfac"]
    igs = rt(Ng / L)
    zz = (np.asarray(cat[2], rt) * igs).astype(np.float64)
    z0 = zz.astype(np.int64)
    w1 = zz - z0
    w0 = 1 - w1
    z0 = np.where(z0 == Ng, 0, z0)
    z1 = np.where(z0 == Ng - 1, 0, z0 + 1)
    den = rho1[z0] * w0 + rho1[z1] * w1
    vx = vfac * (ux[z0] * rho1[z0] * w0 + ux[z1] * rho1[z1] * w1) / den
    vz = vfac * (uz[z0] * rho1[z0] * w0 + uz[z1] * rho1[z1] * w1) / den
    scale = vfac * 0.45 * L / Ng
    tol = 1e-11 if prec == "f64" else 1e-5
    assert np.abs(np.asarray(cat[3], np.float64) - vx).max() <= tol * scale
    assert np.abs(np.asarray(cat[5], np.float64) - vz).max() <= tol * scale
    assert np.abs(np.asarray(cat[4], np.float64)).max() <= tol * scale
    return len(cat[0])


def tracer_keys(nt):
    """(host cell, rank within the cell) of each tracer of a cell-major
    catalogue, as one integer (at most 1024 tracers per cell)."""
    nt = np.asarray(nt, np.int64)
    cells = np.repeat(np.arange(len(nt)), nt)
    start = np.cumsum(nt) - nt
    rank = np.arange(len(cells)) - np.repeat(start, nt)
    return cells * 1024 + rank


def identical_tracer_fraction(nt_a, cat_a, nt_b, cat_b, vtol):
    """Fraction of tracers identical in both catalogues, matched by host cell
    and rank within the cell: positions bit-equal, velocities within vtol
    (relative to the largest speed).  Divided by the larger catalogue."""
    ka, kb = tracer_keys(nt_a), tracer_keys(nt_b)
    assert len(ka) == len(cat_a[0]) and len(kb) == len(cat_b[0])
    _, ia, ib = np.intersect1d(ka, kb, assume_unique=True, return_indices=True)
    same = np.ones(len(ia), bool)
    for a, b in zip(cat_a[:3], cat_b[:3]):
        same &= np.asarray(a)[ia] == np.asarray(b)[ib]
    vmax = max(1.0, max(np.abs(np.asarray(b, np.float64)).max() for b in cat_b[3:]))
    for a, b in zip(cat_a[3:], cat_b[3:]):
        same &= np.abs(np.asarray(a, np.float64)[ia] - np.asarray(b, np.float64)[ib]) <= vtol * vmax
    return same.sum() / max(len(ka), len(kb), 1)


def identical_before_first_difference(nt_a, cat_a, nt_b, cat_b, vtol):
    """FP32 catalogues of two implementations can differ from the first cell
    whose tracer count differs on (a density rounded differently moves a
    cell across a PDF bin boundary, and the position stream offsets of all
    later tracers shift).  Returns the number of tracers before that cell,
    after checking that every one of them is identical in both catalogues."""
    nt_a, nt_b = np.asarray(nt_a, np.int64), np.asarray(nt_b, np.int64)
    diff = np.nonzero(nt_a != nt_b)[0]
    first = diff[0] if len(diff) else len(nt_a)
    n = int(nt_a[:first].sum())
    assert n == int(nt_b[:first].sum())
    vmax = max(1.0, max(np.abs(np.asarray(b, np.float64)).max() for b in cat_b[3:]))
    for a, b in zip(cat_a[:3], cat_b[:3]):
        assert np.array_equal(np.asarray(a)[:n], np.asarray(b)[:n])
    for a, b in zip(cat_a[3:], cat_b[3:]):
        d = np.abs(np.asarray(a, np.float64)[:n] - np.asarray(b, np.float64)[:n])
        assert n == 0 or d.max() <= vtol * vmax
    return n


def att_expected(ez, Ng, L, seed, prec, rng_kind=0):
    """Attached tracers of a populate call with att_part (pop_tracer.c:
    517-582, 804-879), replayed independently: NGP cells of the displaced
    particles in `real` arithmetic (wrapped once), particles per tracer
    cell, and Algorithm S (Bentley, Programming Pearls 12.2) over the
    particles in Lagrangian order with draws from stream position 4N, one
    per particle of a cell that still has fewer tracers than particles.
    Returns (indices of the attached particles in catalogue order, NGP
    coordinate array p (N x 3, real), tracers left per cell)."""
    rt = rtype(prec)
    N = Ng ** 3
    igs = rt(Ng / L)
    q = np.arange(Ng)
    lat = [a.ravel() for a in np.meshgrid(q, q, q, indexing="ij")]
    p = np.empty((N, 3), rt)
    c = np.empty((N, 3), np.int64)
    for d, f in enumerate(("psi_x", "psi_y", "psi_z")):
        x = (ez.field(f).astype(rt) * igs + lat[d].astype(rt)).astype(rt)
        x = np.where(x < 0, x + rt(Ng), np.where(x >= rt(Ng), x - rt(Ng), x)).astype(rt)
        p[:, d] = x
        cd = (x.astype(np.float64) + 0.5).astype(np.int64)
        c[:, d] = np.where(cd >= Ng, 0, cd)
    cell = (c[:, 0] * Ng + c[:, 1]) * Ng + c[:, 2]
    nt = ez.field("nt").astype(np.int64)
    host = nt[cell] > 0
    npart = np.bincount(cell[host], minlength=N)
    t, n = nt.copy(), npart.copy()
    ncontested = int(np.sum(host & (nt[cell] < npart[cell])))
    u = po.Rng(rng_kind, seed, 4 * N, prec="f64").double(ncontested)
    taken, d = [], 0
    for i in np.flatnonzero(host):
        ci = cell[i]
        if t[ci] == 0:
            continue
        if t[ci] < n[ci]:
            ran = u[d]
            d += 1
            if ran * n[ci] >= t[ci]:
                n[ci] -= 1
                continue
        t[ci] -= 1
        n[ci] -= 1
        taken.append(i)
    rest = np.maximum(nt - npart, 0)
    assert np.array_equal(t, rest)
    return np.array(taken, np.int64), p, rest


def check_att(ez, cat, Ng, L, seed, prec, vfac, rng_kind=0):
    """The catalogue of a populate call with att_part and sigma_v = 0:
    attached tracers first (positions p gs, velocities psi vfac of the
    replayed selection), then the others at uniform positions of draws
    5N + 3t' (t' counting only them) in the cells left with tracers."""
    rt = rtype(prec)
    taken, p, rest = att_expected(ez, Ng, L, seed, prec, rng_kind)
    iv = len(taken)
    assert len(cat[0]) == iv + rest.sum()
    gs = rt(L / Ng)
    for d in range(3):
        assert np.array_equal(np.asarray(cat[d])[:iv], (p[taken, d] * gs).astype(rt)), d
    for d, f in enumerate(("psi_x", "psi_y", "psi_z")):
        ref = (ez.field(f)[taken].astype(np.float64) * vfac).astype(rt)
        assert np.array_equal(np.asarray(cat[3 + d])[:iv], ref), d
    # the rest: host cells in cell order, u from 5N + 3t'
    N = Ng ** 3
    cells = np.repeat(np.arange(N), rest)
    idx = np.unravel_index(cells, (Ng, Ng, Ng))
    u = po.Rng(rng_kind, seed, 5 * N, prec="f64").double(3 * len(cells)).reshape(-1, 3)
    tol = 1e-12 if prec == "f64" else 2.5e-7 * Ng
    for col, i in enumerate(idx):
        rec = np.asarray(cat[col], np.float64)[iv:] / float(gs) - i + 0.5
        dd = np.mod(rec - u[:, col] + Ng / 2, Ng) - Ng / 2
        assert len(dd) == 0 or np.abs(dd).max() <= tol, col
    return iv
