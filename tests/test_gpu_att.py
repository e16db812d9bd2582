"""att_part = true on the GPU engine (pop_tracer.c:517-582, 804-879):
tracers attached to particles by the reference's stream-ordered Bentley
selection, the other tracers at uniform positions, velocities psi x vfac
for attached tracers and the CIC gather for the others, and the sigma_v
scatter over the whole catalogue.

 * FP64, MRG32k3a and MT19937: the catalogue equals the oracle's (tracers
   per cell and the positions of the tracers not attached bit-exact; the
   attached tracers' positions, which follow the displacement fields,
   within 1e-12 of the box; velocities within 1e-9 of the largest).
 * FP32 and FP64 up to 128^3: the catalogue (sigma_v = 0) equals an
   independent replay of the NGP counts and the selection on stream 4N,
   with the other tracers' positions on 5N + 3t' (tests/pincheck.py).
 * More than 65535 particles in one tracer cell: EZMOCK_ERR_RHO_NPART, as
   the reference's uint16 counter (pop_tracer.c:569-576)."""
import numpy as np
import pytest

import pincheck as pc
from ezmock_amd import EZmock
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def pair(pk, Ng, L, prec, rng, seed, z=0.0):
    k, P = pk
    g = EZmock(Lbox=L, Ngrid=Ng, rng=rng, seed=seed, precision=prec)
    o = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=rng, seed=seed, nthread=4, prec=prec)
    for e in (g, o):
        e.eval_growth_params(z, 0.31)
        e.setup_linear_pk(k, P)
        e.create_dens_field_from_ic()
    return g, o


@pytest.mark.parametrize("Ng,L,rng,seed", [(32, 128., 0, 5), (48, 192., 1, 4357),
                                           (64, 256., 0, 11)])
def test_att_part_fp64_against_oracle(gpu_available, pk_table, Ng, L, rng, seed):
    g, o = pair(pk_table, Ng, L, "f64", rng, seed)
    nexp = int(0.01 * Ng ** 3)
    cg = g.populate_tracer(0.4, 8., 0.3, 120., nexp, att_part=True)
    co = o.populate_tracer(0.4, 8., 0.3, 120., nexp, att_part=True)
    assert np.array_equal(g.field("nt"), o.field("nt"))
    assert len(cg[0]) == len(co[0])
    iv = o.info()["iv"]
    print(f"{Ng}^3: {len(co[0])} tracers, {iv} attached")
    assert 0 < iv < len(co[0])
    # attached tracers sit at their particles, p gs with p = psi / gs + i:
    # the two displacement fields come from two FFTs and agree to rounding
    # (1e-13 relative), so these positions agree to rounding; the others
    # come from the random streams alone and are bit-exact
    for a, b in zip(cg[:3], co[:3]):
        assert np.abs(a[:iv] - b[:iv]).max() <= 1e-12 * L
        assert np.array_equal(a[iv:], b[iv:])
    vmax = max(np.abs(v).max() for v in co[3:])
    for a, b in zip(cg[3:], co[3:]):
        assert np.abs(a - b).max() <= 1e-9 * vmax


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("Ng,L", [(64, 256.), (128, 512.)])
def test_att_part_replayed_selection(gpu_available, pk_table, prec, Ng, L):
    k, P = pk_table
    g = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=21, precision=prec)
    g.eval_growth_params(0.0, 0.31)
    g.setup_linear_pk(k, P)
    g.create_dens_field_from_ic()
    cat = g.populate_tracer(0.4, 8., 0.3, 0., int(0.01 * Ng ** 3), att_part=True)
    iv = pc.check_att(g, cat, Ng, L, 21, prec, g.info()["vfac"])
    assert 0 < iv < len(cat[0])
    # a second call on the same field gives the same catalogue, and a call
    # without att_part is unaffected by the first
    cat2 = g.populate_tracer(0.4, 8., 0.3, 0., int(0.01 * Ng ** 3), att_part=True)
    for a, b in zip(cat, cat2):
        assert np.array_equal(a, b)
    c3 = g.populate_tracer(0.4, 8., 0.3, 0., int(0.01 * Ng ** 3))
    pc.check_positions(g, c3, Ng, L, 21, prec)


def test_att_part_particle_count_overflow(gpu_available):
    """Three slabs of 69632 particles each pulled onto a point: more than
    65535 particles in each of three dense cells, which host tracers (with
    3000 tracers over the 53248 cells of density 1 left, every seed tried
    from 1 to 7 gives the error in the oracle)."""
    Ng, L = 64, 64.
    q = np.arange(Ng, dtype=np.float64)
    i, j, kk = np.meshgrid(q, q, q, indexing="ij")
    groups = [(i < 17), (i >= 17) & (i < 34), (i >= 34) & (i < 51)]
    points = ((40.2, 20.2, 50.2), (8.2, 40.2, 24.2), (12.2, 30.2, 56.2))
    disp = []
    for a, b in zip((i, j, kk), points):
        d = np.zeros_like(a)
        for gm, t in zip(groups, b):
            d = np.where(gm, t - a, d)
        disp.append(d.ravel())
    codes = []
    for e in (EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=3, precision="f64"),
              po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=3, nthread=4, prec="f64")):
        e.set_growth_params(0.6, 60.)
        e.create_dens_field_from_disp(*disp)
        with pytest.raises(RuntimeError) as ex:
            e.populate_tracer(0.5, 8., 0.3, 10., 3000, att_part=True)
        codes.append(str(ex.value))
    assert codes[0] == codes[1] == "too many particles per cell"
