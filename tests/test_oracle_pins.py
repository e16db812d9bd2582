"""Closed-form pins of the oracle's tracer stages (oracle/orc_tracer.c):
the bias transform and its Box-Muller pairing (pop_tracer.c:70-79,
245-289), tracer positions and their stream offsets (:919, 946-985), the
mass-weighted velocity paint and gather (:593-753) and the sigma_v scatter
order (:1071-1094).  The checks are in tests/pincheck.py; the same checks
run on the GPU engine in tests/test_gpu_pins.py.  The last tests show that
each check fails on a planted mistake (wrong lambda, swapped cos/sin, wrong
stream base, wrong sigma_v order)."""
import numpy as np
import pytest

import pincheck as pc
from oracle import pyoracle as po


def oracle(L, Ng, seed, prec):
    return po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=seed, nthread=1, prec=prec)


def ic_field(L, Ng, seed, prec):
    from ezmock_amd import inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = oracle(L, Ng, seed, prec)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    return ez


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("seed", [3, 77])
def test_bias_box_muller_pairs(prec, seed):
    pc.check_bias(oracle(160., 16, seed, prec), 16, seed, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_positions_from_stream_offsets(prec):
    ez = ic_field(300., 24, 5, prec)
    cat = ez.populate_tracer(0.4, 8., 0.3, 100., 600)
    pc.check_positions(ez, cat, 24, 300., 5, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sigma_v_draw_order(prec):
    ez = ic_field(300., 24, 9, prec)
    assert pc.check_sigma_v(ez, 24, 9, prec) > 100


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_velocity_uniform_displacement(prec):
    pc.check_velocity_uniform(oracle(200., 16, 4, prec), 16, 200., prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_velocity_mass_weighted_gather(prec):
    assert pc.check_velocity_zshear(oracle(240., 16, 4, prec), 16, 240., prec) > 100


# ---- the checks reject planted mistakes
def _bias_normals(prec="f64", seed=3, Ng=16):
    ez = oracle(160., Ng, seed, prec)
    g = pc.check_bias(ez, Ng, seed, prec)
    u = pc.draws(seed, 2 * Ng ** 3, Ng ** 3)
    return g, u


def test_bias_check_rejects_planted_mistakes():
    g, u = _bias_normals()
    pc.check_box_muller_pairs(g[0::2], g[1::2], u[0::2], u[1::2], pc.LAMBDA, 1e-11)
    bad = [
        (g[0::2], g[1::2], u[0::2], u[1::2], 5.0),           # wrong lambda
        (g[1::2], g[0::2], u[0::2], u[1::2], pc.LAMBDA),     # cos/sin swapped
        (g[0::2], g[1::2], u[1::2], u[0::2], pc.LAMBDA),     # radius/angle draws swapped
    ]
    u2 = pc.draws(3, 2 * 16 ** 3 + 2, 16 ** 3)               # base off by one pair
    bad.append((g[0::2], g[1::2], u2[0::2], u2[1::2], pc.LAMBDA))
    for g1, g2, ur, ut, lam in bad:
        with pytest.raises(AssertionError):
            pc.check_box_muller_pairs(g1, g2, ur, ut, lam, 1e-11)


def test_positions_check_rejects_wrong_base():
    ez = ic_field(300., 24, 5, "f64")
    cat = ez.populate_tracer(0.4, 8., 0.3, 100., 600)
    shifted = tuple(np.roll(a, 1) for a in cat[:3]) + cat[3:]
    with pytest.raises(AssertionError):
        pc.check_positions(ez, shifted, 24, 300., 5, "f64")
    with pytest.raises(AssertionError):
        pc.check_positions(ez, cat, 24, 300., 6, "f64")       # another seed's stream


def test_sigma_v_check_rejects_wrong_order():
    class Swapped:
        """Engine whose catalogue has vy and vz exchanged (planted)."""
        def __init__(self, ez):
            self.ez = ez

        def populate_tracer(self, *a):
            c = self.ez.populate_tracer(*a)
            return c[:3] + (c[3], c[5], c[4])

    ez = ic_field(300., 24, 9, "f64")
    with pytest.raises(AssertionError):
        pc.check_sigma_v(Swapped(ez), 24, 9, "f64")


def test_identical_fraction_counts_shifted_tracers():
    ez = ic_field(300., 24, 5, "f64")
    cat = ez.populate_tracer(0.4, 8., 0.3, 100., 600)
    nt = ez.field("nt")
    assert pc.identical_tracer_fraction(nt, cat, nt, cat, 0) == 1.0
    half = len(cat[0]) // 2
    moved = tuple(np.concatenate([a[:half], a[half:] + 1e-3]) for a in cat)
    f = pc.identical_tracer_fraction(nt, moved, nt, cat, 0)
    assert abs(f - half / len(cat[0])) < 1e-12


@pytest.mark.parametrize("prec,rng", [("f64", 0), ("f32", 0), ("f64", 1)])
def test_att_part_selection_replayed(prec, rng):
    """att_part (pop_tracer.c:517-582, 804-879): the oracle's catalogue
    against an independent replay of the NGP counts and the Bentley
    selection on stream 4N, and the other tracers' positions on 5N + 3t'."""
    from ezmock_amd import inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = po.EZmockOracle(Lbox=96., Ngrid=24, rng=rng, seed=5, nthread=1, prec=prec)
    ez.eval_growth_params(0.0, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    cat = ez.populate_tracer(0.4, 8., 0.3, 0., 800, att_part=True)
    iv = pc.check_att(ez, cat, 24, 96., 5, prec, ez.info()["vfac"], rng)
    assert iv == ez.info()["iv"]
    assert 0 < iv < len(cat[0])                       # both kinds present
    taken, _, rest = pc.att_expected(ez, 24, 96., 5, prec, rng)
    assert rest.sum() > 0


def test_att_check_rejects_wrong_stream():
    from ezmock_amd import inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = po.EZmockOracle(Lbox=96., Ngrid=24, rng=0, seed=5, nthread=1, prec="f64")
    ez.eval_growth_params(0.0, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    cat = ez.populate_tracer(0.4, 8., 0.3, 0., 800, att_part=True)
    with pytest.raises(AssertionError):
        pc.check_att(ez, cat, 24, 96., 6, "f64", ez.info()["vfac"])   # another seed's draws
