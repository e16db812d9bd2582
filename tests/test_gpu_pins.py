"""The closed-form tracer pins of tests/pincheck.py on the GPU engine,
through its C ABI: bias transform and Box-Muller pairing
(pop_tracer.c:70-79, 245-289), positions and their stream offsets
(:919, 946-985), the mass-weighted velocity paint and gather (:593-753) and
the sigma_v scatter order (:1071-1094).  These hold independently of the
oracle, so the GPU's agreement with the oracle is not the only evidence."""
import pytest

import pincheck as pc
from ezmock_amd import EZmock

pytestmark = pytest.mark.gpu


def engine(L, Ng, seed, prec):
    return EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=seed, precision=prec)


def ic_field(pk, L, Ng, seed, prec):
    k, P = pk
    ez = engine(L, Ng, seed, prec)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    return ez


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("seed", [3, 77])
def test_bias_box_muller_pairs(gpu_available, prec, seed):
    pc.check_bias(engine(160., 16, seed, prec), 16, seed, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_bias_pairs_large_grid(gpu_available, prec):
    """64^3: 262144 valid cells, pairs across many workgroups."""
    pc.check_bias(engine(640., 64, 21, prec), 64, 21, prec, nexp=20000)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_positions_from_stream_offsets(gpu_available, pk_table, prec):
    ez = ic_field(pk_table, 300., 24, 5, prec)
    cat = ez.populate_tracer(0.4, 8., 0.3, 100., 600)
    pc.check_positions(ez, cat, 24, 300., 5, prec)
    ez = ic_field(pk_table, 640., 64, 6, prec)
    cat = ez.populate_tracer(0.4, 8., 0.3, 100., 20000)
    pc.check_positions(ez, cat, 64, 640., 6, prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_sigma_v_draw_order(gpu_available, pk_table, prec):
    ez = ic_field(pk_table, 300., 24, 9, prec)
    assert pc.check_sigma_v(ez, 24, 9, prec) > 100
    ez = ic_field(pk_table, 640., 64, 10, prec)
    assert pc.check_sigma_v(ez, 64, 10, prec, nexp=20001) > 10000


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_velocity_uniform_displacement(gpu_available, prec):
    pc.check_velocity_uniform(engine(200., 16, 4, prec), 16, 200., prec)
    pc.check_velocity_uniform(engine(640., 64, 4, prec), 64, 640., prec)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_velocity_mass_weighted_gather(gpu_available, prec):
    assert pc.check_velocity_zshear(engine(240., 16, 4, prec), 16, 240., prec) > 100
    assert pc.check_velocity_zshear(engine(640., 64, 4, prec), 64, 640., prec) > 1000
