"""FFT schedules (EZmock_ext_fft_schedule / _set_schedule) and the
autotuner: every candidate schedule of the strided and first passes (column
tile widths 4, 8, 16 or the default; the first pass fused or split into
three single-component passes) gives bit-identical displacement fields, and
the first field of a grid size with autotuning on records a tuned schedule
that reproduces the default fields bit for bit."""
import itertools

import numpy as np
import pytest

from ezmock_amd import EZmock
from ezmock_amd.EZmock import fft_check, fft_schedule, fft_set_schedule

pytestmark = pytest.mark.gpu


def fields(pk, Ng, prec, tune=False):
    k, P = pk
    g = EZmock(Lbox=4.0 * Ng, Ngrid=Ng, rng=0, seed=17, precision=prec)
    g.set_option(6, int(tune))
    g.eval_growth_params(0.5, 0.31)
    g.setup_linear_pk(k, P)
    g.create_dens_field_from_ic()
    return [g.field(f) for f in ("psi_x", "psi_y", "psi_z")]


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_every_schedule_gives_the_same_fields(gpu_available, pk_table, prec):
    Ng = 256
    fft_set_schedule(Ng, 0, 0, 0, prec)
    ref = fields(pk_table, Ng, prec)
    for tk_j, tk_i, split in itertools.product((0, 4, 8, 16), (0, 4, 8, 16), (0, 1)):
        fft_set_schedule(Ng, tk_j, tk_i, split, prec)
        for a, b in zip(fields(pk_table, Ng, prec), ref):
            assert np.array_equal(a, b), (tk_j, tk_i, split)
    fft_set_schedule(Ng, 0, 0, 0, prec)


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_autotune_records_a_schedule(gpu_available, pk_table, prec):
    Ng = 384
    a = fields(pk_table, Ng, prec, tune=True)
    sc = fft_schedule(Ng, prec)
    print(f"384^3 {prec}: tuned schedule {sc}")
    assert sc["tuned"] and sc["split"] in (0, 1)
    assert sc["tk_j"] in (0, 4, 8, 16) and sc["tk_i"] in (0, 4, 8, 16)
    fft_set_schedule(Ng, 0, 0, 0, prec)
    for x, y in zip(a, fields(pk_table, Ng, prec)):
        assert np.array_equal(x, y)
    fft_set_schedule(Ng, sc["tk_j"], sc["tk_i"], sc["split"], prec)


@pytest.mark.parametrize("Ng", [1024, 1536])
def test_large_grid_schedules_plane_waves(gpu_available, Ng):
    """The production first pass under each of its schedules at the large
    sizes, against the plane-wave closed form."""
    h = Ng // 2
    modes = [(3, 5, 7), (Ng - 2, 1, h - 1), (h, Ng - 7, 1)]
    amps = [1.0 + 0.5j, -0.75 + 0.25j, 0.5 - 1.0j]
    keep = fft_schedule(Ng)
    for tk_i, split in itertools.product((0, 4, 8), (0, 1)):
        fft_set_schedule(Ng, 0, tk_i, split)
        err, ref = fft_check(Ng, 1, modes, amps)
        assert np.all(err <= 2e-6 * np.log2(Ng) * ref), (tk_i, split, err / ref)
    fft_set_schedule(Ng, keep["tk_j"], keep["tk_i"], keep["split"])
