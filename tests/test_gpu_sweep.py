"""Calibration sweep (64 mocks at 256^3, Lbox 1000 Mpc/h, seeds 1..64, bias
parameters drawn from a fixed meta-seed in rho_c [0.2, 1.2], rho_exp
[5, 20], pdf_base [0.1, 0.5], sigma_v [0, 300], 3e5 tracers each, FP32)
through the request pool, all submitted at once so that several mocks run
concurrently: every result equals its stand-alone mock (the EZmock class,
one mock at a time) -- same tracer count, bit-identical positions,
velocities within 1e-5 relative (double CIC sums rounded once)."""
import numpy as np
import pytest

from ezmock_amd import EZmock
from ezmock_amd.pool import MockPool

pytestmark = pytest.mark.gpu

NMOCK, NG, L, NEXP = 64, 256, 1000.0, 300000


def sweep_params():
    rs = np.random.default_rng(2024)
    return [(rs.uniform(0.2, 1.2), rs.uniform(5, 20), rs.uniform(0.1, 0.5), rs.uniform(0, 300))
            for _ in range(NMOCK)]


def test_sweep_through_pool_equals_standalone(gpu_available, pk_table):
    k, P = pk_table
    par = sweep_params()
    pool = MockPool(k, P, 0.5, 0.31, nslot=4, mem_budget=200e9)
    ids = [pool.submit(NG, L, i + 1, par[i], NEXP, want_catalog=True) for i in range(NMOCK)]
    res = [pool.wait(i) for i in ids]
    pool.close()
    assert all(r["err"] == 0 for r in res), [r.get("errmsg") for r in res if r["err"]]
    spans = sorted((r["t_start"], r["t_end"]) for r in res)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    for i, r in enumerate(res):
        ez = EZmock(Lbox=L, Ngrid=NG, rng=0, seed=i + 1, precision="f32")
        ez.eval_growth_params(0.5, 0.31)
        ez.setup_linear_pk(k, P)
        ez.create_dens_field_from_ic()
        n = ez.populate_tracer_device(*par[i], NEXP)
        cat = ez.device_catalog()
        assert r["ntracer"] == n, i
        for a, b in zip(r["cat"][:3], cat[:3]):
            assert np.array_equal(a, b), i
        vmax = max(np.abs(v).max() for v in cat[3:])
        for a, b in zip(r["cat"][3:], cat[3:]):
            assert np.abs(a.astype(np.float64) - b).max() <= 1e-5 * vmax, i
        del ez
