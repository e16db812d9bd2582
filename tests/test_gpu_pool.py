"""Request pool (continuous batching): results are those of stand-alone mocks.

* A request computed in a mix of concurrent requests of several grid sizes
  gives the same catalogue as the same mock run alone through ``EZmock``:
  identical tracer count and positions (bit-exact), velocities within 1e-5
  relative (FP32; the CIC sums are order-independent up to the rounding of
  their double accumulation), and the same P0/P2 within 1e-4 relative.
* Admission by memory budget: with a budget that holds one 256^3 request,
  the 256^3 requests run one after another.
* A failing request reports its error code and the pool keeps serving.
"""
import numpy as np
import pytest

from ezmock_amd import EZmock, inputs
from ezmock_amd.pool import MockPool

pytestmark = pytest.mark.gpu

Z, OM = 0.5, 0.31


def _alone(pk, Ng, L, seed, params, nexp, rsd, dk, nbin):
    k, P = pk
    ez = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=seed, precision="f32")
    ez.eval_growth_params(Z, OM)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    ez.populate_tracer_device(*params, nexp)
    cat = ez.device_catalog()
    pkr = ez.powspec(rsd_fac=rsd, interlace=True, dk=dk, nbin=nbin)
    return cat, pkr


@pytest.mark.parametrize("graphs", [False, True])
def test_pool_result_same_alone_and_in_mix(gpu_available, pk_table, graphs):
    """Plain launches (the pool's default) and every instance replaying its
    captured HIP graphs (pool option 8 = 1)."""
    k, P = pk_table
    target = dict(Ngrid=256, Lbox=1000.0, seed=7, params=(0.5, 10.0, 0.3, 150.0),
                  nexp=300000, rsd_fac=0.011384, pk_dk=0.01, pk_nbin=40)
    cat0, pk0 = _alone(pk_table, 256, 1000.0, 7, target["params"], 300000,
                       0.011384, 0.01, 40)
    pool = MockPool(k, P, Z, OM, nslot=4, mem_budget=100e9,
                    options={8: 1} if graphs else None)
    rs = np.random.default_rng(3)
    ids = []
    for i in range(12):
        Ng = int(rs.choice([128, 192, 256, 384]))
        ids.append(pool.submit(Ng, 3.9 * Ng, 100 + i,
                               (rs.uniform(0.2, 1.2), rs.uniform(5, 20), rs.uniform(0.1, 0.5),
                                rs.uniform(0, 300)), int(3e-4 * (3.9 * Ng) ** 3)))
        if i == 5:
            tid = pool.submit(want_catalog=True, **target)
    res = {i: pool.wait(i) for i in ids}
    r = pool.wait(tid)
    pool.close()
    assert all(v["err"] == 0 for v in res.values()), [v.get("errmsg") for v in res.values()]
    assert r["err"] == 0, r.get("errmsg")
    # the mix really overlapped
    spans = sorted((v["t_start"], v["t_end"]) for v in list(res.values()) + [r])
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    cat = r["cat"]
    assert r["ntracer"] == len(cat0[0])
    for a, b in zip(cat[:3], cat0[:3]):
        assert np.array_equal(a, b)
    for a, b in zip(cat[3:], cat0[3:]):
        assert np.all(np.abs(a - b) <= 1e-5 * (np.abs(b) + 1))
    for name, ref in zip(("k", "P0", "P2", "nmodes"), pk0):
        assert np.allclose(r[name], ref, rtol=1e-4, atol=1e-4 * np.abs(pk0[1]).max()), name


def test_pool_admission_by_budget(gpu_available, pk_table):
    k, P = pk_table
    probe = MockPool(k, P, Z, OM, nslot=1, mem_budget=1e12)
    one = probe.footprint(256)
    probe.close()
    pool = MockPool(k, P, Z, OM, nslot=3, mem_budget=int(1.5 * one))
    ids = [pool.submit(256, 1000.0, 11 + i, (0.5, 10.0, 0.3, 150.0), 300000) for i in range(4)]
    res = sorted((pool.wait(i) for i in ids), key=lambda v: v["t_start"])
    pool.close()
    assert all(v["err"] == 0 for v in res)
    for a, b in zip(res, res[1:]):
        assert b["t_start"] >= a["t_end"] - 1e-6


def test_pool_reports_errors_and_keeps_serving(gpu_available, pk_table):
    k, P = pk_table
    pool = MockPool(k, P, Z, OM, nslot=2, mem_budget=50e9)
    bad = pool.submit(128, 500.0, 5, (0.5, 10.0, 1.5, 150.0), 10000)   # pdf_base >= 1
    good = pool.submit(128, 500.0, 5, (0.5, 10.0, 0.3, 150.0), 10000)
    rb, rg = pool.wait(bad), pool.wait(good)
    pool.close()
    assert rb["err"] != 0 and "base" in rb["errmsg"].lower()
    assert rg["err"] == 0 and rg["ntracer"] > 0
    with pytest.raises(RuntimeError):
        MockPool(k, P, Z, OM, nslot=1, mem_budget=1e6).submit(256, 1000.0, 1, (0.5, 10, 0.3, 150), 1000)
