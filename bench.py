"""Benchmark of the MI355X EZmock engine.

One step = one complete mock of the single-mock drop-in workload
(512^3, Lbox 2500 Mpc/h, z 0.5, MRG32k3a seed 42, bias (0.8, 8, 0.25, 200),
2e-4 (h/Mpc)^3 -> 3,125,000 expected tracers, FP32): create_dens_field_from_ic
followed by populate_tracer with the catalogue kept in device memory, with the
linear P(k) table and cosmology already set (inputs resident on the device
when the timed region starts).  Prints one JSON line on stdout.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]

The roofline entry is for the dominant kernel of the timed region: its
algorithmic bytes per launch (the per-unit model below x the units one launch
processes) over its average launch time, measured with HIP events recorded on
the engine's stream around every launch during the timed steps.
The cpu_baseline entry times the CPU oracle (a C + OpenMP restatement of the
reference path) on this host for one mock of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms/mock at 512^3 & 1024^3 (FP32 vs FP64 GPU), batched mocks/s, p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, Ng, s, ntr):
    """Algorithmic HBM bytes of one launch of a kernel class.  F = s N is
    one real field; the half-complex spectrum is the same size."""
    N = Ng ** 3
    F = s * N
    table = {
        "gen_spectrum": F,                     # write the spectrum once
        "fft_pass_i": 4 * F,                   # read the spectrum, write 3 components
        "fft_pass_i_gen": 3 * F,               # generate the spectrum, write 3 components
        "fft_pass_j": 2 * F,
        "fft_rows_c2r": 2 * F,
        "bias_count": F,
        "bias": 2 * F,                         # read rho, write rhot
        "compact": F,
        "assign_nt": F + 2 * N,                # read rhot, write int16 counts
        "positions": 2 * N + 3 * s * ntr,
        "cic_bin_count": 3 * F,                # read psi x3
        "cic_bin_scatter": 3 * F + 4 * N,      # read psi x3, write indices
        "cic_tiles_rho_mom": 8 * F,            # read psi x3, write rho + {rho, u x3}
        "cic_tiles_rho": 4 * F,
        "cic_tiles_mom": 8 * F,                # read psi x3 + rho, write {rho, u x3}
        "cic_bounds": 3 * F,                   # tile-scan option: read psi x3
        "cic_scan_rho_mom": 8 * F,
        "cic_scan_rho": 4 * F,
        "cic_scan_mom": 8 * F,
        "velocities": (32 + 6) * s * ntr,      # 8 corners x 4 fields + r/w
    }
    return table.get(kernel)


def run_gpu(Ng, Lbox, seed, params, density, precision, steps, warmup, profile=True):
    from ezmock_amd import EZmock, inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = EZmock(Lbox=Lbox, Ngrid=Ng, rng=0, seed=seed, precision=precision)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    nexp = int(round(density * Lbox ** 3))
    ntr = 0
    for _ in range(warmup):
        ez.create_dens_field_from_ic()
        ntr = ez.populate_tracer_device(*params, nexp)
    ez.profile(profile)
    ez.reset_kernel_times()
    ez.sync()
    t0 = time.perf_counter()
    lat = []
    for _ in range(steps):
        t1 = time.perf_counter()
        ez.create_dens_field_from_ic()
        ntr = ez.populate_tracer_device(*params, nexp)
        lat.append(time.perf_counter() - t1)
    ez.sync()
    dt = time.perf_counter() - t0
    kt = ez.kernel_times()
    return dt, ntr, kt, lat


def run_large(Ng=1536, L=3000.0):
    """Large single mock: 1536^3, Lbox 3000 Mpc/h, z 1.0, seed 7, 5e-4 (h/Mpc)^3
    tracers (bias (0.5, 10, 0.3, 150)), FP32; one warm-up and one timed mock,
    with the algorithmic GB/s of the FFT passes, the CIC and the bias kernel."""
    from ezmock_amd import EZmock, inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=7, precision="f32")
    ez.eval_growth_params(1.0, 0.31)
    ez.setup_linear_pk(k, P)
    nexp = int(round(5e-4 * L ** 3))
    ez.create_dens_field_from_ic()
    ntr = ez.populate_tracer_device(0.5, 10.0, 0.3, 150.0, nexp)
    ez.profile(True)
    ez.reset_kernel_times()
    ez.sync()
    t0 = time.perf_counter()
    ez.create_dens_field_from_ic()
    ntr = ez.populate_tracer_device(0.5, 10.0, 0.3, 150.0, nexp)
    ez.sync()
    dt = time.perf_counter() - t0
    kt = ez.kernel_times()
    gbs = {}
    for name in ("fft_pass_i", "fft_pass_i_gen", "fft_pass_j", "fft_rows_c2r", "cic_bin_count",
                 "cic_bin_scatter", "cic_tiles_rho_mom", "bias"):
        if name in kt:
            t, c = kt[name]
            b = algorithmic_bytes(name, Ng, 4, ntr)
            gbs[name] = round(b * c / (t * 1e-3) / 1e9, 1)
    del ez
    return {"ms_per_mock": round(dt * 1e3, 1), "tracers": ntr,
            "alg_GBps": gbs, "peak_GBps": HBM_PEAK_GBS}


def run_sweep(par, Ng, L, graph):
    """The sweep on one instance, one mock after another through
    EZmock_ext_mock: replayed from its captured HIP graphs (graph) or by the
    plain create + populate calls (option 8 off)."""
    from ezmock_amd import EZmock, inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = EZmock(Lbox=L, Ngrid=Ng, rng=0, seed=1, precision="f32")
    if not graph:
        ez.set_option(8, 0)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    for i in range(2):                                      # warm-up (and capture)
        ez.set_seed(1000 + i)
        ez.mock(*par[i], 300000)
    ez.sync()
    t0 = time.perf_counter()
    for i, p in enumerate(par):
        ez.set_seed(i + 1)
        ez.mock(*p, 300000)
    ez.sync()
    dt = time.perf_counter() - t0
    c = ez.path_counts()
    del ez
    return len(par) / dt, c


def run_batch(nmock=64, Ng=256, L=1000.0, nslot=4):
    """Calibration sweep: nmock mocks at Ng^3 with seeds 1..nmock and bias
    parameters drawn from a fixed meta-seed (rho_c in [0.2,1.2], rho_exp in
    [5,20], pdf_base in [0.1,0.5], sigma_v in [0,300]), 3e5 tracers each,
    through the request pool (all submitted at once; plain launches, and
    again with every worker's instance replaying its captured HIP graphs),
    and on one instance with and without the graphs."""
    import numpy as np
    from ezmock_amd import inputs
    from ezmock_amd.pool import MockPool
    k, P = inputs.eh_nowiggle_pk()
    rs = np.random.default_rng(2024)
    par = [(rs.uniform(0.2, 1.2), rs.uniform(5, 20), rs.uniform(0.1, 0.5), rs.uniform(0, 300))
           for _ in range(nmock)]
    pool = MockPool(k, P, 0.5, 0.31, nslot=nslot, mem_budget=200e9)
    for w in [pool.submit(Ng, L, 1000 + i, par[i], 300000) for i in range(nslot)]:
        pool.wait(w)                                        # warm the workers
    t0 = time.perf_counter()
    ids = [pool.submit(Ng, L, i + 1, par[i], 300000) for i in range(nmock)]
    res = [pool.wait(i) for i in ids]
    dt = time.perf_counter() - t0
    pool.close()
    bad = [r["errmsg"] for r in res if r["err"]]
    # one instance, one mock after another: its wall time, and the sum of
    # its kernel times (the rate the pool would reach with no host gaps)
    d1, _, _, _ = run_gpu(Ng, L, 1, par[0], 3e-4, "f32", 8, 2, profile=False)
    _, _, kt1, _ = run_gpu(Ng, L, 1, par[0], 3e-4, "f32", 8, 1, profile=True)
    kms = sum(t for t, c in kt1.values()) / 8
    rg, cg = run_sweep(par, Ng, L, True)
    rp, _ = run_sweep(par, Ng, L, False)
    pool = MockPool(k, P, 0.5, 0.31, nslot=nslot, mem_budget=200e9, options={8: 1})
    for w in [pool.submit(Ng, L, 1000 + i, par[i], 300000) for i in range(nslot)]:
        pool.wait(w)
    t0 = time.perf_counter()
    ids = [pool.submit(Ng, L, i + 1, par[i], 300000) for i in range(nmock)]
    resg = [pool.wait(i) for i in ids]
    dtg = time.perf_counter() - t0
    pool.close()
    bad += [r["errmsg"] for r in resg if r["err"]]
    return {"mocks": nmock, "grid": Ng, "slots": nslot, "seconds": round(dt, 3),
            "mocks_per_s": round(nmock / dt, 2), "errors": bad[:3],
            "pool_graph_mocks_per_s": round(nmock / dtg, 2),
            "single_instance_graph_mocks_per_s": round(rg, 1),
            "single_instance_plain_mocks_per_s": round(rp, 1),
            "single_instance_graph_replays": cg["replay"],
            "single_instance_ms_per_mock": round(d1 * 1e3 / 8, 3),
            "kernel_ms_per_mock": round(kms, 3),
            "kernel_bound_mocks_per_s": round(1e3 / kms, 1)}


def run_serving(nreq=2000, rate=20.0, nslot=4, budget=200e9, seed=7):
    """Serving loop: nreq requests arriving Poisson at `rate` per second,
    grids {128,192,256,384,512}^3 with weights {.3,.25,.2,.15,.1},
    Lbox = 3.9 Ng, bias parameters uniform in the calibration ranges,
    3e-4 (h/Mpc)^3 tracers, admitted by a device-memory budget."""
    import numpy as np
    from ezmock_amd import inputs
    from ezmock_amd.pool import MockPool
    k, P = inputs.eh_nowiggle_pk()
    rs = np.random.default_rng(seed)
    grids = rs.choice([128, 192, 256, 384, 512], size=nreq, p=[.3, .25, .2, .15, .1])
    gaps = rs.exponential(1.0 / rate, size=nreq)
    pool = MockPool(k, P, 0.5, 0.31, nslot=nslot, mem_budget=budget)
    for Ng in (128, 512):                                    # warm the workers
        pool.wait(pool.submit(Ng, 3.9 * Ng, 99, (0.5, 10, 0.3, 150), int(3e-4 * (3.9 * Ng) ** 3)))
    ids = []
    t0 = time.perf_counter()
    tnext = t0
    for i in range(nreq):
        tnext += gaps[i]
        d = tnext - time.perf_counter()
        if d > 0:
            time.sleep(d)
        Ng = int(grids[i])
        L = 3.9 * Ng
        ids.append(pool.submit(Ng, L, 1 + i, (rs.uniform(0.2, 1.2), rs.uniform(5, 20),
                                              rs.uniform(0.1, 0.5), rs.uniform(0, 300)),
                               int(round(3e-4 * L ** 3))))
        if (i + 1) % 100 == 0:
            log(f"  serving: {i + 1}/{nreq} submitted")
    res = [pool.wait(i) for i in ids]
    dt = time.perf_counter() - t0
    pool.close()
    lat = np.array([r["latency"] for r in res]) * 1e3
    ok = [r for r in res if not r["err"]]
    # the slowest requests, split into queue wait and compute time
    slow = [{"Ng": int(grids[i]), "wait_ms": round((res[i]["t_start"] - res[i]["t_submit"]) * 1e3, 1),
             "run_ms": round((res[i]["t_end"] - res[i]["t_start"]) * 1e3, 1)}
            for i in np.argsort(-lat)[:5]]
    return {"requests": nreq, "rate_per_s": rate, "slots": nslot,
            "errors": len(res) - len(ok), "seconds": round(dt, 2),
            "mocks_per_s": round(len(ok) / dt, 2),
            "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
            "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
            "latency_ms_max": round(float(lat.max()), 2), "slowest": slow}


def host_threads():
    """Host cores this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when the environment sets it (a share of a larger host)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_dropin(steps=3, precision="f32"):
    """The drop-in path: create_dens_field_from_ic + EZmock_populate_tracer,
    which returns the catalogue in malloc'd host arrays (device-to-host copy
    included), on the same 512^3 workload."""
    from ezmock_amd import EZmock, inputs
    k, P = inputs.eh_nowiggle_pk()
    ez = EZmock(Lbox=2500., Ngrid=512, rng=0, seed=42, precision=precision)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    ez.create_dens_field_from_ic()
    cat = ez.populate_tracer(0.8, 8., 0.25, 200., 3125000)
    del cat
    ez.sync()
    t0 = time.perf_counter()
    tp = 0.0
    for _ in range(steps):
        ez.create_dens_field_from_ic()
        t1 = time.perf_counter()
        cat = ez.populate_tracer(0.8, 8., 0.25, 200., 3125000)
        tp += time.perf_counter() - t1
        n = len(cat[0])
        del cat
    dt = time.perf_counter() - t0
    return {"ms_per_mock": round(dt * 1e3 / steps, 2),
            "populate_tracer_ms": round(tp * 1e3 / steps, 2), "tracers": n,
            "host_bytes": n * 6 * (4 if precision == "f32" else 8)}


def cpu_baseline(precision):
    """The oracle on the host cores, one mock of the same workload."""
    from oracle import pyoracle as po
    from ezmock_amd import inputs
    threads = host_threads()
    k, P = inputs.eh_nowiggle_pk()
    Ng, L = 512, 2500.
    ez = po.EZmockOracle(Lbox=L, Ngrid=Ng, rng=0, seed=42, nthread=threads, prec=precision)
    ez.eval_growth_params(0.5, 0.31)
    ez.setup_linear_pk(k, P)
    t0 = time.perf_counter()
    ez.create_dens_field_from_ic()
    ez.populate_tracer(0.8, 8., 0.25, 200., int(round(2e-4 * L ** 3)))
    dt = time.perf_counter() - t0
    stages = {k_: round(v, 2) for k_, v in ez.stage_times().items()}
    ez.close()
    return {"value": round(dt * 1e3, 1), "unit": "ms/mock", "cores": threads,
            "kind": "port",
            "sample": f"one whole mock of the same workload (512^3, {precision}, not scaled) "
                      f"by the oracle, a C+OpenMP restatement of the reference CPU path with "
                      f"its own FFT, on {threads} OpenMP threads ({cpu_model()}, "
                      f"{os.cpu_count()} logical CPUs on the host, {threads} available to this "
                      f"process); stage seconds {stages}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="f32", choices=["f32", "f64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the secondary lines (1024^3 FP32, 512^3 FP64, batch, serving)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the serving trace (Poisson, 20 per second)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs one mock per GPU on a single GPU; runs on more "
              "than one GPU are not part of this build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    Ng, Lbox, seed = 512, 2500., 42
    params = (0.8, 8., 0.25, 200.)
    density = 2e-4
    s = 4 if args.precision == "f32" else 8
    log(f"bench: 512^3 {args.precision}, warmup {args.warmup}, steps {args.steps}")
    # the timed run without the per-kernel timers (which put the three FFT
    # components on one stream and add an event pair per kernel); then the
    # same steps with the timers, for the kernel breakdown and the roofline
    dt, ntr, _, lat = run_gpu(Ng, Lbox, seed, params, density, args.precision,
                              args.steps, args.warmup, profile=False)
    ms = dt * 1e3 / args.steps
    log(f"bench: {ms:.2f} ms/mock, {ntr} tracers")
    dtp, _, kt, _ = run_gpu(Ng, Lbox, seed, params, density, args.precision,
                            args.steps, 1, profile=True)
    log(f"bench: with per-kernel timers {dtp * 1e3 / args.steps:.2f} ms/mock")
    for name, (t, c) in sorted(kt.items(), key=lambda x: -x[1][0]):
        log(f"  {name:16s} {t / args.steps:9.3f} ms/mock  ({c} launches)")

    # dominant kernel with a byte model
    dom = max((x for x in kt.items() if algorithmic_bytes(x[0], Ng, s, ntr)),
              key=lambda x: x[1][0])
    name, (tot_ms, cnt) = dom
    avg_s = tot_ms / cnt * 1e-3
    bytes_ = algorithmic_bytes(name, Ng, s, ntr)
    achieved = bytes_ / avg_s / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get(f"{name}_{Ng}_{args.precision}")
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "algorithmic_bytes": bytes_,
            "avg_launch_ms": round(avg_s * 1e3, 4)}
    total_alg = sum((algorithmic_bytes(n, Ng, s, ntr) or 0) * c for n, (t, c) in kt.items())
    kernel_ms = sum(t for t, c in kt.values())
    lat_ms = sorted(x * 1e3 for x in lat)
    extra = {
        "mock_latency_ms_p50": round(lat_ms[len(lat_ms) // 2], 2),
        "mock_latency_ms_max": round(lat_ms[-1], 2),
        "mocks_per_s": round(1e3 / ms, 3),
        "tracers": ntr,
        "all_kernels_alg_GBps": round(total_alg / (kernel_ms * 1e-3) / 1e9, 1),
        "kernel_ms_per_mock": {n: round(t / args.steps, 3) for n, (t, c) in kt.items()},
    }
    if not args.no_extra:
        # the pool runs before the large single mocks, so that the serving
        # latencies do not follow the release of a 150 GB instance
        try:
            extra["batch_64x256"] = run_batch()
            log(f"bench: batch {extra['batch_64x256']}")
        except Exception as e:
            extra["batch_64x256"] = f"failed: {e}"
        try:
            extra["dropin_512_f32"] = run_dropin()
            log(f"bench: drop-in {extra['dropin_512_f32']}")
        except Exception as e:
            extra["dropin_512_f32"] = f"failed: {e}"
        try:
            extra["serving"] = run_serving(nreq=args.serve_requests)
            log(f"bench: serving {extra['serving']}")
        except Exception as e:
            extra["serving"] = f"failed: {e}"
        try:
            d2, n2, _, _ = run_gpu(1024, 2500., 42, params, density, "f32", 2, 1, profile=False)
            extra["ms_per_mock_1024_f32"] = round(d2 * 1e3 / 2, 2)
            log(f"bench: 1024^3 f32 {extra['ms_per_mock_1024_f32']} ms/mock")
            # per-kernel times of one more mock with the timers on
            _, _, kt2, _ = run_gpu(1024, 2500., 42, params, density, "f32", 1, 1, profile=True)
            extra["kernel_ms_1024_f32"] = {n: round(t, 3) for n, (t, c) in kt2.items()}
            log(f"bench: 1024^3 kernels {extra['kernel_ms_1024_f32']}")
        except Exception as e:          # report, do not hide
            extra["ms_per_mock_1024_f32"] = f"failed: {e}"
        try:
            d3, n3, _, _ = run_gpu(512, 2500., 42, params, density, "f64", 2, 1, profile=False)
            extra["ms_per_mock_512_f64"] = round(d3 * 1e3 / 2, 2)
            log(f"bench: 512^3 f64 {extra['ms_per_mock_512_f64']} ms/mock")
        except Exception as e:
            extra["ms_per_mock_512_f64"] = f"failed: {e}"
        try:
            extra["large_1536_f32"] = run_large()
            log(f"bench: 1536^3 {extra['large_1536_f32']}")
        except Exception as e:
            extra["large_1536_f32"] = f"failed: {e}"
    try:                                # the autotuner's picks (tk 0 = default)
        from ezmock_amd.EZmock import fft_schedule
        extra["fft_schedules"] = {str(n): fft_schedule(n, args.precision)
                                  for n in (256, 384, 512, 1024, 1536)}
    except Exception as e:
        extra["fft_schedules"] = f"failed: {e}"
    cpu = None
    if not args.no_cpu_baseline:
        log("bench: timing the CPU oracle sample ...")
        cpu = cpu_baseline(args.precision)
        log(f"bench: cpu baseline {cpu['value']} ms/mock on {cpu['cores']} threads")
    line = {
        "metric": METRIC, "value": round(ms, 3), "unit": "ms/mock", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision, "data": "synthetic (EH no-wiggle linear P(k), seeded noise)",
        "config": {"workload": "single mock 512^3, Lbox 2500 Mpc/h, z 0.5, MRG32k3a seed 42, "
                               "bias (0.8, 8, 0.25, 200), 2e-4 (h/Mpc)^3",
                   "Ng": Ng, "Lbox": Lbox, "nexp": int(round(density * Lbox ** 3)),
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
