#!/usr/bin/env python3
"""Benchmark: local WLS solves/s of GWR at n = 100,000, k = 8 on one MI355X.

Workload (one step): GWRBasic on n points (lon ~ U[-10,30], lat ~ U[35,60]),
k = 8 (intercept + 7 equicorrelated covariates), great-circle distance,
Gaussian kernel with a fixed bandwidth selected by an AICc golden-section
search over [0.5, 20] degrees of arc (55.66 .. 2226.39 km), then the final fit
with hat-matrix diagnostics (betas, SE, tr S, tr S^T S, qDiag).  A local solve
is one (focus point, bandwidth) pair; every criterion evaluation and the fit
count n solves each.

The timed path is the mixed-precision one (FP32 weights, FP16 hi+lo split
contraction on the matrix cores, FP64 partial sums and solves, certified
golden-section decisions); the same step is run on the FP64 path for the
mixed/FP64 ratio and the beta error.  Inputs are resident on the device
before the timed region starts.

`value` counts each (focus point, bandwidth) evaluation of the search and the
fit once; the mixed search takes the FP64 search's iterates, so both paths
count the same solves per step and mixed_over_fp64 is their wall-time ratio.

More workloads ride along in the same JSON line:
- "cfg1": n = 2000 uniform square, k = 4, bisquare adaptive CV search + fit,
  GPU FP64 / mixed and the oracle's full OpenMP run (iterates compared);
- "refbench": the reference's own benchmark shape (a fit at n = 50,000,
  k = 3, Gaussian bandwidth 0.2), GPU and a sampled oracle fit;
- "mgwr_cpu_parity": GWRMultiscale on the multiscale-cosines workload at
  n = 1000, the oracle's backfit against the GPU's (bandwidths bit-exact);
- "mgwr": GWRMultiscale backfitting at n = 50,000, k = 6 (bisquare adaptive
  bandwidths chosen by CV for every parameter, CVR threshold 1e-5) in mixed
  precision and FP64, with the adaptive select's rate;
- "mgwr_hat": the same with the hat matrix at n = 4000, k = 3;
- "batch": 512 uneven problems (n log-uniform in [200, 20000], k in 3..12,
  Gaussian / bisquare / tricube, adaptive or fixed), each a CV golden-section
  search plus fit, through the batching server against a loop of
  single-problem calls;
- "serve": a Poisson stream of 500 fit and CV-search requests (n log-uniform
  in [500, 200000], k in 2..16, clustered coordinates) replayed in real time
  against the continuous-batching server at two rates, with latency
  percentiles and the beta error on requests spread over all sizes;
- "predict", "varsel": prediction at n new locations and variable selection.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "local WLS solves/s @n=100k,k=8 mixed vs FP64 (β err≤1e-6); serve p99 ms"
PEAK = {"mixed": 2.5e15, "fp64": 78.6e12}   # dense FP16 MFMA / FP64 MFMA, FLOP/s
HBM_PEAK = 8.0e12


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def jsonable(o):
    """numpy scalars and arrays in the JSON line."""
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError(type(o).__name__)


def leg(name, fn, *a):
    """One side workload of the JSON line; an exception ends that leg only."""
    try:
        return fn(*a)
    except Exception as e:   # noqa: BLE001 -- reported in the line, the headline stands
        import traceback
        traceback.print_exc()
        log("%s failed: %r" % (name, e))
        return {"error": repr(e)}


def beta_rel_err(b, ref):
    """max over columns of max|b - ref| / max|ref| (column-scaled)."""
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    return float(np.max(np.max(np.abs(b - ref), axis=0) / scale))


def run_steps(ctx, capi, prec, steps, lo, hi):
    out = None
    for _ in range(steps):
        out = ctx.gwr_run(capi.GAUSSIAN, False, 0.0, select_bw=True, crit=capi.CRIT_AIC, precision=prec,
                          lower=lo, upper=hi)
    return out


def cpu_threads():
    """Host threads of the CPU legs: every core of this process's affinity
    mask, capped by OMP_NUM_THREADS when it is set (the host share one GPU
    gets where several GPUs share a host)."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    return (min(aff, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else aff), aff


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(d, bw, sample, threads):
    from oracle import pyoracle
    t0 = time.perf_counter()
    pyoracle.criterion(d["coords"], d["X"], d["y"], pyoracle.GAUSSIAN, False, bw, pyoracle.CRIT_AIC, geographic=True,
                       nthreads=threads, focus_count=sample)
    dt = time.perf_counter() - t0
    return sample / dt, dt


def col_rel(b, ref):
    """(column-scaled, element-wise) relative errors of b against ref; the
    element-wise one over entries with |ref| >= 1e-3 max|ref| of their column
    (smaller coefficients carry the column's absolute error)."""
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    cs = float(np.max(np.max(np.abs(b - ref), axis=0) / scale))
    big = np.abs(ref) >= 1e-3 * scale
    el = float(np.max(np.abs(b - ref)[big] / np.abs(ref)[big])) if big.any() else 0.0
    return cs, el


def run_cfg1(capi, synth, threads):
    """n = 2000 uniform square, k = 4, Euclidean bisquare adaptive CV search over
    [20, 2000] neighbours + fit with hat diagnostics: GPU FP64 and mixed, and
    the oracle's full OpenMP run on the host."""
    from oracle import pyoracle
    d = synth.uniform_square(n=2000)
    n = d["X"].shape[0]
    out = {"n": n, "k": d["X"].shape[1]}
    res = {}
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"])
        for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
            ctx.gwr_run(capi.BISQUARE, True, 0, select_bw=True, crit=capi.CRIT_CV, precision=prec, lower=20,
                        upper=2000)   # warm (the search's selects are cached from here on: cleared below)
            ctx.set_data(d["coords"], d["X"], d["y"])
            ctx.reset_timers()
            t0 = time.perf_counter()
            res[name] = ctx.gwr_run(capi.BISQUARE, True, 0, select_bw=True, crit=capi.CRIT_CV, precision=prec,
                                    lower=20, upper=2000)
            dt = time.perf_counter() - t0
            out[name + "_seconds"] = dt
            out[name + "_solves_per_s"] = ctx.solve_count() / dt
    log("cfg1: oracle on %d threads" % threads)
    t0 = time.perf_counter()
    o = pyoracle.gwr_run(d["coords"], d["X"], d["y"], pyoracle.BISQUARE, True, 0, select_bw=True,
                         crit=pyoracle.CRIT_CV, lower=20, upper=2000, nthreads=threads)
    dt = time.perf_counter() - t0
    evals = len(o["trace"])
    out["cpu_seconds"] = dt
    out["cpu_solves_per_s"] = (evals + 1) * n / dt
    out["cpu_threads"] = threads
    out["bandwidth"] = float(o["bandwidth"])
    out["iterates_bitexact_fp64"] = [x for x, _ in res["fp64"]["trace"]] == [x for x, _ in o["trace"]]
    out["iterates_bitexact_mixed"] = [x for x, _ in res["mixed"]["trace"]] == [x for x, _ in o["trace"]]
    out["bandwidth_identical"] = res["fp64"]["bandwidth"] == o["bandwidth"] == res["mixed"]["bandwidth"]
    errs = {}
    for q in ("betas", "betasSE", "qDiag"):
        errs[q] = max(col_rel(res["fp64"][q].reshape(n, -1), o[q].reshape(n, -1)))
    errs["diagnostic"] = max(abs(res["fp64"]["diagnostic"][key] - o["diagnostic"][key]) / abs(o["diagnostic"][key])
                             for key in o["diagnostic"])
    out["fp64_max_rel_err_vs_oracle"] = errs
    out["mixed_beta_rel_err_vs_oracle"] = col_rel(res["mixed"]["betas"], o["betas"])
    return out


def run_refbench(capi, synth, threads, sample):
    """The reference's own benchmark shape (test/testGWRBasic.cpp:380-419): a
    GWRBasic fit with hat diagnostics at n = 50,000, k = 3, Euclidean Gaussian
    fixed bandwidth 0.2; the oracle's fit on a sample of focus points."""
    from oracle import pyoracle
    d = synth.reference_benchmark()
    n = d["X"].shape[0]
    out = {"n": n, "k": 3, "bandwidth": 0.2}
    res = {}
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"])
        for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
            ctx.gwr_run(capi.GAUSSIAN, False, 0.2, precision=prec)   # warm
            t0 = time.perf_counter()
            res[name] = ctx.gwr_run(capi.GAUSSIAN, False, 0.2, precision=prec)
            dt = time.perf_counter() - t0
            out[name + "_seconds"] = dt
            out[name + "_solves_per_s"] = n / dt
    log("refbench: oracle fit of %d focus points on %d threads" % (sample, threads))
    t0 = time.perf_counter()
    o = pyoracle.fit_core(d["coords"], d["X"], d["y"], pyoracle.GAUSSIAN, False, 0.2, nthreads=threads,
                          focus_count=sample)
    dt = time.perf_counter() - t0
    out["cpu_solves_per_s"] = sample / dt
    out["cpu_sample"] = sample
    out["cpu_threads"] = threads
    out["fp64_beta_rel_err_vs_oracle_sample"] = col_rel(res["fp64"]["betas"][:sample], o["betas"][:sample])
    out["mixed_beta_rel_err_vs_fp64"] = col_rel(res["mixed"]["betas"], res["fp64"]["betas"])
    out["mixed_over_fp64"] = out["fp64_seconds"] / out["mixed_seconds"]
    return out


def run_mgwr_cpu_parity(capi, synth, n, threads):
    """Multiscale cosines at reduced n (the point density of n = 50,000): the
    oracle's backfit on the host against the GPU's FP64 and mixed runs."""
    from oracle import pyoracle
    d = synth.multiscale_cosines(n=n, extent=1000.0 * np.sqrt(n / 50_000))
    k = d["X"].shape[1]
    args = ([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k, [0] + [1] * (k - 1), [1e-5] * k)
    out = {"n": n, "k": k}
    res = {}
    with capi.Context(0) as ctx:
        for name, prec in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
            ctx.set_data(d["coords"], d["X"], d["y"])
            t0 = time.perf_counter()
            res[name] = ctx.mgwr_run(*args, crit_type=0, crit_threshold=1e-5, precision=prec)
            out[name + "_seconds"] = time.perf_counter() - t0
    log("mgwr parity: oracle n=%d on %d threads" % (n, threads))
    t0 = time.perf_counter()
    o = pyoracle.mgwr_run(d["coords"], d["X"], d["y"], *args, has_hat=False, crit_type=0, crit_threshold=1e-5,
                          nthreads=threads)
    out["cpu_seconds"] = time.perf_counter() - t0
    out["cpu_threads"] = threads
    out["bandwidths"] = [float(b) for b in o["bandwidths"]]
    out["bandwidths_bitexact_fp64"] = bool(np.array_equal(res["fp64"]["bandwidths"], o["bandwidths"]))
    out["bandwidths_bitexact_mixed"] = bool(np.array_equal(res["mixed"]["bandwidths"], o["bandwidths"]))
    out["fp64_beta_rel_err_vs_oracle"] = col_rel(res["fp64"]["betas"], o["betas"])
    out["mixed_beta_rel_err_vs_oracle"] = col_rel(res["mixed"]["betas"], o["betas"])
    return out


def run_batch(capi, synth, num, loop):
    """Uneven batch: every problem is a CV search + fit, mixed precision."""
    probs = synth.uneven_batch(num)
    sum_n = sum(p["X"].shape[0] for p in probs)
    log("batch: %d problems, sum n %d" % (num, sum_n))
    with capi.Server(0, capi.MIXED, max_inflight=num) as srv:
        t0 = time.perf_counter()
        ids = [srv.submit(p["coords"], p["X"], p["y"], p["kernel"], p["adaptive"], 0.0, select_bw=True,
                          crit=capi.CRIT_CV) for p in probs]
        res = [srv.wait(i) for i in ids]
        dt = time.perf_counter() - t0
        st = srv.stats()
        tuned = srv.autotune_report()
    bad = sum(r["status"] != capi.GWM_OK for r in res)
    solves = sum(r["solves"] for r in res)
    out = {"problems": num, "sum_n": sum_n, "solves": solves, "seconds": dt, "solves_per_s": solves / dt,
           "rounds": st["rounds"], "device_busy_frac": st["busy_ms"] * 1e-3 / dt,
           "peak_device_bytes": st["peak_bytes"], "failed": bad, "server_autotune": tuned}
    if loop:
        log("batch: loop of single-problem calls over the same %d problems" % num)
        ctx = capi.Context(0)
        lsolves = 0
        t0 = time.perf_counter()
        lres = []
        for p in probs:
            ctx.set_data(p["coords"], p["X"], p["y"])
            ctx.reset_timers()
            g = ctx.gwr_run(p["kernel"], p["adaptive"], 0.0, select_bw=True, crit=capi.CRIT_CV,
                            precision=capi.MIXED)
            lsolves += ctx.solve_count()
            lres.append(g["bandwidth"])
        ldt = time.perf_counter() - t0
        ctx.close()
        out["loop_seconds"] = ldt
        out["loop_solves_per_s"] = lsolves / ldt
        out["batch_over_loop"] = (solves / dt) / (lsolves / ldt)
        out["bandwidths_equal_loop"] = int(sum(a == r["bandwidth"] for a, r in zip(lres, res)))
    return out


def run_mgwr(capi, synth, n):
    """Multiscale-cosines backfitting, mixed precision and FP64."""
    d = synth.multiscale_cosines(n=n)
    k = d["X"].shape[1]
    out = {"n": n, "k": k}
    res = {}
    with capi.Context(0) as ctx:
        for name, prec in (("mixed", capi.MIXED), ("fp64", capi.FP64)):
            log("mgwr %s n=%d" % (name, n))
            # fresh problem state for each run: no adaptive-bandwidth vectors
            # cached by the other precision's run
            ctx.set_data(d["coords"], d["X"], d["y"])
            ctx.reset_timers()
            t0 = time.perf_counter()
            r = ctx.mgwr_run([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k,
                             [0] + [1] * (k - 1), [1e-5] * k, crit_type=0, crit_threshold=1e-5, precision=prec)
            dt = time.perf_counter() - t0
            res[name] = r
            out[name + "_seconds"] = dt
            out[name + "_solves_per_s"] = ctx.solve_count() / dt
            out[name + "_iterations"] = r["iterations"]
        # a third, kernel-timed run (HIP events synchronise after every phase,
        # so the two runs above are untimed): the adaptive select's share
        log("mgwr mixed n=%d, kernel timers on" % n)
        ctx.set_data(d["coords"], d["X"], d["y"])
        ctx.reset_timers()
        ctx.set_timing(True)
        ctx.mgwr_run([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k,
                     [0] + [1] * (k - 1), [1e-5] * k, crit_type=0, crit_threshold=1e-5, precision=capi.MIXED)
        ctx.set_timing(False)
        sel, sel_ms = ctx.kernel_time("select")
        look, look_ms = ctx.kernel_time("select_lookup")
        tab, tab_ms = ctx.kernel_time("nbr_table")
    # one select finds the thresholds of n focus points among n data points:
    # n^2 element visits in the reference's sort-based algorithm.  Bandwidths
    # below the neighbour table's K are read from it instead (one build).
    out["selects"] = sel
    out["select_ms"] = sel_ms
    out["select_elements_per_s"] = sel * float(n) * n / (sel_ms * 1e-3) if sel_ms > 0 else None
    out["select_lookups"] = look
    out["select_lookup_ms"] = look_ms
    out["neighbour_table_builds"] = tab
    out["neighbour_table_ms"] = tab_ms
    out["bandwidths"] = [float(b) for b in res["mixed"]["bandwidths"]]
    out["bandwidths_identical"] = bool(np.array_equal(res["mixed"]["bandwidths"], res["fp64"]["bandwidths"]))
    out["beta_rel_err_vs_fp64"] = beta_rel_err(res["mixed"]["betas"], res["fp64"]["betas"])
    out["beta_elementwise_rel_err_vs_fp64"] = col_rel(res["mixed"]["betas"], res["fp64"]["betas"])[1]
    out["aicc_rel_err_vs_fp64"] = abs(res["mixed"]["diagnostic"]["AICc"] - res["fp64"]["diagnostic"]["AICc"]) / \
        abs(res["fp64"]["diagnostic"]["AICc"])
    return out


def run_mgwr_hat(capi, synth, n):
    """Multiscale backfitting with the hat matrix (S0 and one n x n operator
    per variable on the device, an FP64 GEMM per backfitting step)."""
    d = synth.multiscale_cosines(n=n)
    X = d["X"][:, :3]
    k = X.shape[1]
    out = {"n": n, "k": k}
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], X, d["y"])
        ctx.mgwr_set_hat_matrix(True)
        log("mgwr with hat matrix n=%d k=%d" % (n, k))
        ctx.reset_timers()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        r = ctx.mgwr_run([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k,
                         [0] + [1] * (k - 1), [1e-5] * k, crit_type=0, crit_threshold=1e-5, precision=capi.MIXED)
        dt = time.perf_counter() - t0
        ctx.set_timing(False)
        steps, ms = ctx.kernel_time("mgwr_hat")
        shat = ctx.mgwr_last_shat()
    out["seconds"] = dt
    out["iterations"] = r["iterations"]
    out["hat_steps"] = steps
    out["hat_ms_per_step"] = ms / max(steps, 1)
    # one step: S_i (n^2 weights), M = R_i - S0, R_i = S_i M + S_i (2 n^3 flops), S0 = R_i - M
    out["hat_step_tflops"] = 2.0 * float(n) ** 3 / (out["hat_ms_per_step"] * 1e-3) / 1e12 if steps else None
    out["hat_step_frac_of_fp64_peak"] = out["hat_step_tflops"] * 1e12 / PEAK["fp64"] if steps else None
    out["trS"], out["trStS"] = float(shat[0]), float(shat[1])
    out["AICc"] = r["diagnostic"]["AICc"]
    out["bandwidths"] = [float(b) for b in r["bandwidths"]]
    return out


def run_predict(ctx, capi, d, bw, m, seed=2024):
    """Betas at m new locations of the sphere-box workload at the selected bandwidth."""
    rng = np.random.default_rng(seed)
    loc = np.column_stack([rng.uniform(-10, 30, m), rng.uniform(35, 60, m)])
    out = {"m": m, "bandwidth_km": bw}
    res = {}
    for name, prec in (("mixed", capi.MIXED), ("fp64", capi.FP64)):
        ctx.predict(loc[:1024], capi.GAUSSIAN, False, bw, precision=prec)   # warm
        log("predict %s m=%d" % (name, m))
        ctx.reset_timers()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        res[name] = ctx.predict(loc, capi.GAUSSIAN, False, bw, precision=prec)
        dt = time.perf_counter() - t0
        ctx.set_timing(False)
        _, kms = ctx.kernel_time("contract_%s_predict" % name)
        out[name + "_seconds"] = dt
        out[name + "_solves_per_s"] = m / dt
        out[name + "_sweep_ms"] = kms
    out["mixed_over_fp64"] = out["mixed_solves_per_s"] / out["fp64_solves_per_s"]
    out["beta_rel_err_vs_fp64"] = beta_rel_err(res["mixed"], res["fp64"])
    return out


def run_varsel(ctx):
    """Forward selection over the sphere-box covariates (unit-weight AICc per candidate set)."""
    log("variable selection")
    t0 = time.perf_counter()
    sel, models = ctx.indep_vars_select(threshold=3.0)
    dt = time.perf_counter() - t0
    return {"models": len(models), "selected": [int(v) for v in sel], "seconds": dt}


def autotune_summary(lines):
    """Classes the sweep autotuner has seen, and the largest gains of the
    shapes it chose over the default shape (time of shape 0 / chosen)."""
    gains = []
    for l in lines:
        chosen = int(l.split("chosen=")[1].split()[0])
        ms = [float(v) for v in l.split("ms=")[1].split(",")]
        if chosen > 0 and ms[0] > 0 and ms[chosen] > 0:
            gains.append((ms[0] / ms[chosen], l))
    gains.sort(reverse=True)
    return {"classes": len(lines), "tuned_to_non_default": len(gains),
            "median_gain": float(np.median([g for g, _ in gains])) if gains else None,
            "top": [l for _, l in gains[:3]]}


def serve_submit(capi, srv, r):
    n = r["X"].shape[0]
    if r["search"]:
        return srv.submit(r["coords"], r["X"], r["y"], capi.BISQUARE, True, 0.0, select_bw=True, crit=capi.CRIT_CV)
    return srv.submit(r["coords"], r["X"], r["y"], capi.BISQUARE, True, float(min(n - 1, 100)))


def run_serve(capi, synth, num, rates, check):
    """Serving stream replayed in real time at each rate (the same requests,
    arrival times scaled); fits use an adaptive bisquare kernel of 100
    neighbours, searches CV over [20, n] neighbours.  The beta error is taken
    against the FP64 server on `check` requests spread evenly over the sizes
    of the stream (searches and fits)."""
    reqs = synth.serving_stream(num, rate=1.0)
    log("serve: %d requests (%d searches), rates %s/s" % (num, sum(r["search"] for r in reqs), rates))
    out = {"requests": num, "searches": int(sum(r["search"] for r in reqs)),
           "n_max": int(max(r["X"].shape[0] for r in reqs)), "k_max": int(max(r["X"].shape[1] for r in reqs)),
           "rates": []}
    fits = np.array([not r["search"] for r in reqs])
    last = None
    for rate in rates:
        log("serve: rate %.2f/s" % rate)
        with capi.Server(0, capi.MIXED, max_inflight=64) as srv:
            start = time.perf_counter()
            ids = []
            for q, r in enumerate(reqs):
                delay = r["t"] / rate - (time.perf_counter() - start)
                if delay > 0:
                    time.sleep(delay)
                ids.append(serve_submit(capi, srv, r))
                if (q + 1) % 100 == 0:
                    log("serve: rate %.2f/s, %d requests submitted" % (rate, q + 1))
            res = [srv.wait(i) for i in ids]
            wall = time.perf_counter() - start
            st = srv.stats()
        lat = np.array([r["latency_ms"] for r in res])
        solves = sum(r["solves"] for r in res)
        out["rates"].append({
            "rate_per_s": rate, "seconds": wall, "p50_ms": float(np.percentile(lat, 50)),
            "p90_ms": float(np.percentile(lat, 90)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "fit_p50_ms": float(np.percentile(lat[fits], 50)),
            "fit_p99_ms": float(np.percentile(lat[fits], 99)),
            "search_p50_ms": float(np.percentile(lat[~fits], 50)) if (~fits).any() else None,
            "search_p99_ms": float(np.percentile(lat[~fits], 99)) if (~fits).any() else None,
            "solves": solves, "solves_per_s": solves / wall, "rounds": st["rounds"],
            "device_busy_frac": st["busy_ms"] * 1e-3 / wall, "peak_inflight": st["peak_inflight"],
            "peak_device_bytes": st["peak_bytes"],
            "failed": int(sum(r["status"] != capi.GWM_OK for r in res))})
        last = res
    if check:
        idx = sorted(range(num), key=lambda i: reqs[i]["X"].shape[0])
        pick = sorted(set(idx[int(round(q))] for q in np.linspace(0, num - 1, check)))
        cs, el, same_bw = [], [], 0
        with capi.Server(0, capi.FP64, max_inflight=16) as srv:
            ref = [srv.wait(serve_submit(capi, srv, reqs[i])) for i in pick]
        for i, g in zip(pick, ref):
            a, b = col_rel(last[i]["betas"], g["betas"])
            cs.append(a)
            el.append(b)
            same_bw += last[i]["bandwidth"] == g["bandwidth"]
        out["checked"] = len(pick)
        out["checked_n"] = [int(reqs[i]["X"].shape[0]) for i in pick]
        out["checked_searches"] = int(sum(reqs[i]["search"] for i in pick))
        out["beta_rel_err_vs_fp64"] = max(cs)
        out["beta_elementwise_rel_err_vs_fp64"] = max(el)
        out["bandwidths_identical"] = same_bw
    return out


def load_profile(name):
    """A committed counter summary, profiles/<name>.json, or None."""
    path = os.path.join(ROOT, "profiles", "%s.json" % name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=16384,
                    help="focus points of the CPU baseline sample (about 10 s of host work)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mgwr-n", type=int, default=50_000, help="multiscale-cosines size (0: skip)")
    ap.add_argument("--mgwr-hat-n", type=int, default=4000, help="multiscale size with the hat matrix (0: skip)")
    ap.add_argument("--predict-m", type=int, default=-1, help="prediction locations (default n; 0: skip)")
    ap.add_argument("--batch-problems", type=int, default=512, help="uneven batch size (0: skip)")
    ap.add_argument("--no-loop", action="store_true", help="skip the single-problem loop of the batch")
    ap.add_argument("--serve-requests", type=int, default=500, help="serving stream length (0: skip)")
    ap.add_argument("--serve-rates", type=str, default="4,6",
                    help="serving-stream arrival rates, requests/s (comma-separated; one replay each)")
    ap.add_argument("--serve-check", type=int, default=12,
                    help="requests spread over the stream's sizes re-run in FP64 for the beta error")
    ap.add_argument("--cfg1", type=int, default=1, help="n = 2000 uniform-square leg (0: skip)")
    ap.add_argument("--refbench-n", type=int, default=50_000,
                    help="the reference's own benchmark shape, n = 50,000, k = 3 (0: skip)")
    ap.add_argument("--refbench-sample", type=int, default=8192, help="focus points of its CPU sample")
    ap.add_argument("--mgwr-parity-n", type=int, default=1000,
                    help="multiscale CPU parity size (the oracle's full backfit; 0: skip)")
    ap.add_argument("--sweeps-only", type=int, default=0,
                    help="profiling mode: run this many mixed AICc sweeps (and as many FP64 ones) and exit")
    ap.add_argument("--sweep-bw", type=float, default=61.890021059082855,
                    help="bandwidth (km) of the profiling sweeps (default: the selected one)")
    ap.add_argument("--profile-step", action="store_true",
                    help="profiling mode: one mixed step (search + fit) after a warm-up sweep, its kernel "
                         "launches, times and pairs as JSON, then exit")
    ap.add_argument("--skip-headline", action="store_true",
                    help="profiling mode: run only the multiscale / batch / serve legs and print their JSON")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="profiling mode: one device-to-device copy of this many bytes (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        src = torch.ones(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return

    from libgwmodel_amd import capi, synth
    rates = [float(r) for r in args.serve_rates.split(",") if r]
    threads, aff = cpu_threads()

    if args.skip_headline:
        extra = {"mgwr": run_mgwr(capi, synth, args.mgwr_n) if args.mgwr_n else None,
                 "mgwr_hat": run_mgwr_hat(capi, synth, args.mgwr_hat_n) if args.mgwr_hat_n else None,
                 "batch": run_batch(capi, synth, args.batch_problems, not args.no_loop) if args.batch_problems else None,
                 "serve": run_serve(capi, synth, args.serve_requests, rates, args.serve_check)
                 if args.serve_requests else None}
        print(json.dumps(extra, default=jsonable), flush=True)
        return

    n = args.n
    lo, hi = synth.SPHERE_BOX_BW_KM
    log("generating sphere-box workload n=%d" % n)
    d = synth.sphere_box(n=n)
    k = d["X"].shape[1]
    ctx = capi.Context(0)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)

    if args.profile_step:
        ctx.criterion(capi.GAUSSIAN, False, 300.0, capi.CRIT_AIC, capi.MIXED)
        ctx.reset_timers()
        ctx.set_timing(True)
        t0 = time.perf_counter()
        r = run_steps(ctx, capi, capi.MIXED, 1, lo, hi)
        dt = time.perf_counter() - t0
        ctx.set_timing(False)
        rep = {"step_ms": dt * 1e3, "bandwidth_km": r["bandwidth"], "n": n}
        for name in ("d2_table", "contract_mixed_plain", "contract_fp64_plain", "contract_mixed_hat",
                     "contract_mixed_qdiag", "solve"):
            launches, ms = ctx.kernel_time(name)
            rep[name] = {"launches": launches, "ms": ms}
            if name.startswith("contract"):
                rep[name]["pairs"] = ctx.kernel_pairs(name)
        print(json.dumps(rep))
        ctx.close()
        return

    if args.sweeps_only:
        ctx.reset_timers()
        ctx.set_timing(True)
        for prec in (capi.MIXED, capi.FP64):
            for _ in range(args.sweeps_only):
                ctx.criterion(capi.GAUSSIAN, False, args.sweep_bw, capi.CRIT_AIC, prec)
        ctx.set_timing(False)
        rep = {"sweeps": args.sweeps_only, "n": n, "bandwidth_km": args.sweep_bw}
        for name in ("contract_mixed_plain", "contract_fp64_plain"):
            launches, ms = ctx.kernel_time(name)
            rep[name] = {"launches": launches, "avg_ms": ms / max(launches, 1),
                         "pairs_per_launch": ctx.kernel_pairs(name) / max(launches, 1)}
        print(json.dumps(rep))
        ctx.close()
        return

    # warmup (kernel loads, allocation of the workspaces)
    log("warmup x%d" % args.warmup)
    ctx.criterion(capi.GAUSSIAN, False, 300.0, capi.CRIT_AIC, capi.MIXED)
    ctx.criterion(capi.GAUSSIAN, False, 300.0, capi.CRIT_AIC, capi.FP64)
    run_steps(ctx, capi, capi.MIXED, args.warmup, lo, hi)

    # timed region: mixed precision, kernel timing with HIP events on the context stream
    log("timed mixed steps x%d" % args.steps)
    ctx.reset_timers()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    mixed = run_steps(ctx, capi, capi.MIXED, args.steps, lo, hi)
    t_mixed = time.perf_counter() - t0
    ctx.set_timing(False)
    solves_mixed = ctx.solve_count()
    tab_l, tab_ms = ctx.kernel_time("d2_table")
    launches, ms = ctx.kernel_time("contract_mixed_plain")
    swept_pairs = ctx.kernel_pairs("contract_mixed_plain")
    l64_in_mixed, ms64_in_mixed = ctx.kernel_time("contract_fp64_plain")
    hat_l, hat_ms = ctx.kernel_time("contract_mixed_hat")

    # FP64 path, same step
    log("fp64 steps x%d" % args.fp64_steps)
    ctx.reset_timers()
    ctx.set_timing(True)
    t0 = time.perf_counter()
    fp64 = run_steps(ctx, capi, capi.FP64, args.fp64_steps, lo, hi)
    t_fp64 = time.perf_counter() - t0
    ctx.set_timing(False)
    solves_fp64 = ctx.solve_count()
    l64, ms64 = ctx.kernel_time("contract_fp64_plain")
    swept64 = ctx.kernel_pairs("contract_fp64_plain")
    # Solves per step = the (focus, bandwidth) evaluations of the search and
    # the fit, each counted once: the FP64 step's count.  The mixed search takes
    # the same iterates; its FP64 re-evaluations of uncertain decisions repeat
    # candidates it evaluated already and are not counted again.
    solves_per_step = solves_fp64 // args.fp64_steps
    value = solves_per_step * args.steps / t_mixed
    value_fp64 = solves_fp64 / t_fp64
    m = n if args.predict_m < 0 else args.predict_m
    predict = run_predict(ctx, capi, d, mixed["bandwidth"], m) if m else None
    varsel = run_varsel(ctx)
    ctx.close()

    P = k * (k + 1) // 2 + k
    flops_per_launch = 2.0 * n * n * P          # n focus x n data x P columns, multiply-add = 2
    avg_ms = ms / max(launches, 1)
    # algorithmic work: every (focus, data) pair of the n solves a launch
    # performs; executed: the pairs it actually streamed (pruning skips data
    # chunks whose weights are exactly zero)
    swept_per_launch = swept_pairs / max(launches, 1)
    executed_flops = 2.0 * swept_per_launch * P
    achieved = flops_per_launch / (avg_ms * 1e-3)
    avg_ms64 = ms64 / max(l64, 1)
    traffic = load_profile("traffic_contract_mixed_plain_n%d" % n)
    valu = load_profile("valu_contract_mixed_plain_n%d" % n)
    # VALU issue bound of the launch: instructions per executed pair x pairs /
    # the chip's lane-instruction rate (1024 SIMDs x 16 lanes x 2.4 GHz)
    lane_rate = 1024 * 16 * 2.4e9
    valu_bound_ms = (valu["valu_insts_per_pair"] * swept_per_launch / lane_rate * 1e3) if valu else None

    cpu = None
    cpu_lin = None
    if not args.no_cpu:
        log("cpu baseline: oracle, %d threads (affinity %d), %d focus points" % (threads, aff, args.cpu_sample))
        lin = [cpu_baseline(d, mixed["bandwidth"], s, threads)[1] for s in (1024, 2048)]
        cpu_lin = {"seconds_1024": lin[0], "seconds_2048": lin[1], "ratio": lin[1] / lin[0]}
        cv, dt = cpu_baseline(d, mixed["bandwidth"], args.cpu_sample, threads)
        cpu = {"value": cv, "unit": "local WLS solves/s", "cores": threads, "kind": "port",
               "sample": "%d focus points x n=%d data, one AICc evaluation at the selected bandwidth, "
                         "great-circle Gaussian, OpenMP (%.1f s); 2048 vs 1024 focus points took %.2fx the time"
                         % (args.cpu_sample, n, dt, cpu_lin["ratio"]),
               "affinity_cores": aff, "cpu_model": cpu_model(), "linearity": cpu_lin}

    cfg1 = leg("cfg1", run_cfg1, capi, synth, threads) if args.cfg1 else None
    refbench = leg("refbench", run_refbench, capi, synth, threads, args.refbench_sample) if args.refbench_n else None
    mgwr_parity = leg("mgwr_cpu_parity", run_mgwr_cpu_parity, capi, synth, args.mgwr_parity_n, threads) \
        if args.mgwr_parity_n else None
    mgwr = leg("mgwr", run_mgwr, capi, synth, args.mgwr_n) if args.mgwr_n else None
    mgwr_hat = leg("mgwr_hat", run_mgwr_hat, capi, synth, args.mgwr_hat_n) if args.mgwr_hat_n else None
    batch = leg("batch", run_batch, capi, synth, args.batch_problems, not args.no_loop) \
        if args.batch_problems else None
    serve = leg("serve", run_serve, capi, synth, args.serve_requests, rates, args.serve_check) \
        if args.serve_requests else None
    cs, el = col_rel(mixed["betas"], fp64["betas"])
    big = np.abs(fp64["betas"]) >= 0.1 * np.max(np.abs(fp64["betas"]), axis=0)
    el_big = float(np.max(np.abs(mixed["betas"] - fp64["betas"])[big] / np.abs(fp64["betas"])[big]))
    rel_all = np.abs(mixed["betas"] - fp64["betas"]) / np.maximum(np.abs(fp64["betas"]), 1e-300)
    el_p99 = float(np.percentile(rel_all, 99))

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "local WLS solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t_mixed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: fp32 weights, fp16 hi+lo split MFMA (fp32 acc, fp64 partials), fp64 solves",
        "data": "synthetic (SplitMix64 seed 1002)",
        "config": {"workload": "GWRBasic AICc golden-section bandwidth search + fit with hat diagnostics, "
                               "great-circle, Gaussian fixed bandwidth in [55.66, 2226.39] km",
                   "n": n, "k": k, "selected_bandwidth_km": mixed["bandwidth"],
                   "solves_per_step": solves_per_step,
                   "solves_executed_per_step_mixed": solves_mixed // args.steps,
                   "fp64_certify_sweeps_per_step_in_mixed": l64_in_mixed / args.steps,
                   "fp64_certify_ms_per_step_in_mixed": ms64_in_mixed / args.steps},
        "fp64_value": value_fp64,
        "fp64_ms_per_step": 1e3 * t_fp64 / args.fp64_steps,
        # same search and fit, same solves: the wall-time ratio per step
        "mixed_over_fp64": value / value_fp64,
        "mixed_sweep_over_fp64_sweep": avg_ms64 / avg_ms if avg_ms > 0 else None,
        "bandwidth_identical": mixed["bandwidth"] == fp64["bandwidth"],
        "iterates_identical": [x for x, _ in mixed["trace"]] == [x for x, _ in fp64["trace"]],
        "beta_rel_err_vs_fp64": cs,
        "beta_elementwise_rel_err_vs_fp64": el,
        "beta_elementwise_floor": "entries with |beta| >= 1e-3 max|beta| of their column",
        "beta_elementwise_rel_err_vs_fp64_floor_0.1": el_big,
        "beta_elementwise_rel_err_vs_fp64_p99": el_p99,
        "se_rel_err_vs_fp64": beta_rel_err(mixed["betasSE"], fp64["betasSE"]),
        "aicc_rel_err_vs_fp64": abs(mixed["diagnostic"]["AICc"] - fp64["diagnostic"]["AICc"]) /
        abs(fp64["diagnostic"]["AICc"]),
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["mixed"] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK["mixed"],
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": (traffic.get("source"), traffic.get("kernel")) if traffic else None,
                     "kernel": "contract_mixed_kernel (AICc sweep)", "avg_launch_ms": avg_ms,
                     "launches": launches,
                     "flops_per_launch": flops_per_launch,
                     "executed_pair_fraction": swept_per_launch / (float(n) * n),
                     "executed_flops_per_launch": executed_flops,
                     "executed_tflops": executed_flops / (avg_ms * 1e-3) / 1e12,
                     "valu_insts_per_pair": valu["valu_insts_per_pair"] if valu else None,
                     "valu_issue_bound_ms": valu_bound_ms,
                     "valu_source": valu.get("source") if valu else None,
                     # the sweeps read each executed pair's FP32 squared distance
                     # from the step's distance table (4 B per pair)
                     "table_stream_tb_per_s": 4.0 * swept_per_launch / (avg_ms * 1e-3) / 1e12 if tab_l else None,
                     "table_stream_frac_of_hbm": 4.0 * swept_per_launch / (avg_ms * 1e-3) / HBM_PEAK if tab_l else None,
                     "table_build_ms_per_step": tab_ms / args.steps if tab_l else None,
                     "fp64_kernel_avg_launch_ms": avg_ms64,
                     "fp64_frac_of_fp64_peak": flops_per_launch / (avg_ms64 * 1e-3) / PEAK["fp64"] if l64 else None,
                     "fp64_executed_pair_fraction": swept64 / max(l64, 1) / (float(n) * n)},
        "cpu_baseline": cpu,
        "serve_p99_ms": serve["rates"][0]["p99_ms"] if serve and "rates" in serve else None,
        "serve_p50_ms": serve["rates"][0]["p50_ms"] if serve and "rates" in serve else None,
        "serve_rate_per_s": serve["rates"][0]["rate_per_s"] if serve and "rates" in serve else None,
        "serve": serve,
        "batch": batch,
        "cfg1": cfg1,
        "refbench": refbench,
        "mgwr_cpu_parity": mgwr_parity,
        "mgwr": mgwr,
        "mgwr_hat": mgwr_hat,
        "predict": predict,
        "varsel": varsel,
        "autotune": autotune_summary(capi.autotune_report()),
    }
    print(json.dumps(line, default=jsonable), flush=True)


if __name__ == "__main__":
    main()
