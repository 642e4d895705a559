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

Three more workloads ride along in the same JSON line (keys "mgwr", "batch"
and "serve"): GWRMultiscale backfitting at n = 50,000, k = 6 (bisquare
adaptive bandwidths chosen by CV for every parameter, CVR threshold 1e-5) in
mixed precision and FP64; 512 uneven problems (n log-uniform in [200, 20000], k in 3..12,
Gaussian / bisquare / tricube, adaptive or fixed), each a CV golden-section
search plus fit, run through the batching server against a loop of
single-problem calls; and a Poisson stream of fit and CV-search requests
(n log-uniform in [500, 200000], k in 2..16, clustered coordinates) replayed in
real time against the continuous-batching server, with p50/p99 latency.

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


def cpu_baseline(d, bw, sample, threads):
    from oracle import pyoracle
    t0 = time.perf_counter()
    pyoracle.criterion(d["coords"], d["X"], d["y"], pyoracle.GAUSSIAN, False, bw, pyoracle.CRIT_AIC, geographic=True,
                       nthreads=threads, focus_count=sample)
    dt = time.perf_counter() - t0
    return sample / dt, dt


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
    bad = sum(r["status"] != capi.GWM_OK for r in res)
    solves = sum(r["solves"] for r in res)
    out = {"problems": num, "sum_n": sum_n, "solves": solves, "seconds": dt, "solves_per_s": solves / dt,
           "rounds": st["rounds"], "failed": bad}
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
    out["bandwidths"] = [float(b) for b in res["mixed"]["bandwidths"]]
    out["bandwidths_identical"] = bool(np.array_equal(res["mixed"]["bandwidths"], res["fp64"]["bandwidths"]))
    out["beta_rel_err_vs_fp64"] = beta_rel_err(res["mixed"]["betas"], res["fp64"]["betas"])
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


def run_serve(capi, synth, num, rate, check):
    """Serving stream replayed in real time; fits use an adaptive bisquare
    kernel of 100 neighbours, searches CV over [20, n] neighbours."""
    reqs = synth.serving_stream(num, rate=rate)
    log("serve: %d requests at %.2f/s (%d searches)" % (num, rate, sum(r["search"] for r in reqs)))
    with capi.Server(0, capi.MIXED, max_inflight=64) as srv:
        start = time.perf_counter()
        ids = []
        for r in reqs:
            delay = r["t"] - (time.perf_counter() - start)
            if delay > 0:
                time.sleep(delay)
            ids.append(serve_submit(capi, srv, r))
        res = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - start
        st = srv.stats()
    lat = np.array([r["latency_ms"] for r in res])
    fits = np.array([not r["search"] for r in reqs])
    solves = sum(r["solves"] for r in res)
    out = {"requests": num, "rate_per_s": rate, "searches": int((~fits).sum()), "seconds": wall,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(lat.max()), "fit_p50_ms": float(np.percentile(lat[fits], 50)),
           "fit_p99_ms": float(np.percentile(lat[fits], 99)),
           "search_p50_ms": float(np.percentile(lat[~fits], 50)) if (~fits).any() else None,
           "solves": solves, "solves_per_s": solves / wall, "rounds": st["rounds"],
           "peak_inflight": st["peak_inflight"], "failed": int(sum(r["status"] != capi.GWM_OK for r in res))}
    if check:
        # beta error vs the FP64 server on the smallest requests of each type
        idx = sorted(range(num), key=lambda i: reqs[i]["X"].shape[0])
        pick = [i for i in idx if reqs[i]["search"]][:check] + [i for i in idx if not reqs[i]["search"]][:check]
        errs, same_bw = [], 0
        with capi.Server(0, capi.FP64, max_inflight=16) as srv:
            ref = [srv.wait(serve_submit(capi, srv, reqs[i])) for i in pick]
        for i, g in zip(pick, ref):
            errs.append(beta_rel_err(res[i]["betas"], g["betas"]))
            same_bw += res[i]["bandwidth"] == g["bandwidth"]
        out["beta_rel_err_vs_fp64"] = max(errs)
        out["checked"] = len(pick)
        out["bandwidths_identical"] = same_bw
    return out


def load_traffic(name):
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % name)
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
    ap.add_argument("--serve-requests", type=int, default=64, help="serving stream length (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0, help="serving-stream arrival rate, requests/s")
    ap.add_argument("--serve-check", type=int, default=3,
                    help="requests of each type re-run in FP64 for the beta error")
    ap.add_argument("--sweeps-only", type=int, default=0,
                    help="profiling mode: run this many mixed AICc sweeps (and as many FP64 ones) and exit")
    ap.add_argument("--sweep-bw", type=float, default=61.890021059082855,
                    help="bandwidth (km) of the profiling sweeps (default: the selected one)")
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

    if args.skip_headline:
        extra = {"mgwr": run_mgwr(capi, synth, args.mgwr_n) if args.mgwr_n else None,
                 "mgwr_hat": run_mgwr_hat(capi, synth, args.mgwr_hat_n) if args.mgwr_hat_n else None,
                 "batch": run_batch(capi, synth, args.batch_problems, not args.no_loop) if args.batch_problems else None,
                 "serve": run_serve(capi, synth, args.serve_requests, args.serve_rate, args.serve_check)
                 if args.serve_requests else None}
        print(json.dumps(extra), flush=True)
        return

    n = args.n
    lo, hi = synth.SPHERE_BOX_BW_KM
    log("generating sphere-box workload n=%d" % n)
    d = synth.sphere_box(n=n)
    k = d["X"].shape[1]
    ctx = capi.Context(0)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)

    if args.sweeps_only:
        for prec in (capi.MIXED, capi.FP64):
            for _ in range(args.sweeps_only):
                ctx.criterion(capi.GAUSSIAN, False, args.sweep_bw, capi.CRIT_AIC, prec)
        print(json.dumps({"sweeps": args.sweeps_only, "n": n}))
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
    launches, ms = ctx.kernel_time("contract_mixed_plain")
    swept_pairs = ctx.kernel_pairs("contract_mixed_plain")
    l64_in_mixed, _ = ctx.kernel_time("contract_fp64_plain")
    hat_l, hat_ms = ctx.kernel_time("contract_mixed_hat")
    value = solves_mixed / t_mixed

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
    value_fp64 = solves_fp64 / t_fp64
    m = n if args.predict_m < 0 else args.predict_m
    predict = run_predict(ctx, capi, d, mixed["bandwidth"], m) if m else None
    varsel = run_varsel(ctx)
    ctx.close()

    P = k * (k + 1) // 2 + k
    flops_per_launch = 2.0 * n * n * P          # n focus x n data x P columns, multiply-add = 2
    avg_ms = ms / max(launches, 1)
    # executed work: the pairs the sweeps processed (pruning skips chunks whose
    # weights are exactly zero); algorithmic: all n^2 pairs
    swept_per_launch = swept_pairs / max(launches, 1)
    executed_flops = 2.0 * swept_per_launch * P
    achieved = executed_flops / (avg_ms * 1e-3)
    effective = flops_per_launch / (avg_ms * 1e-3)
    avg_ms64 = ms64 / max(l64, 1)
    traffic = load_traffic("contract_mixed_plain_n%d" % n)

    cpu = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        log("cpu baseline: oracle, %d threads, %d focus points" % (threads, args.cpu_sample))
        cv, dt = cpu_baseline(d, mixed["bandwidth"], args.cpu_sample, threads)
        cpu = {"value": cv, "unit": "local WLS solves/s", "cores": threads, "kind": "port",
               "sample": "%d focus points x n=%d data, one AICc evaluation at the selected bandwidth, "
                         "great-circle Gaussian, OpenMP (%.1f s)" % (args.cpu_sample, n, dt)}

    mgwr = run_mgwr(capi, synth, args.mgwr_n) if args.mgwr_n else None
    mgwr_hat = run_mgwr_hat(capi, synth, args.mgwr_hat_n) if args.mgwr_hat_n else None
    batch = run_batch(capi, synth, args.batch_problems, not args.no_loop) if args.batch_problems else None
    serve = run_serve(capi, synth, args.serve_requests, args.serve_rate, args.serve_check) \
        if args.serve_requests else None

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
                   "solves_per_step": solves_mixed // args.steps,
                   "fp64_certify_sweeps_in_mixed": l64_in_mixed},
        "fp64_value": value_fp64,
        "fp64_ms_per_step": 1e3 * t_fp64 / args.fp64_steps,
        "mixed_over_fp64": value / value_fp64,
        # same search and fit, wall time per step: the mixed step's FP64
        # re-evaluations of certified decisions count as solves above, not here
        "mixed_step_speedup": (t_fp64 / args.fp64_steps) / (t_mixed / args.steps),
        "mixed_sweep_over_fp64_sweep": avg_ms64 / avg_ms if avg_ms > 0 else None,
        "bandwidth_identical": mixed["bandwidth"] == fp64["bandwidth"],
        "beta_rel_err_vs_fp64": beta_rel_err(mixed["betas"], fp64["betas"]),
        "se_rel_err_vs_fp64": beta_rel_err(mixed["betasSE"], fp64["betasSE"]),
        "aicc_rel_err_vs_fp64": abs(mixed["diagnostic"]["AICc"] - fp64["diagnostic"]["AICc"]) /
        abs(fp64["diagnostic"]["AICc"]),
        "roofline": {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK["mixed"] / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK["mixed"],
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "kernel": "contract_mixed_kernel (AICc sweep)", "avg_launch_ms": avg_ms,
                     "flops_per_launch": executed_flops,
                     "swept_pair_fraction": swept_per_launch / (float(n) * n),
                     "effective_tflops_all_pairs": effective / 1e12,
                     "fp64_kernel_avg_launch_ms": avg_ms64,
                     "fp64_frac_of_fp64_peak": 2.0 * (swept64 / max(l64, 1)) * P / (avg_ms64 * 1e-3) / PEAK["fp64"]
                     if l64 else None,
                     "fp64_swept_pair_fraction": swept64 / max(l64, 1) / (float(n) * n)},
        "cpu_baseline": cpu,
        "serve_p99_ms": serve["p99_ms"] if serve else None,
        "serve_p50_ms": serve["p50_ms"] if serve else None,
        "serve": serve,
        "batch": batch,
        "mgwr": mgwr,
        "mgwr_hat": mgwr_hat,
        "predict": predict,
        "varsel": varsel,
        "autotune": autotune_summary(capi.autotune_report()),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
