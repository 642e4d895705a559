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
    ap.add_argument("--cpu-sample", type=int, default=512, help="focus points of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sweeps-only", type=int, default=0,
                    help="profiling mode: run this many mixed AICc sweeps (and as many FP64 ones) and exit")
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

    n = args.n
    lo, hi = synth.CONFIG2_BW_KM
    log("generating config-2 workload n=%d" % n)
    d = synth.config2(n=n)
    k = d["X"].shape[1]
    ctx = capi.Context(0)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)

    if args.sweeps_only:
        for prec in (capi.MIXED, capi.FP64):
            for _ in range(args.sweeps_only):
                ctx.criterion(capi.GAUSSIAN, False, 100.0, capi.CRIT_AIC, prec)
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
    value_fp64 = solves_fp64 / t_fp64
    ctx.close()

    P = k * (k + 1) // 2 + k
    flops_per_launch = 2.0 * n * n * P          # n focus x n data x P columns, multiply-add = 2
    avg_ms = ms / max(launches, 1)
    achieved = flops_per_launch / (avg_ms * 1e-3)
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
                     "flops_per_launch": flops_per_launch,
                     "fp64_kernel_avg_launch_ms": avg_ms64,
                     "fp64_frac_of_fp64_peak": flops_per_launch / (avg_ms64 * 1e-3) / PEAK["fp64"] if l64 else None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
