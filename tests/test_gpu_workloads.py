"""Parity on the five benchmark workloads (libgwmodel_amd/synth.py), through
the C ABI, against the CPU oracle and between the precisions.

- n = 2,000 uniform square (k = 4, bisquare adaptive CV over [20, 2000], fit
  with hat diagnostics and S): full size against the oracle.
- n = 100,000 sphere box (k = 8, great circle, Gaussian, AICc over
  [55.66, 2226.39] km): full-size FP64 and mixed searches (same bandwidth and
  iterates), then betas, SE and qDiag of the first 1,024 focus points against
  the oracle's fit at the selected bandwidth (GWRBasic.cpp:386-438 restricted
  to those focus points).
- GWRMultiscale on the multiscale-cosines workload at n = 3,000 (k = 6,
  bisquare adaptive CV per parameter, CVR 1e-5, point density of the n =
  50,000 workload): against the oracle's result in tests/golden/
  cfg3_mgwr_n3000.npz (the oracle needs minutes at this size;
  tests/golden/make_cfg3_fixture.py regenerates it).
- 512 uneven problems: every one through the batching server and through the
  single-problem API (bitwise equal), the 16 smallest against the oracle.
- A 24-request serving stream (n up to 20,000, k up to 16): every result of
  the mixed server against the FP64 single-problem API, 4 against the oracle.

Tolerances: FP64 vs the oracle -- bandwidths, golden-section iterates and
adaptive thresholds bit-exact, betas / SE / qDiag / S within 1e-9 relative
(+1e-12 absolute), diagnostics within 1e-9; mixed vs FP64 -- identical
bandwidths, betas within 1e-6 normwise per focus point and column-scaled.
"""
import os
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def close(a, b, rtol, atol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b) + atol))


def col_rel(b, ref):
    return float(np.max(np.max(np.abs(b - ref), axis=0) / np.maximum(np.max(np.abs(ref), axis=0), 1e-300)))


def normwise_rel(b, ref):
    return float(np.max(np.linalg.norm(b - ref, axis=1) / np.linalg.norm(ref, axis=1)))


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


DIAG_KEYS = ["RSS", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"]


def test_uniform_square_n2000_full(ctx, oracle):
    say("case: uniform square n=2000 bisquare adaptive CV search + fit, FP64 and mixed vs oracle")
    d = synth.uniform_square(n=2000)
    o = oracle.gwr_run(d["coords"], d["X"], d["y"], oracle.BISQUARE, True, 0, select_bw=True, crit=oracle.CRIT_CV,
                       lower=20, upper=2000, want_S=True)
    ctx.set_data(d["coords"], d["X"], d["y"])
    g = ctx.gwr_run(capi.BISQUARE, True, 0, select_bw=True, crit=capi.CRIT_CV, lower=20, upper=2000, want_S=True)
    assert g["bandwidth"] == o["bandwidth"]
    assert [x for x, _ in g["trace"]] == [x for x, _ in o["trace"]]
    for (x, f), (_, fo) in zip(g["trace"], o["trace"]):
        assert f == pytest.approx(fo, rel=1e-11), x
    for key in DIAG_KEYS:
        assert g["diagnostic"][key] == pytest.approx(o["diagnostic"][key], rel=1e-9), key
    for key in ["betas", "betasSE", "qDiag", "S"]:
        assert close(g[key], o[key], 1e-9, 1e-12), key
    assert close(g["shat"], o["shat"], 1e-10)
    m = ctx.gwr_run(capi.BISQUARE, True, 0, select_bw=True, crit=capi.CRIT_CV, precision=capi.MIXED, lower=20,
                    upper=2000)
    assert m["bandwidth"] == g["bandwidth"]
    assert normwise_rel(m["betas"], g["betas"]) <= 1e-6 and col_rel(m["betas"], g["betas"]) <= 1e-6


def test_sphere_box_n100000_full(ctx, oracle):
    say("case: sphere box n=100000 AICc searches, FP64 and mixed; 1024 focus points vs oracle")
    d = synth.sphere_box(n=100_000)
    lo, hi = synth.SPHERE_BOX_BW_KM
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    bf, tf = ctx.select_bandwidth(capi.GAUSSIAN, False, capi.CRIT_AIC, capi.FP64, lo, hi)
    bm, tm = ctx.select_bandwidth(capi.GAUSSIAN, False, capi.CRIT_AIC, capi.MIXED, lo, hi)
    assert bm == bf
    assert [x for x, _ in tm] == [x for x, _ in tf]
    gf = ctx.fit(capi.GAUSSIAN, False, bf, capi.FP64)
    gm = ctx.fit(capi.GAUSSIAN, False, bf, capi.MIXED)
    assert normwise_rel(gm["betas"], gf["betas"]) <= 1e-6 and col_rel(gm["betas"], gf["betas"]) <= 1e-6
    m = 1024
    o = oracle.fit_core(d["coords"], d["X"], d["y"], oracle.GAUSSIAN, False, bf, geographic=True, focus_count=m)
    assert o is not None
    assert close(gf["betas"][:m], o["betas"][:m], 1e-9, 1e-12)
    assert close(gf["betasSE_raw"][:m], o["betasSE_raw"][:m], 1e-9, 1e-14)
    assert normwise_rel(gm["betas"][:m], o["betas"][:m]) <= 1e-6


def test_multiscale_cosines_n3000_vs_oracle_fixture(ctx):
    say("case: GWRMultiscale multiscale cosines n=3000, FP64 vs oracle fixture, mixed bandwidths")
    sys.path.insert(0, GOLDEN)
    import make_cfg3_fixture as fx
    ref = np.load(os.path.join(GOLDEN, "cfg3_mgwr_n3000.npz"))
    d = fx.inputs()
    assert np.array_equal(fx.fingerprint(d), ref["fingerprint"])
    k = d["X"].shape[1]
    args = ([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k, [0] + [1] * (k - 1), [1e-5] * k)
    ctx.set_data(d["coords"], d["X"], d["y"])
    g = ctx.mgwr_run(*args, crit_type=0, crit_threshold=1e-5)
    assert np.array_equal(g["bandwidths"], ref["bandwidths"])
    assert g["iterations"] == int(ref["iterations"][0])
    assert close(g["betas"], ref["betas"], 1e-9, 1e-12)
    diag = [g["diagnostic"][key] for key in ["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"]]
    assert close(diag, ref["diag"], 1e-9)
    gm = ctx.mgwr_run(*args, crit_type=0, crit_threshold=1e-5, precision=capi.MIXED)
    assert np.array_equal(gm["bandwidths"], g["bandwidths"])
    assert col_rel(gm["betas"], g["betas"]) <= 1e-6


def test_uneven_batch_512_server_vs_single_and_oracle(ctx, oracle):
    say("case: 512 uneven problems, server vs single-problem API; 16 smallest vs oracle")
    probs = synth.uneven_batch(512)
    with capi.Server(0, capi.MIXED, max_inflight=512) as srv:
        ids = [srv.submit(p["coords"], p["X"], p["y"], p["kernel"], p["adaptive"], 0.0, select_bw=True,
                          crit=capi.CRIT_CV) for p in probs]
        res = [srv.wait(i) for i in ids]
        report = srv.autotune_report()
    # the server's shape tuner ran both shapes on classes of this batch
    # (launches of either shape are among the bit-exact results below)
    say("\n".join(report))
    assert any(" rounds=0," not in l and not l.endswith(",0") for l in report)
    for pi, (p, r) in enumerate(zip(probs, res)):
        assert r["status"] == capi.GWM_OK, pi
        ctx.set_data(p["coords"], p["X"], p["y"])
        g = ctx.gwr_run(p["kernel"], p["adaptive"], 0.0, select_bw=True, crit=capi.CRIT_CV, precision=capi.MIXED)
        assert r["bandwidth"] == g["bandwidth"], pi
        for key in ["betas", "betasSE", "qDiag"]:
            assert np.array_equal(r[key], g[key]), (pi, key)
    smallest = sorted(range(len(probs)), key=lambda i: probs[i]["X"].shape[0])[:16]
    for pi in smallest:
        p = probs[pi]
        say("  oracle: problem %d n=%d k=%d kernel=%d adaptive=%d" % (pi, p["X"].shape[0], p["X"].shape[1],
                                                                    p["kernel"], p["adaptive"]))
        o = oracle.gwr_run(p["coords"], p["X"], p["y"], p["kernel"], p["adaptive"], 0, select_bw=True,
                           crit=oracle.CRIT_CV)
        ctx.set_data(p["coords"], p["X"], p["y"])
        g = ctx.gwr_run(p["kernel"], p["adaptive"], 0.0, select_bw=True, crit=capi.CRIT_CV)
        assert g["bandwidth"] == o["bandwidth"], pi
        assert close(g["betas"], o["betas"], 1e-9, 1e-12), pi
        assert res[pi]["bandwidth"] == o["bandwidth"], pi
        assert normwise_rel(res[pi]["betas"], o["betas"]) <= 1e-6, pi


def serve_request(srv, r):
    n = r["X"].shape[0]
    if r["search"]:
        return srv.submit(r["coords"], r["X"], r["y"], capi.BISQUARE, True, 0.0, select_bw=True, crit=capi.CRIT_CV)
    return srv.submit(r["coords"], r["X"], r["y"], capi.BISQUARE, True, float(min(n - 1, 100)))


def single_request(ctx, r, precision):
    n = r["X"].shape[0]
    ctx.set_data(r["coords"], r["X"], r["y"])
    if r["search"]:
        return ctx.gwr_run(capi.BISQUARE, True, 0.0, select_bw=True, crit=capi.CRIT_CV, precision=precision)
    return ctx.gwr_run(capi.BISQUARE, True, float(min(n - 1, 100)), precision=precision)


def test_serving_stream_24_vs_fp64_and_oracle(ctx, oracle):
    say("case: serving stream, 24 requests n <= 20000, mixed server vs FP64 single; 4 vs oracle")
    reqs = synth.serving_stream(24, n_range=(500, 20_000))
    with capi.Server(0, capi.MIXED, max_inflight=24) as srv:
        ids = [serve_request(srv, r) for r in reqs]
        res = [srv.wait(i) for i in ids]
    for ri, (r, s) in enumerate(zip(reqs, res)):
        assert s["status"] == capi.GWM_OK, ri
        g = single_request(ctx, r, capi.FP64)
        assert s["bandwidth"] == g["bandwidth"], ri
        assert normwise_rel(s["betas"], g["betas"]) <= 1e-6, ri
        assert col_rel(s["betas"], g["betas"]) <= 2e-6, ri
    smallest = sorted(range(len(reqs)), key=lambda i: reqs[i]["X"].shape[0])[:4]
    for ri in smallest:
        r = reqs[ri]
        n = r["X"].shape[0]
        say("  oracle: request %d n=%d k=%d search=%d" % (ri, n, r["X"].shape[1], r["search"]))
        if r["search"]:
            o = oracle.gwr_run(r["coords"], r["X"], r["y"], oracle.BISQUARE, True, 0, select_bw=True,
                               crit=oracle.CRIT_CV)
        else:
            o = oracle.gwr_run(r["coords"], r["X"], r["y"], oracle.BISQUARE, True, float(min(n - 1, 100)))
        g = single_request(ctx, r, capi.FP64)
        assert g["bandwidth"] == o["bandwidth"], ri
        assert close(g["betas"], o["betas"], 1e-9, 1e-12), ri
        for key in DIAG_KEYS:
            assert g["diagnostic"][key] == pytest.approx(o["diagnostic"][key], rel=1e-9), (ri, key)
