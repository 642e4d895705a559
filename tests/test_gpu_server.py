"""Batched / continuous-batching server (gwm_hip_server_*) on the device.

- Requests in a batch reproduce the single-problem API bit for bit (both use
  the same data segments and kernels).
- A request returns the same bits alone and mixed with other requests of
  different n, k, kernels and types, in FP64 and in mixed precision.
- A CV bandwidth search run through the server matches the CPU oracle's
  bandwidth and diagnostics.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def problem(seed, n, k):
    rng = np.random.default_rng(seed)
    coords = rng.uniform(0, 100, (n, 2))
    X = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(k - 1)])
    beta = np.column_stack([1 + coords[:, 0] / 50] + [np.sin(coords[:, 1] / (10 + 5 * c)) for c in range(k - 1)])
    y = np.sum(X * beta, axis=1) + rng.normal(0, 0.3, n)
    return coords, X, y


MIX = [
    # seed, n, k, kernel, adaptive, bw, select_bw, crit
    (1, 700, 3, capi.BISQUARE, True, 80.0, True, capi.CRIT_CV),
    (2, 1500, 6, capi.GAUSSIAN, False, 12.0, False, capi.CRIT_CV),
    (3, 300, 2, capi.TRICUBE, True, 60.0, True, capi.CRIT_AIC),
    (4, 2500, 12, capi.GAUSSIAN, True, 200.0, False, capi.CRIT_CV),
    (5, 900, 8, capi.BISQUARE, False, 25.0, True, capi.CRIT_AIC),
]


def submit(srv, case):
    seed, n, k, kern, adaptive, bw, sel, crit = case
    coords, X, y = problem(seed, n, k)
    lo, hi = (20, n) if adaptive else (1.0, 150.0)
    return srv.submit(coords, X, y, kern, adaptive, bw, select_bw=sel, crit=crit, lower=lo if sel else None,
                      upper=hi if sel else None)


def test_server_matches_single_problem_api():
    say("case: server vs single-problem API (fp64)")
    with capi.Server(0, capi.FP64) as srv, capi.Context(0) as ctx:
        ids = [submit(srv, c) for c in MIX]
        res = [srv.wait(i) for i in ids]
        for case, r in zip(MIX, res):
            seed, n, k, kern, adaptive, bw, sel, crit = case
            assert r["status"] == capi.GWM_OK
            coords, X, y = problem(seed, n, k)
            ctx.set_data(coords, X, y)
            lo, hi = (20, n) if adaptive else (1.0, 150.0)
            g = ctx.gwr_run(kern, adaptive, bw, select_bw=sel, crit=crit, lower=lo if sel else None,
                            upper=hi if sel else None)
            assert r["bandwidth"] == g["bandwidth"], case
            assert np.array_equal(r["betas"], g["betas"]), case
            assert np.array_equal(r["betasSE"], g["betasSE"]), case
            assert np.array_equal(r["qDiag"], g["qDiag"]), case
            assert r["diagnostic"]["AICc"] == g["diagnostic"]["AICc"]
        st = srv.stats()
        say("rounds %d solves %d peak inflight %d" % (st["rounds"], st["solves"], st["peak_inflight"]))
        assert st["peak_inflight"] == len(MIX)


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_alone_equals_in_the_mix(precision):
    say("case: alone vs mix precision=%d" % precision)
    target = MIX[0]
    with capi.Server(0, precision) as srv:
        alone = srv.wait(submit(srv, target))
    with capi.Server(0, precision, max_inflight=8) as srv:
        ids = [submit(srv, c) for c in MIX[1:3]] + [submit(srv, target)] + [submit(srv, c) for c in MIX[3:]]
        res = [srv.wait(i) for i in ids]
    mixed = res[2]
    assert mixed["status"] == capi.GWM_OK and alone["status"] == capi.GWM_OK
    assert mixed["bandwidth"] == alone["bandwidth"]
    assert np.array_equal(mixed["betas"], alone["betas"])
    assert np.array_equal(mixed["betasSE"], alone["betasSE"])
    assert np.array_equal(mixed["qDiag"], alone["qDiag"])
    assert mixed["diagnostic"] == alone["diagnostic"]


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_sliced_rounds_equal_whole_rounds(precision):
    """Evaluations cut into one chunk of focus points per round (a tiny round
    budget) give the same bits as evaluations run whole in one round."""
    say("case: sliced vs whole rounds precision=%d" % precision)
    cases = [MIX[0], MIX[3], MIX[4]]   # adaptive CV search, Gaussian fit k=12, fixed AIC search
    out = {}
    for ms in (0.0, 1e-6):
        with capi.Server(0, precision, max_inflight=8) as srv:
            srv.set_round_ms(ms)
            ids = [submit(srv, c) for c in cases]
            out[ms] = [srv.wait(i) for i in ids]
            st = srv.stats()
        say("round_ms %g: rounds %d" % (ms, st["rounds"]))
    for a, b in zip(out[0.0], out[1e-6]):
        assert a["status"] == capi.GWM_OK and b["status"] == capi.GWM_OK
        assert a["bandwidth"] == b["bandwidth"]
        assert np.array_equal(a["betas"], b["betas"])
        assert np.array_equal(a["betasSE"], b["betasSE"])
        assert np.array_equal(a["qDiag"], b["qDiag"])
        assert a["diagnostic"] == b["diagnostic"]


def test_server_search_matches_oracle(oracle):
    say("case: server CV search vs oracle (uniform_square n=800)")
    d = synth.uniform_square(n=800)
    with capi.Server(0, capi.MIXED) as srv:
        r = srv.wait(srv.submit(d["coords"], d["X"], d["y"], capi.BISQUARE, True, 0, select_bw=True,
                                crit=capi.CRIT_CV, lower=20, upper=800))
    o = oracle.gwr_run(d["coords"], d["X"], d["y"], oracle.BISQUARE, True, 0, select_bw=True, crit=oracle.CRIT_CV,
                       lower=20, upper=800)
    assert r["bandwidth"] == o["bandwidth"]
    assert r["diagnostic"]["RSquare"] == pytest.approx(o["diagnostic"]["RSquare"], rel=1e-7)
    scale = np.max(np.abs(o["betas"]), axis=0)
    assert np.max(np.max(np.abs(r["betas"] - o["betas"]), axis=0) / scale) < 1e-6


def test_admission_by_memory_budget():
    say("case: admission under a small memory budget")
    # a budget below two requests' footprint admits them one at a time
    with capi.Server(0, capi.FP64, hbm_budget_bytes=3e6, max_inflight=16) as srv:
        ids = [submit(srv, c) for c in MIX[:3]]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert all(r["status"] == capi.GWM_OK for r in res)
    assert st["peak_inflight"] == 1


@pytest.mark.parametrize("precision", [capi.FP64, capi.MIXED])
def test_partial_sum_slices_equal_single_problem(precision):
    """A partial-sum budget of one byte holds one chunk of 128 focus points
    per request: every phase (main sweep, FP64 recomputation, qDiag) runs as
    many slices, each written at its own offset of a buffer of 128 columns.
    The results must be the single-problem API's bits (mixed: the same
    certified search)."""
    say("case: one-chunk partial-sum slices precision=%d" % precision)
    cases = [MIX[0], MIX[1], MIX[3], MIX[4]]
    with capi.Server(0, precision, max_inflight=8) as srv:
        srv.set_partial_budget(1)
        ids = [submit(srv, c) for c in cases]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    say("rounds %d peak bytes %.3g" % (st["rounds"], st["peak_bytes"]))
    with capi.Context(0) as ctx:
        for case, r in zip(cases, res):
            seed, n, k, kern, adaptive, bw, sel, crit = case
            coords, X, y = problem(seed, n, k)
            ctx.set_data(coords, X, y)
            lo, hi = (20, n) if adaptive else (1.0, 150.0)
            g = ctx.gwr_run(kern, adaptive, bw, select_bw=sel, crit=crit, precision=precision,
                            lower=lo if sel else None, upper=hi if sel else None)
            assert r["status"] == capi.GWM_OK, case
            assert r["bandwidth"] == g["bandwidth"], case
            assert np.array_equal(r["betas"], g["betas"]), case
            assert np.array_equal(r["betasSE"], g["betasSE"]), case
            assert np.array_equal(r["qDiag"], g["qDiag"]), case
            assert r["diagnostic"]["AICc"] == g["diagnostic"]["AICc"], case
    with pytest.raises(capi.GwmError):
        with capi.Server(0, precision) as srv:
            srv.set_partial_budget(0.5)
