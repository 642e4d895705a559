"""The adaptive-bandwidth neighbour table (gwm_hip_set_neighbour_table) on the
GPU through the C ABI.

BandwidthWeight::weight (BandwidthWeight.cpp:40-55) resolves an adaptive
bandwidth from the floor(bw)-th and floor(bw)+1-th smallest distances of each
focus point.  From the second such select of a problem, the context answers
every bandwidth with floor(bw) < K from a table of each point's K smallest
selection keys in ascending order, built once.  The table holds the very
order statistics the select finds, so thresholds -- and every search, fit and
backfit built on them -- must be bit-identical with the table on and off, and
equal to the oracle's.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu

CAP = 16 * 2 ** 30


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


def table_k(n):
    K = 8192
    while K >= n:
        K //= 2
    return K


def thresholds(ctx, coords, X, y, dist, bws, cap, prune=True):
    ctx.set_neighbour_table(cap)
    ctx.set_data(coords, X, y, dist)   # fresh problem: empty bandwidth cache, no table
    ctx.set_pruning(prune)
    ctx.reset_timers()
    ctx.set_timing(True)
    out = [ctx.adaptive_bandwidths(bw) for bw in bws]
    ctx.set_timing(False)
    counts = {nm: ctx.kernel_time(nm)[0] for nm in ("select", "select_lookup", "nbr_table")}
    ctx.set_pruning(True)
    return out, counts


def bandwidth_list(n, K):
    return [1.0, 1.5, 2.0, 2.75, 7.0, 31.4, 100.0, K / 3 + 0.3, K - 1.0, K - 0.5, float(K), K + 0.5, n - 1.0,
            n - 0.25, float(n), 1.5 * n]


@pytest.mark.parametrize("n,dist,prune", [(3000, capi.EUCLIDEAN, True), (20011, capi.EUCLIDEAN, True),
                                          (3000, capi.GREAT_CIRCLE, True), (12000, capi.GREAT_CIRCLE, True),
                                          (3000, capi.EUCLIDEAN, False)])
def test_thresholds_equal_select(ctx, n, dist, prune):
    say("case: neighbour table n=%d dist=%d prune=%d" % (n, dist, prune))
    rng = np.random.default_rng(n + dist)
    if dist == capi.EUCLIDEAN:
        coords = np.round(rng.uniform(0, 40, (n, 2)), 1)   # many exact ties
        coords[7] = coords[8]
        coords[100:120] = coords[99]                     # a stack of coincident points
    else:
        coords = np.column_stack([rng.uniform(-10, 30, n), rng.uniform(35, 60, n)])
        coords[100:110] = coords[99]
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = rng.normal(size=n)
    K = table_k(n)
    bws = bandwidth_list(n, K)
    off, c_off = thresholds(ctx, coords, X, y, dist, bws, 0, prune)
    on, c_on = thresholds(ctx, coords, X, y, dist, bws, CAP, prune)
    eligible = sum(1 for b in bws if np.floor(b) <= K - 1 and b < n)
    assert c_off["nbr_table"] == 0 and c_off["select_lookup"] == 0
    assert c_on["nbr_table"] == 1
    assert c_on["select_lookup"] == eligible - 1
    assert c_on["select"] == len(bws) - eligible + 1
    for bw, a, b in zip(bws, off, on):
        assert np.array_equal(a, b), "bw=%r" % bw


def test_thresholds_match_oracle(ctx, oracle):
    """Lookups against the oracle's sorted-distance thresholds, ties included."""
    say("case: neighbour table vs oracle")
    n = 2500
    rng = np.random.default_rng(5)
    coords = np.round(rng.uniform(0, 25, (n, 2)), 1)
    coords[40:60] = coords[39]
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    bws = [3.0, 18.0, 19.5, 21.0, 150.25, 1023.0, 1023.9]
    on, counts = thresholds(ctx, coords, X, rng.normal(size=n), capi.EUCLIDEAN, bws, CAP)
    assert counts["select_lookup"] == len(bws) - 1
    focus = list(range(30, 70)) + list(range(0, n, 61))
    for bw, g in zip(bws, on):
        o = np.array([oracle.adaptive_threshold(coords, bw, f) for f in focus])
        assert np.array_equal(g[focus], o), "bw=%r" % bw


def test_searches_identical(ctx):
    """A golden-section CV search over adaptive bandwidths and a multiscale
    backfit, table on and off: bandwidths, betas and criteria bit for bit."""
    say("case: neighbour table, GWRBasic search and GWRMultiscale backfit")
    d = synth.multiscale_cosines(n=4000)
    k = d["X"].shape[1]
    res = {}
    for cap in (0, CAP):
        ctx.set_neighbour_table(cap)
        ctx.set_data(d["coords"], d["X"], d["y"])
        ctx.reset_timers()
        ctx.set_timing(True)
        g = ctx.gwr_run(capi.BISQUARE, True, 0.0, select_bw=True, crit=capi.CRIT_CV, precision=capi.FP64, lower=20,
                        upper=3000)
        args = ([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k, [0] + [1] * (k - 1),
                [1e-5] * k)
        m = ctx.mgwr_run(*args, crit_type=0, crit_threshold=1e-5, precision=capi.MIXED)
        ctx.set_timing(False)
        res[cap] = (g, m, ctx.kernel_time("nbr_table")[0], ctx.kernel_time("select_lookup")[0])
    (g0, m0, b0, l0), (g1, m1, b1, l1) = res[0], res[CAP]
    assert b0 == 0 and l0 == 0 and b1 == 1 and l1 > 10
    assert g0["bandwidth"] == g1["bandwidth"]
    for key in ("betas", "betasSE", "shat"):
        assert np.array_equal(g0[key], g1[key]), key
    assert np.array_equal(m0["bandwidths"], m1["bandwidths"])
    assert np.array_equal(m0["betas"], m1["betas"])
    assert m0["iterations"] == m1["iterations"]


def test_cap_and_release(ctx):
    """A cap below the table's size never builds it; lowering the cap below a
    built table's size releases it, after which selects answer again."""
    say("case: neighbour table cap")
    n = 3000
    rng = np.random.default_rng(9)
    coords = rng.uniform(0, 10, (n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = rng.normal(size=n)
    bws = [10.0, 20.0, 30.0]
    _, c = thresholds(ctx, coords, X, y, capi.EUCLIDEAN, bws, 64 * n * 8 - 1)
    assert c["nbr_table"] == 0 and c["select"] == 3
    # 64 * n * 8 bytes fits exactly a table of K = 64 rows
    a, c = thresholds(ctx, coords, X, y, capi.EUCLIDEAN, bws, 64 * n * 8)
    assert c["nbr_table"] == 1 and c["select_lookup"] == 2
    ctx.set_neighbour_table(1e3)
    ctx.reset_timers()
    ctx.set_timing(True)
    b = ctx.adaptive_bandwidths(40.0)
    ctx.set_timing(False)
    assert ctx.kernel_time("select")[0] == 1 and ctx.kernel_time("select_lookup")[0] == 0
    ctx.set_neighbour_table(CAP)
    with pytest.raises(capi.GwmError):
        ctx.set_neighbour_table(-1.0)
    assert b.shape == (n,) and np.all(b >= 0)
