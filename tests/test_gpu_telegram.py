"""Messages, progress and per-variable distances through the C ABI.

- The ITelegram messages of a bandwidth search (GWRBasic.cpp:76-102,
  :477-484): for the londonhp100 CV search that selects 67
  (test/testGWRBasic.cpp:105-137) the "#bandwidth-criterion <bw>,<value>"
  messages come in the order BandwidthSelector::optimize
  (BandwidthSelector.cpp:6-81) evaluates the candidates, replayed here on the
  oracle's criterion values, with the stage tags around them.
- Variable selection (GWRBasic.cpp:50-72, :338-342) and GWRMultiscale
  (GWRMultiscale.cpp:105-175, :208-317) report their tags in the reference's
  order.
- GWRMultiscale with one distance per variable (Euclidean and great circle
  mixed) against the oracle, FP64 bit-exact bandwidths and betas within 1e-9.
"""
import math
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Recorder:
    def __init__(self):
        self.messages = []
        self.progress = []

    def print(self, level, message):
        self.messages.append((level, message))

    def prog(self, current, total):
        self.progress.append((current, total))

    def texts(self):
        return [m for _, m in self.messages]


@pytest.fixture
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def golden_order(lower, upper, f):
    """Candidates in the order BandwidthSelector::optimize (BandwidthSelector.cpp:6-81)
    evaluates them for an adaptive bandwidth, f giving the criterion."""
    xU, xL, eps, R = upper, lower, 1e-4, (math.sqrt(5) - 1) / 2
    d = R * (xU - xL)
    x1, x2 = math.floor(xL + d), math.floor(xU - d + 0.5)   # floor, std::round (half away from zero)
    seq = [x1, x2]
    f1, f2 = f(x1), f(x2)
    d1, it = f2 - f1, 0
    while abs(d) > eps and abs(d1) > eps and it < 100:
        d = R * d
        if f1 < f2:
            xL, x2, f2 = x2, x1, f1
            x1 = math.floor(xL + d + 0.5)
            seq.append(x1)
            f1 = f(x1)
        else:
            xU, x1, f1 = x1, x2, f2
            x2 = math.floor(xU - d)
            seq.append(x2)
            f2 = f(x2)
        it += 1
        d1 = f2 - f1
    return seq


def parse_criterion(msg):
    assert msg.startswith("#bandwidth-criterion ")
    bw, value = msg[len("#bandwidth-criterion "):].split(",")
    return bw, float(value)


def test_bandwidth_search_messages_follow_the_reference_order(ctx, londonhp100, oracle):
    say("case: telegram of the londonhp100 CV search (bw = 67), FP64")
    rec = Recorder()
    ctx.set_telegram(rec.print, rec.prog)
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    g = ctx.gwr_run(capi.GAUSSIAN, True, 0.0, select_bw=True, crit=capi.CRIT_CV)
    assert g["bandwidth"] == 67
    texts = rec.texts()
    assert texts[:2] == ["#stage Bandwidth selection", "#bandwidth-criterion adaptive,criterion"]
    assert texts[-2:] == ["#stage Model fitting", "#stage Model Diagnostic"]
    crit = texts[2:-2]
    assert all(lv == capi.LOG_INFO for lv, _ in rec.messages)
    # order: replay the reference's loop on the oracle's values
    cache = {}

    def f(bw):
        if bw not in cache:
            cache[bw] = oracle.criterion(londonhp100["coords"], londonhp100["X"], londonhp100["y"], oracle.GAUSSIAN,
                                         True, bw, oracle.CRIT_CV)
        return cache[bw]

    order = golden_order(20.0, 100.0, f)
    assert [int(b) for b, _ in map(parse_criterion, crit)] == [int(x) for x in order]
    for (b, v), x in zip(map(parse_criterion, crit), order):
        assert b == str(int(x))                           # std::to_string of an integer bandwidth
        assert v == pytest.approx(cache[x], rel=1e-11, abs=2e-6)   # "%f" of the value
    # progress: exp(-|last - value|) per evaluation (the first against DBL_MAX), then the fit
    fr = [p for p in rec.progress if p[1] == 0.0]
    assert len(fr) == len(crit)
    assert fr[0][0] == 0.0
    for q in range(1, len(fr)):
        assert fr[q][0] == pytest.approx(math.exp(-abs(cache[order[q - 1]] - cache[order[q]])), rel=1e-6, abs=1e-300)
    assert rec.progress[-1] == (100.0, 100.0)
    # the criterion list holds the distinct candidates
    assert sorted(set(int(x) for x in order)) == [int(x) for x, _ in g["trace"]]


def test_bandwidth_search_messages_mixed(ctx, londonhp100):
    say("case: telegram of the londonhp100 CV search, mixed precision")
    rec = Recorder()
    ctx.set_telegram(rec.print, rec.prog)
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    g = ctx.gwr_run(capi.GAUSSIAN, True, 0.0, select_bw=True, crit=capi.CRIT_CV, precision=capi.MIXED)
    assert g["bandwidth"] == 67
    crit = [parse_criterion(m) for m in rec.texts() if m.startswith("#bandwidth-criterion ") and "criterion" != m[-9:]]
    assert {int(b) for b, _ in crit} == {int(x) for x, _ in g["trace"]}
    ctx.set_telegram(None, None)
    n0 = len(rec.messages)
    ctx.gwr_run(capi.GAUSSIAN, True, 67.0)
    assert len(rec.messages) == n0


def test_variable_selection_messages(ctx, londonhp100):
    say("case: telegram of the londonhp100 variable selection")
    rec = Recorder()
    ctx.set_telegram(rec.print, rec.prog)
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    sel, models = ctx.indep_vars_select(True, 3.0)
    texts = rec.texts()
    assert texts[:2] == ["#stage Independent variable selection", "#variable-criterion variables,criterion"]
    body = texts[2:]
    assert len(body) == 6 and len(rec.progress) == 6
    assert [p for p in rec.progress] == [(float(q + 1), 6.0) for q in range(6)]
    got = {}
    for m in body:
        vars_s, v = m[len("#variable-criterion "):].split(",")
        got[tuple(sorted(int(t) for t in vars_s.split("+")))] = float(v)
    for vars_, value in models:
        assert got[tuple(sorted(vars_))] == pytest.approx(value, rel=1e-11, abs=2e-6)
    # the first step tries the columns in order (VariableForwardSelector.cpp:15-24)
    assert [m.split(",")[0] for m in body[:3]] == ["#variable-criterion 1", "#variable-criterion 2",
                                                   "#variable-criterion 3"]


def test_multiscale_messages(ctx, londonhp100):
    say("case: telegram of a londonhp100 GWRMultiscale run (adaptive bisquare CV, no hat matrix)")
    rec = Recorder()
    ctx.set_telegram(rec.print, rec.prog)
    ctx.set_data(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"])
    r = ctx.mgwr_run([capi.BISQUARE] * 3, [1] * 3, [0.0] * 3, [0] * 3, [capi.CRIT_CV] * 3, [0, 1, 1], [1e-5] * 3)
    assert list(r["bandwidths"]) == [35, 98, 98]
    t = rec.texts()
    assert t[0] == "#stage Calculating initial bandwidths"
    for i in range(3):
        a = t.index("#initial-bandwidth %d" % i)
        b = t.index("#initial-bandwidth %d,%s" % (i, "%f" % r_initial(t, i)))
        assert a < b and all(m.startswith("#bandwidth-criterion ") for m in t[a + 1:b])
    s = t.index("#stage Calculating initial bandwidth")
    assert t[s + 1] == "#bandwidth-criterion adaptive,criterion"
    m = t.index("#stage Model fitting")
    assert t[m + 1] == "#back-fitting Model fitting with inital bandwidth"
    assert t[m + 2] == "#back-fitting Selecting the optimum bandwidths for each independent variable"
    iters = [q for q, x in enumerate(t) if x.startswith("#back-fitting #iteration ")]
    assert len(iters) == r["iterations"]
    assert t[iters[0]] == "#back-fitting #iteration 1"
    crits = [x for x in t if x.startswith("#back-fitting #backfitting-criterion ")]
    assert len(crits) == r["iterations"]
    assert t[-2:] == ["#back-fitting Finished", "#stage Model Diagnostic"]
    sel = [x for x in t if x.startswith("#back-fitting #variable-bandwidth-selection ")]
    # "<i>,<old>,<new>,<|new - old|>,<false | true | count,left>" after each search
    done = [x[len("#back-fitting #variable-bandwidth-selection "):].split(",") for x in sel if x.count(",") >= 4]
    for i in range(3):
        f = [q for q in done if int(q[0]) == i][-1]
        assert float(f[2]) == r["bandwidths"][i]
        assert f[4] in ("false", "true") or int(f[4]) + int(f[5]) == 5
    pct = [p for p, tot in rec.progress if tot == 0.0]
    assert all(0.0 <= p <= 1.0 for p in pct)


def r_initial(t, i):
    """The bandwidth a '#initial-bandwidth i,<bw>' message reports."""
    for x in t:
        if x.startswith("#initial-bandwidth %d," % i):
            return float(x.split(",")[1])
    raise AssertionError("no initial bandwidth message for variable %d" % i)


@pytest.mark.parametrize("dists", [[0, 1, 0], [1, 0, 1]])
def test_multiscale_mixed_distances_vs_oracle(ctx, oracle, dists):
    say("case: GWRMultiscale with per-variable distances %s, n=400, FP64 and mixed vs oracle" % dists)
    rng = np.random.default_rng(5)
    n = 400
    lon, lat = rng.uniform(100.0, 112.0, n), rng.uniform(20.0, 32.0, n)
    coords = np.column_stack([lon, lat])
    X = np.column_stack([np.ones(n), rng.normal(size=n), rng.normal(size=n)])
    y = 1.0 + np.sin((lon - 100) / 3) * X[:, 1] + np.cos((lat - 20) / 4) * X[:, 2] + 0.2 * rng.normal(size=n)
    k = 3
    args = ([capi.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [capi.CRIT_CV] * k, [0, 1, 1], [1e-5] * k)
    o = oracle.mgwr_run(coords, X, y, [oracle.BISQUARE] * k, [1] * k, [0.0] * k, [0] * k, [oracle.CRIT_CV] * k,
                        [0, 1, 1], [1e-5] * k, has_hat=False, geographic=dists)
    ctx.set_data(coords, X, y, capi.EUCLIDEAN)
    ctx.mgwr_set_distances(dists)
    g = ctx.mgwr_run(*args)
    assert np.array_equal(g["bandwidths"], o["bandwidths"])
    assert g["iterations"] == o["iterations"]
    assert np.all(np.abs(g["betas"] - o["betas"]) <= 1e-9 * np.abs(o["betas"]) + 1e-12)
    gm = ctx.mgwr_run(*args, precision=capi.MIXED)
    assert np.array_equal(gm["bandwidths"], g["bandwidths"])
    ref = np.max(np.abs(g["betas"]), axis=0)
    assert np.max(np.max(np.abs(gm["betas"] - g["betas"]), axis=0) / ref) <= 1e-6
    # the context keeps the data's distance for single-scale calls afterwards
    ctx.mgwr_set_distances(None)
    assert ctx.criterion(capi.BISQUARE, True, 50, capi.CRIT_CV) == pytest.approx(
        oracle.criterion(coords, X, y, oracle.BISQUARE, True, 50, oracle.CRIT_CV), rel=1e-11)
