"""GPU parity of gwm_hip_predict (GWRBasic::predict, GWRBasic.cpp:130-191)
against the oracle's predictSerial restatement, through the C ABI.

Tolerances: FP64 betas within 1e-9 |beta| + 1e-9 max_locations |beta_a| of the
oracle (a location away from the data can have a local design of condition
~1e4, where the order of the weighted sums alone moves small components by
more than 1e-9 of themselves; the column-scaled term stays ~1e3 above that
noise); mixed precision within 1e-6 of the FP64 betas (max |delta| /
max(1, |beta|)); at the data coordinates the prediction equals the fit's
betas; pruning on and off bit-identical.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def close_cols(a, b, tol=1e-9):
    """|a - b| <= tol |b| + tol * (largest |b| of the column)."""
    a, b = np.asarray(a), np.asarray(b)
    bound = tol * np.abs(b) + tol * np.abs(b).max(0)[None, :]
    say("  worst |a - b| / bound = %.3g" % float(np.max(np.abs(a - b) / bound)))
    return bool(np.all(np.abs(a - b) <= bound))


def mixed_err(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(1.0, np.abs(b))))


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


def locations_in(coords, m, seed, pad=0.1):
    rng = np.random.default_rng(seed)
    lo, hi = coords.min(0), coords.max(0)
    span = hi - lo
    return rng.uniform(lo - pad * span, hi + pad * span, (m, 2))


CASES = [(capi.GAUSSIAN, True, 36), (capi.BISQUARE, False, 5000.0), (capi.TRICUBE, True, 50.5),
         (capi.EXPONENTIAL, False, 3000.0), (capi.BOXCAR, True, 40), (capi.GAUSSIAN, True, 150)]


def test_predict_londonhp100_vs_oracle(ctx, oracle, londonhp100):
    c, X, y = londonhp100["coords"], londonhp100["X"], londonhp100["y"]
    ctx.set_data(c, X, y)
    loc = locations_in(c, 257, seed=21)
    for kern, adaptive, bw in CASES:
        say("case: predict londonhp100 m=257 kernel=%d adaptive=%d bw=%g" % (kern, adaptive, bw))
        o = oracle.predict(c, X, y, loc, kern, adaptive, bw)
        assert o is not None
        g = ctx.predict(loc, kern, adaptive, bw)
        assert close_cols(g, o)
        gm = ctx.predict(loc, kern, adaptive, bw, precision=capi.MIXED)
        assert mixed_err(gm, g) <= 1e-6


def test_predict_at_data_points_is_fit(ctx, londonhp100):
    c, X, y = londonhp100["coords"], londonhp100["X"], londonhp100["y"]
    ctx.set_data(c, X, y)
    for kern, adaptive, bw in CASES[:3]:
        say("case: predict at data points kernel=%d adaptive=%d bw=%g" % (kern, adaptive, bw))
        f = ctx.fit(kern, adaptive, bw, want_qdiag=False)["betas"]
        g = ctx.predict(c, kern, adaptive, bw)
        assert close_cols(g, f, 1e-12)


def test_predict_more_locations_than_data_and_single(ctx, oracle):
    d = synth.uniform_square(n=300)
    ctx.set_data(d["coords"], d["X"], d["y"])
    loc = locations_in(d["coords"], 2500, seed=22)
    say("case: predict uniform_square n=300 m=2500 adaptive gaussian 60")
    o = oracle.predict(d["coords"], d["X"], d["y"], loc, capi.GAUSSIAN, True, 60)
    g = ctx.predict(loc, capi.GAUSSIAN, True, 60)
    assert close_cols(g, o)
    say("case: predict m=1")
    g1 = ctx.predict(loc[7:8], capi.GAUSSIAN, True, 60)
    assert np.array_equal(g1[0], g[7])


def test_predict_great_circle_vs_oracle(ctx, oracle):
    d = synth.sphere_box(n=3000)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    loc = locations_in(d["coords"], 700, seed=23, pad=0.0)
    for kern, adaptive, bw in [(capi.GAUSSIAN, True, 80), (capi.BISQUARE, False, 400.0)]:
        say("case: predict great circle n=3000 m=700 kernel=%d adaptive=%d bw=%g" % (kern, adaptive, bw))
        o = oracle.predict(d["coords"], d["X"], d["y"], loc, kern, adaptive, bw, geographic=True)
        assert o is not None
        g = ctx.predict(loc, kern, adaptive, bw)
        assert close_cols(g, o)
        gm = ctx.predict(loc, kern, adaptive, bw, precision=capi.MIXED)
        assert mixed_err(gm, g) <= 1e-6


def test_predict_pruning_exact(ctx, oracle):
    d = synth.uniform_square(n=20000)
    ctx.set_data(d["coords"], d["X"], d["y"])
    loc = locations_in(d["coords"], 3000, seed=24)
    try:
        for kern, adaptive, bw in [(capi.BISQUARE, True, 120), (capi.GAUSSIAN, False, 4.0)]:
            for prec in (capi.FP64, capi.MIXED):
                say("case: predict pruning n=20000 m=3000 kernel=%d adaptive=%d prec=%d" % (kern, adaptive, prec))
                ctx.set_pruning(True)
                a = ctx.predict(loc, kern, adaptive, bw, precision=prec)
                ctx.set_pruning(False)
                b = ctx.predict(loc, kern, adaptive, bw, precision=prec)
                assert np.array_equal(a, b)
        say("case: predict n=20000 vs oracle (first 400 locations)")
        o = oracle.predict(d["coords"], d["X"], d["y"], loc[:400], capi.BISQUARE, True, 120)
        assert close_cols(ctx.predict(loc[:400], capi.BISQUARE, True, 120), o)
    finally:
        ctx.set_pruning(True)


def test_predict_not_spd(ctx, oracle):
    rng = np.random.default_rng(7)
    n = 40
    coords = rng.uniform(0, 10, (n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = rng.normal(size=n)
    ctx.set_data(coords, X, y)
    loc = np.array([[5.0, 5.0], [500.0, 500.0]])
    say("case: predict not SPD")
    with pytest.raises(capi.GwmError) as e:
        ctx.predict(loc, capi.BISQUARE, False, 20.0)
    assert e.value.status == capi.GWM_ERR_NOT_SPD
