"""The autotuner of sweep launch shapes changes no result bit: every
candidate shape (one or two focus tiles per wave, tiles in order or by a
stride) sums each focus point's data in the same order.  At n = 3000 the
default grid is small, so the first sweep of each class times every shape."""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
def test_shapes_are_bit_identical(prec):
    d = synth.uniform_square(n=3000)
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"])
        ctx.set_autotune(False)
        say("case: autotune off, precision %d" % prec)
        ref = [ctx.criterion(capi.BISQUARE, True, bw, capi.CRIT_CV, prec) for bw in (60, 90)]
        fit_ref = ctx.fit(capi.GAUSSIAN, False, 8.0, prec)
        ctx.set_autotune(True)
        say("case: autotune on, precision %d" % prec)
        for rep in range(3):
            got = [ctx.criterion(capi.BISQUARE, True, bw, capi.CRIT_CV, prec) for bw in (60, 90)]
            assert got == ref
        for rep in range(3):
            fit = ctx.fit(capi.GAUSSIAN, False, 8.0, prec)
            for key in ("betas", "betasSE_raw", "shat", "qDiag"):
                assert np.array_equal(fit[key], fit_ref[key]), key
    lines = capi.autotune_report()
    say("\n".join(lines))
    assert any("euclidean cv" in l and "n<2^12" in l for l in lines)


def _fit_arrays(fit):
    return [fit[key] for key in ("betas", "betasSE_raw", "shat", "qDiag")]


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("geographic", [False, True])
def test_work_order_is_bit_identical(prec, geographic):
    """Large pruned sweeps (here 16 segments x 157 tiles) run their (tile,
    segment) items longest first when launch shaping is on; every item
    writes its own partial sums, so criteria and fits (hat-mode and qDiag
    sweeps included, with the per-chunk bandwidths of an adaptive qDiag) are
    the same bits as in plain launch order."""
    n = 20000
    if geographic:
        d = synth.sphere_box(n=n)
        dist, fixed_bw, ad_bw = capi.GREAT_CIRCLE, 90.0, 300.0
    else:
        d = synth.uniform_square(n=n)
        dist, fixed_bw, ad_bw = capi.EUCLIDEAN, 2.5, 300.0
    cases = [(capi.GAUSSIAN, False, fixed_bw), (capi.BISQUARE, True, ad_bw)]
    out = {}
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"], dist)
        for on in (False, True, True):
            ctx.set_autotune(on)
            say("case: work order %s, geographic %s, precision %d" % (on, geographic, prec))
            res = []
            for kern, ad, bw in cases:
                res.append(ctx.criterion(kern, ad, bw, capi.CRIT_CV, prec))
                res.append(ctx.criterion(kern, ad, bw, capi.CRIT_AIC, prec))
                res.extend(_fit_arrays(ctx.fit(kern, ad, bw, prec)))
            out.setdefault(on, []).append(res)
    ref = out[False][0]
    for got in out[True]:
        assert len(got) == len(ref)
        for a, b in zip(got, ref):
            assert np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def test_work_order_single_group():
    """n = 150000: 19 data segments, not a multiple of the 8 XCD groups, so
    the order is one longest-first list."""
    d = synth.uniform_square(n=150000)
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"])
        res = {}
        for on in (False, True):
            ctx.set_autotune(on)
            say("case: work order %s, n = 150000" % on)
            res[on] = [ctx.criterion(capi.GAUSSIAN, False, 1.5, capi.CRIT_AIC, prec) for prec in (capi.FP64, capi.MIXED)]
        assert res[True] == res[False]


@pytest.mark.parametrize("geographic", [False, True])
def test_predict_work_order_is_bit_identical(geographic):
    """Prediction sweeps over locations of their own (their own spatial
    order and tile boxes) run longest first too, with the same bits."""
    n, m = 20000, 15000
    if geographic:
        d = synth.sphere_box(n=n)
        dist, cases = capi.GREAT_CIRCLE, [(capi.GAUSSIAN, False, 90.0), (capi.BISQUARE, True, 300.0)]
    else:
        d = synth.uniform_square(n=n)
        dist, cases = capi.EUCLIDEAN, [(capi.GAUSSIAN, False, 2.5), (capi.BISQUARE, True, 300.0)]
    rng = np.random.default_rng(11)
    lo, hi = d["coords"].min(0), d["coords"].max(0)
    loc = rng.uniform(lo, hi, (m, 2))
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"], dist)
        res = {}
        for on in (False, True):
            ctx.set_autotune(on)
            say("case: predict work order %s, geographic %s" % (on, geographic))
            res[on] = [ctx.predict(loc, kern, ad, bw, prec) for kern, ad, bw in cases for prec in (capi.FP64, capi.MIXED)]
        for a, b in zip(res[True], res[False]):
            assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
@pytest.mark.parametrize("geographic", [False, True])
def test_forced_shapes_are_bit_identical(prec, geographic):
    """Every launch shape, forced on every sweep: one or two focus tiles per
    wave, stride or plain order.  CV and AICc criteria and fits (hat-mode and
    qDiag sweeps), Euclidean and great circle, k = 8 (three column tiles: the
    great-circle CV and qDiag sweeps default to one tile per wave, the hat
    sweep to two), equal bit for bit."""
    n = 3000
    if geographic:
        d = synth.sphere_box(n=n)
        dist, bw, ad = capi.GREAT_CIRCLE, 400.0, 150.0
    else:
        d = synth.sphere_box(n=n)
        d["coords"] = d["coords"] * 10.0   # planar coordinates, same design
        dist, bw, ad = capi.EUCLIDEAN, 40.0, 150.0
    with capi.Context(0) as ctx:
        ctx.set_data(d["coords"], d["X"], d["y"], dist)
        runs = []
        for shape in range(4):
            say("case: forced shape %d precision %d geographic %d" % (shape, prec, geographic))
            ctx.force_shape(shape)
            crit = [ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_CV, prec),
                    ctx.criterion(capi.BISQUARE, True, ad, capi.CRIT_AIC, prec)]
            fits = [_fit_arrays(ctx.fit(capi.GAUSSIAN, False, bw, prec)),
                    _fit_arrays(ctx.fit(capi.BISQUARE, True, ad, prec))]
            runs.append((crit, fits))
        for crit, fits in runs[1:]:
            assert crit == runs[0][0]
            for f, f0 in zip(fits, runs[0][1]):
                for a, b in zip(f, f0):
                    assert np.array_equal(a, b)
