"""The FP64 distance table of certified searches (gwm_hip_set_certify_table)
on the GPU through the C ABI.

A mixed-precision golden-section search re-takes its uncertain decisions on
FP64 criteria (BandwidthSelector.cpp:6-81 takes every decision in FP64).  On
great-circle data at fixed bandwidths those certification sweeps read the
FP64 squared distances from a table built by the search's first one, for
every pair the bracket's upper end reaches.  The table holds the values the
FP64 sweep forms itself, so the search must return the same bits with the
table on and off, and its FP64-evaluated candidates must equal the FP64
path's criteria at those bandwidths bit for bit.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.set_certify_table(48 * 2 ** 30)
    c.close()


def run(ctx, cap, kernel, crit, precision, lo, hi):
    ctx.set_certify_table(cap)
    ctx.reset_timers()
    ctx.set_timing(True)
    r = ctx.gwr_run(kernel, False, 0.0, select_bw=True, crit=crit, precision=precision, lower=lo, upper=hi)
    ctx.set_timing(False)
    return r, ctx.kernel_time("f64_table")[0]


@pytest.mark.parametrize("n,kernel,crit", [(20011, capi.GAUSSIAN, capi.CRIT_AIC), (20011, capi.GAUSSIAN, capi.CRIT_CV),
                                           (8000, capi.BISQUARE, capi.CRIT_CV),
                                           (8000, capi.EXPONENTIAL, capi.CRIT_AIC),
                                           (6000, capi.TRICUBE, capi.CRIT_CV)])
def test_certified_search_same_bits(ctx, n, kernel, crit):
    say("case: certify table n=%d kernel=%d crit=%d" % (n, kernel, crit))
    d = synth.sphere_box(n=n)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    lo, hi = synth.SPHERE_BOX_BW_KM
    if kernel in (capi.BISQUARE, capi.TRICUBE):
        lo = 150.0
    off, b_off = run(ctx, 0, kernel, crit, capi.MIXED, lo, hi)
    on, b_on = run(ctx, 48 * 2 ** 30, kernel, crit, capi.MIXED, lo, hi)
    f64, _ = run(ctx, 48 * 2 ** 30, kernel, crit, capi.FP64, lo, hi)
    assert b_off == 0 and b_on == 1
    assert on["bandwidth"] == off["bandwidth"] == f64["bandwidth"]
    assert on["trace"] == off["trace"]
    for key in ("betas", "betasSE", "shat", "qDiag"):
        assert np.array_equal(on[key], off[key]), key
    # the certified candidates are the FP64 path's criteria, bit for bit
    ref = dict(f64["trace"])
    exact = [x for x, v in on["trace"] if ref.get(x) == v]
    assert len(exact) >= 2
    assert on["bandwidth"] in exact
    # the others are mixed-precision values within the search's bound (3 tol |f| for CV)
    tol = 3e-7 if crit == capi.CRIT_CV else 1e-7
    for x, v in on["trace"]:
        assert v == pytest.approx(ref[x], rel=tol), x


def test_table_too_large_falls_back(ctx):
    """A cap below every table the search would need: no table, same bits."""
    say("case: certify table cap too small")
    d = synth.sphere_box(n=6000, seed=5)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    lo, hi = synth.SPHERE_BOX_BW_KM
    small, b_small = run(ctx, 1e6, capi.GAUSSIAN, capi.CRIT_AIC, capi.MIXED, lo, hi)
    big, b_big = run(ctx, 48 * 2 ** 30, capi.GAUSSIAN, capi.CRIT_AIC, capi.MIXED, lo, hi)
    assert b_small == 0 and b_big == 1
    assert small["trace"] == big["trace"]
    assert np.array_equal(small["betas"], big["betas"])
    with pytest.raises(capi.GwmError):
        ctx.set_certify_table(-1.0)
