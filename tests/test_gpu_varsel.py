"""GPU parity of the independent-variable selection (GWRBasic.cpp:50-72,
309-346; VariableForwardSelector.cpp:9-77) through the C ABI.

The reference's KATs (testGWRBasic.cpp:139-240) pin the criterion list at
1e-8 absolute and the selected columns exactly; other cases compare with the
oracle at 1e-10 relative (the criterion is an AICc of order 1e3).
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu

VARSEL_KAT = [([2], 2551.61359020599), ([3], 2551.30032201349), ([1], 2468.93236280013),
              ([1, 3], 2452.86447942033), ([1, 2], 2450.59642666509), ([1, 2, 3], 2452.80388934625)]


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


def test_kat_londonhp100(ctx, londonhp100):
    say("case: variable selection KAT londonhp100")
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    sel, models = ctx.indep_vars_select(threshold=3.0)
    assert sel == [1, 3]
    assert len(models) == len(VARSEL_KAT)
    for (v, c), (kv, kc) in zip(models, VARSEL_KAT):
        assert v == kv
        assert abs(c - kc) <= 1e-8


def test_criterion_vs_oracle(ctx, oracle, londonhp100):
    X, y = londonhp100["X"], londonhp100["y"]
    ctx.set_data(londonhp100["coords"], X, y)
    for vars in ([1], [3, 1], [1, 2, 3]):
        say("case: criterion columns %s" % vars)
        assert ctx.indep_vars_criterion(vars) == pytest.approx(oracle.indep_vars_criterion(X, y, vars), rel=1e-12)


def test_singular_design_is_dbl_max(ctx, londonhp100):
    say("case: singular design (a zero column)")
    X = londonhp100["X"]
    Xd = np.column_stack([X, np.zeros(len(X))])
    ctx.set_data(londonhp100["coords"], Xd, londonhp100["y"])
    assert ctx.indep_vars_criterion([1, 4]) == sys.float_info.max
    assert ctx.indep_vars_criterion([1, 2]) < sys.float_info.max
    sel, models = ctx.indep_vars_select(threshold=3.0)
    assert len(models) == 4 * 5 // 2
    assert all((c == sys.float_info.max) == (4 in v) for v, c in models)


@pytest.mark.parametrize("has_intercept", [True, False])
def test_select_vs_oracle_larger(ctx, oracle, has_intercept):
    d = synth.uniform_square(n=20000)
    rng = np.random.default_rng(31)
    X = np.column_stack([d["X"], rng.normal(size=(20000, 4))])   # 4 pure-noise columns
    say("case: selection n=20000 k=%d intercept=%s" % (X.shape[1], has_intercept))
    ctx.set_data(d["coords"], X, d["y"])
    sel, models = ctx.indep_vars_select(threshold=3.0, has_intercept=has_intercept)
    osel, omodels = oracle.indep_vars_select(X, d["y"], threshold=3.0, has_intercept=has_intercept)
    assert sel == osel
    assert [v for v, c in models] == [v for v, c in omodels]
    for (v, c), (ov, oc) in zip(models, omodels):
        assert c == pytest.approx(oc, rel=1e-10)
