"""GPU multiscale GWR with the hat matrix (GWRMultiscale.cpp:163-179,
215-225, 296-301) through the C ABI: the reference's KATs
(testGWRMultiscale.cpp:45-133, 175-257; AICc/R2/adjR2 at 1e-6, bandwidths
exact or at 0.1) and a larger case against the oracle's hat branch
(tr S0 and tr S0^T S0 within 1e-9 relative, bandwidths exact)."""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu

COMMON = dict(centered=[0, 1, 1], bw_threshold=[1e-5] * 3)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    c.mgwr_set_hat_matrix(True)
    yield c
    c.close()


def run(ctx, d, kernels, adaptive, bws, init, approach, **kw):
    ctx.set_data(d["coords"], d["Xm"], d["y"])
    return ctx.mgwr_run(kernels, adaptive, bws, init, approach, COMMON["centered"], COMMON["bw_threshold"], **kw)


def check_diag(d, aicc, r2, adjr2):
    assert abs(d["AICc"] - aicc) < 1e-6
    assert abs(d["RSquare"] - r2) < 1e-6
    assert abs(d["RSquareAdjust"] - adjr2) < 1e-6


def test_kat_fixed_bisquare_cv_hat(ctx, londonhp100):
    say("case: KAT MGWR fixed bisquare CV with hat (testGWRMultiscale.cpp:45-88)")
    r = run(ctx, londonhp100, [capi.BISQUARE] * 3, [0] * 3, [0] * 3, [0] * 3, [capi.CRIT_CV] * 3)
    bw = r["bandwidths"]
    assert abs(bw[0] - 4623.78) < 0.1 and abs(bw[1] - 12665.70) < 0.1 and abs(bw[2] - 12665.70) < 0.1
    check_diag(r["diagnostic"], 2437.09277417389, 0.744649364494, 0.712344894394)


def test_kat_adaptive_aic_initial_dcvr_hat(ctx, londonhp100):
    say("case: KAT MGWR adaptive AIC initial dCVR with hat (testGWRMultiscale.cpp:90-133)")
    r = run(ctx, londonhp100, [capi.BISQUARE] * 3, [1] * 3, [36] * 3, [1] * 3, [capi.CRIT_AIC] * 3, crit_type=1)
    assert list(r["bandwidths"]) == [45, 98, 98]
    check_diag(r["diagnostic"], 2437.935218705351, 0.7486787930045755, 0.7118919517893492)


def test_kat_adaptive_cv_cvr_hat(ctx, londonhp100):
    say("case: KAT MGWR adaptive CV CVR with hat (testGWRMultiscale.cpp:175-218)")
    r = run(ctx, londonhp100, [capi.BISQUARE] * 3, [1] * 3, [36] * 3, [1] * 3, [capi.CRIT_CV] * 3, crit_type=0)
    assert list(r["bandwidths"]) == [35, 98, 98]
    check_diag(r["diagnostic"], 2438.256543496552, 0.757377391669, 0.715598248225)


def test_kat_bounds_hat(ctx, londonhp100):
    say("case: KAT MGWR bounds with hat (testGWRMultiscale.cpp:220-257)")
    r = run(ctx, londonhp100, [capi.BISQUARE] * 3, [1] * 3, [36] * 3, [0] * 3, [capi.CRIT_CV] * 3, crit_type=0,
            lower=50, upper=100)
    assert list(r["bandwidths"]) == [52, 99, 99]


def test_hat_vs_oracle_n600(ctx, oracle):
    say("case: MGWR with hat vs oracle uniform_square n=600")
    d = synth.uniform_square(n=600)
    Xm = d["X"][:, :3]
    ctx.set_data(d["coords"], Xm, d["y"])
    kw = dict(crit_type=1, lower=None, upper=None)
    g = ctx.mgwr_run([capi.BISQUARE, capi.GAUSSIAN, capi.BISQUARE], [1, 1, 1], [0] * 3, [0] * 3,
                     [capi.CRIT_CV, capi.CRIT_AIC, capi.CRIT_CV], [0, 1, 1], [1e-5] * 3, **kw)
    o = oracle.mgwr_run(d["coords"], Xm, d["y"], [oracle.BISQUARE, oracle.GAUSSIAN, oracle.BISQUARE], [1, 1, 1],
                        [0] * 3, [0] * 3, [oracle.CRIT_CV, oracle.CRIT_AIC, oracle.CRIT_CV], [0, 1, 1], [1e-5] * 3,
                        has_hat=True, crit_type=1)
    assert list(g["bandwidths"]) == list(o["bandwidths"])
    for key in ["ENP", "EDF", "AICc", "RSquareAdjust"]:
        say("  %s gpu %.15g oracle %.15g" % (key, g["diagnostic"][key], o["diagnostic"][key]))
        assert g["diagnostic"][key] == pytest.approx(o["diagnostic"][key], rel=1e-9)
    shat = ctx.mgwr_last_shat()
    assert shat[0] > 3 and shat[1] > 0


def test_hat_off_gives_zero_shat(ctx, londonhp100):
    say("case: MGWR hat switched off")
    ctx.mgwr_set_hat_matrix(False)
    try:
        r = run(ctx, londonhp100, [capi.BISQUARE] * 3, [1] * 3, [0] * 3, [0] * 3, [capi.CRIT_CV] * 3)
        assert list(r["bandwidths"]) == [35, 98, 98]
        assert list(ctx.mgwr_last_shat()) == [0.0, 0.0]
    finally:
        ctx.mgwr_set_hat_matrix(True)
