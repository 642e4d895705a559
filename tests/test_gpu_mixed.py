"""Mixed-precision path against the FP64 path (both through the C ABI).

Stated bounds, mixed against FP64:
- beta: per focus point ||db_i|| / ||b_i|| <= 1e-6 (normwise) and, per
  coefficient, max|db| / max|b| <= 1e-6 (column-scaled); element-wise
  |db| / max(|b|, 1e-2 max|b|) <= 1e-4 (a coefficient near zero carries the
  absolute error of its column);
- criteria within 1e-7 relative (AICc) and the certified bound (CV);
- tr S within 1e-6; the raw local SE diag(A^-1 B A^-1) within 5e-6 and qDiag
  within 1e-6 relative, element by element;
- with certified golden-section decisions the selected bandwidth and every
  iterate equal the FP64 search's.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def col_rel(b, ref):
    scale = np.maximum(np.max(np.abs(ref), axis=0), 1e-300)
    return float(np.max(np.max(np.abs(b - ref), axis=0) / scale))


def normwise_rel(b, ref):
    return float(np.max(np.linalg.norm(b - ref, axis=1) / np.linalg.norm(ref, axis=1)))


def elem_rel(b, ref, floor=1e-2):
    scale = np.maximum(np.abs(ref), floor * np.max(np.abs(ref), axis=0))
    return float(np.max(np.abs(b - ref) / scale))


def check_fit(gm, g64):
    """The stated bounds on one fit; returns the measured errors."""
    e = dict(col=col_rel(gm["betas"], g64["betas"]), norm=normwise_rel(gm["betas"], g64["betas"]),
             elem=elem_rel(gm["betas"], g64["betas"]),
             se=float(np.max(np.abs(gm["betasSE_raw"] - g64["betasSE_raw"]) / np.abs(g64["betasSE_raw"]))),
             qdiag=float(np.max(np.abs(gm["qDiag"] - g64["qDiag"]) / np.abs(g64["qDiag"]))),
             shat=float(np.max(np.abs(gm["shat"] - g64["shat"]) / np.abs(g64["shat"]))))
    say("errors: " + "  ".join("%s %.2e" % kv for kv in e.items()))
    assert e["col"] <= 1e-6 and e["norm"] <= 1e-6 and e["elem"] <= 1e-4
    assert e["se"] <= 5e-6 and e["qdiag"] <= 1e-6 and e["shat"] <= 1e-6
    return e


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


MIX_CASES = [
    (capi.GAUSSIAN, False, 12.0),
    (capi.GAUSSIAN, True, 150.0),
    (capi.BISQUARE, True, 120.0),
    (capi.TRICUBE, False, 25.0),
    (capi.EXPONENTIAL, False, 8.0),
    (capi.BOXCAR, False, 30.0),
]


@pytest.mark.parametrize("kernel,adaptive,bw", MIX_CASES)
def test_mixed_criteria_and_fit_euclidean(ctx, kernel, adaptive, bw):
    say("case: mixed euclidean kernel=%d adaptive=%d bw=%g" % (kernel, adaptive, bw))
    d = synth.uniform_square(n=3000)
    ctx.set_data(d["coords"], d["X"], d["y"])
    for crit in [capi.CRIT_CV, capi.CRIT_AIC]:
        f64 = ctx.criterion(kernel, adaptive, bw, crit, capi.FP64)
        mix = ctx.criterion(kernel, adaptive, bw, crit, capi.MIXED)
        assert mix == pytest.approx(f64, rel=1e-7), (crit, mix, f64)
    check_fit(ctx.fit(kernel, adaptive, bw, capi.MIXED), ctx.fit(kernel, adaptive, bw, capi.FP64))


def test_mixed_great_circle_sphere_box(ctx):
    say("case: mixed great-circle sphere_box n=6000")
    d = synth.sphere_box(n=6000)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    # at n = 6000 over the 40 x 25 degree box a 60 km bandwidth leaves ~8
    # effective neighbours for k = 8 (near-singular local designs); bandwidths
    # here keep the designs conditioned, the bench reports the n = 1e5 error
    for bw in [150.0, 400.0, 1500.0]:
        f64 = ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_AIC, capi.FP64)
        mix = ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_AIC, capi.MIXED)
        assert mix == pytest.approx(f64, rel=1e-7), (bw, mix, f64)
        g64 = ctx.fit(capi.GAUSSIAN, False, bw, capi.FP64)
        gm = ctx.fit(capi.GAUSSIAN, False, bw, capi.MIXED)
        assert col_rel(gm["betas"], g64["betas"]) < 1e-6, bw


def test_mixed_great_circle_cv_se_qdiag(ctx):
    """Great circle at the n = 100,000 workload's selected bandwidth, on a
    20,000-point sample of the same box: CV criteria (one focus tile per wave
    by default for k = 8), and the fit's SE and qDiag (the w^2-weighted hat
    and qDiag sweeps), against FP64."""
    say("case: mixed great-circle CV / SE / qDiag sphere_box n=20000")
    d = synth.sphere_box(n=20000)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    for bw in [140.0, 400.0]:
        f64 = ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_CV, capi.FP64)
        mix = ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_CV, capi.MIXED)
        assert abs(mix - f64) <= ctx.last_error_bound() <= 1e-6 * f64, (bw, mix, f64)
        check_fit(ctx.fit(capi.GAUSSIAN, False, bw, capi.MIXED), ctx.fit(capi.GAUSSIAN, False, bw, capi.FP64))
    check_fit(ctx.fit(capi.BISQUARE, True, 200.0, capi.MIXED), ctx.fit(capi.BISQUARE, True, 200.0, capi.FP64))


def test_certified_search_matches_fp64(ctx):
    say("case: certified mixed search sphere_box n=4000")
    d = synth.sphere_box(n=4000)
    lo, hi = synth.SPHERE_BOX_BW_KM
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    b64, t64 = ctx.select_bandwidth(capi.GAUSSIAN, False, capi.CRIT_AIC, capi.FP64, lo, hi)
    bm, tm = ctx.select_bandwidth(capi.GAUSSIAN, False, capi.CRIT_AIC, capi.MIXED, lo, hi)
    assert bm == b64
    assert [x for x, _ in tm] == [x for x, _ in t64]
    say("case: certified mixed search uniform_square adaptive bisquare CV n=2000")
    d = synth.uniform_square(n=2000)
    ctx.set_data(d["coords"], d["X"], d["y"])
    b64, t64 = ctx.select_bandwidth(capi.BISQUARE, True, capi.CRIT_CV, capi.FP64, 20, 2000)
    bm, tm = ctx.select_bandwidth(capi.BISQUARE, True, capi.CRIT_CV, capi.MIXED, 20, 2000)
    assert bm == b64
    assert [x for x, _ in tm] == [x for x, _ in t64]


def test_mixed_mgwr_bandwidths_match(ctx, londonhp100):
    say("case: mixed MGWR londonhp100")
    ctx.set_data(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"])
    r = ctx.mgwr_run([capi.BISQUARE] * 3, [1] * 3, [0] * 3, [0] * 3, [capi.CRIT_CV] * 3, [0, 1, 1], [1e-5] * 3,
                     precision=capi.MIXED)
    assert list(r["bandwidths"]) == [35, 98, 98]
    assert abs(r["diagnostic"]["RSquare"] - 0.757377391669) < 1e-6


def test_certified_searches_match_fp64_on_uneven_batch(ctx):
    """Certified mixed searches return the FP64 search's bandwidth on the
    first 96 problems of the uneven batch workload, for CV and for AICc
    (whose value crosses zero on some of them)."""
    say("case: certified mixed vs FP64 searches, 96 uneven-batch problems x {CV, AICc}")
    probs = synth.uneven_batch(512)[:96]
    differ = []
    for pi, p in enumerate(probs):
        ctx.set_data(p["coords"], p["X"], p["y"])
        for crit in (capi.CRIT_CV, capi.CRIT_AIC):
            b64, _ = ctx.select_bandwidth(p["kernel"], p["adaptive"], crit, capi.FP64)
            bmx, _ = ctx.select_bandwidth(p["kernel"], p["adaptive"], crit, capi.MIXED)
            if b64 != bmx:
                differ.append((pi, crit, b64, bmx))
    say("  differing: %s" % differ)
    assert not differ
