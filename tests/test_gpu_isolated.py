"""Focus points whose largest weight is far below 1 (CV, whose own weight is
zeroed, and prediction locations), on the GPU through the C ABI.

Two regimes the reference handles with plain FP64 arithmetic
(BandwidthWeight.h:68-121, GWRBasic.cpp:233-268):
- a point 13-30 bandwidths from every other point: its Gaussian or
  exponential weights lie below e^-40 (the FP64 cut), yet they are all it has, and the
  reference's criterion is finite.  The FP64 path cuts weights relative to
  the nearest point's weight, so value and status match the oracle;
- a bandwidth below a fifth of the point spacing: every non-self weight is
  below 2^-12, which the FP16 split of the mixed path cannot resolve.  Those
  focus points are recomputed in FP64, so the mixed values match FP64 and the
  oracle.

Tolerances: FP64 vs the oracle 1e-11 relative on criteria, 1e-9 on betas;
mixed vs FP64 within the error bound the certified search assumes for the
value (3e-7 relative for CV at the default tolerance), 1e-6 on betas.
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def close(a, b, rtol, atol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b) + atol))


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


def with_outliers(n=2000, seed=1001):
    """The n = 2000 workload on [0,100]^2 plus three points about 30 units
    outside its edges: each has a few dozen points 30-35 units away, enough
    for a well-conditioned local design at Gaussian bandwidth 2."""
    d = synth.uniform_square(n=n, seed=seed)
    out = np.array([[130.0, 50.0], [50.0, 131.0], [-30.0, 40.0]])
    rng = np.random.default_rng(seed)
    m = len(out)
    coords = np.vstack([d["coords"], out])
    X = np.vstack([d["X"], np.column_stack([np.ones(m), rng.normal(size=(m, d["X"].shape[1] - 1))])])
    y = np.concatenate([d["y"], rng.normal(size=m)])
    nn = np.min(np.hypot(out[:, None, 0] - d["coords"][None, :, 0], out[:, None, 1] - d["coords"][None, :, 1]), 1)
    return coords, X, y, nn


@pytest.mark.parametrize("kernel,bw", [(capi.GAUSSIAN, 2.0), (capi.EXPONENTIAL, 0.4)])
def test_isolated_points_cv_match_oracle(ctx, oracle, kernel, bw):
    say("case: isolated points CV kernel=%d bw=%g" % (kernel, bw))
    coords, X, y, nn = with_outliers()
    # the outliers' nearest points lie 13-30 bandwidths away, where every
    # weight is below the FP64 cut e^-40 relative to a weight of 1
    e_nn = -0.5 * (nn / bw) ** 2 if kernel == capi.GAUSSIAN else -nn / bw
    assert np.all(e_nn < -40) and np.all(nn / bw >= 13) and np.all(nn / bw <= 100)
    o = oracle.criterion(coords, X, y, kernel, False, bw, oracle.CRIT_CV)
    assert o < oracle.DBL_MAX
    ctx.set_data(coords, X, y)
    g = ctx.criterion(kernel, False, bw, capi.CRIT_CV)
    assert g == pytest.approx(o, rel=1e-11), (g, o)
    gm = ctx.criterion(kernel, False, bw, capi.CRIT_CV, precision=capi.MIXED)
    cond, refined = ctx.last_condition()
    assert refined >= 3   # at least the outliers went to FP64
    # within the bound the certified search assumes for this value (at least
    # 3e-7 |CV|); exponential weights at small bandwidths carry the FP32
    # rounding of exponents of up to -26
    bound = ctx.last_error_bound()
    assert bound <= 1e-5 * abs(g)
    assert abs(gm - g) <= bound, (gm, g, bound)


def jittered_grid(side=24, spacing=1.0, seed=7):
    """A side x side grid with jitter of 0.5% of the spacing and k = 2
    (intercept + one covariate): below the spacing every local design rests
    on the 2-4 nearest grid neighbours, of nearly equal weight, which
    determine both coefficients (a third covariate would leave corner points
    with two points for three unknowns)."""
    rng = np.random.default_rng(seed)
    u, v = np.meshgrid(np.arange(side) * spacing, np.arange(side) * spacing)
    coords = np.column_stack([u.ravel(), v.ravel()]) + rng.uniform(-0.005, 0.005, (side * side, 2)) * spacing
    n = side * side
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = X @ np.array([1.0, 0.5]) + np.sin(coords[:, 0] / 3) + rng.normal(0, 0.2, n)
    return coords, X, y


def test_bandwidth_below_spacing_cv(ctx, oracle):
    """Fixed Gaussian CV at a fifth and a sixth of the point spacing: the
    nearest weight is about e^-12.5 and e^-18 (below 2^-12) for every point."""
    say("case: CV below the point spacing")
    coords, X, y = jittered_grid()
    ctx.set_data(coords, X, y)
    for bw in [0.2, 1.0 / 6.0]:
        o = oracle.criterion(coords, X, y, capi.GAUSSIAN, False, bw, oracle.CRIT_CV)
        g = ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_CV)
        gm = ctx.criterion(capi.GAUSSIAN, False, bw, capi.CRIT_CV, precision=capi.MIXED)
        _, refined = ctx.last_condition()
        assert refined == len(y)
        assert g == pytest.approx(o, rel=1e-11), (bw, g, o)
        assert gm == g, (bw, gm, g)   # every point recomputed: the FP64 bits


@pytest.mark.parametrize("lower,upper", [(0.1, 0.2), (0.1, 12.0)])
def test_search_through_small_bandwidths(ctx, oracle, lower, upper):
    """Certified mixed CV searches over an interval below a fifth of the
    spacing, and over one that starts there, take the FP64 search's iterates
    and the oracle's bandwidth."""
    say("case: CV search over [%g, %g]" % (lower, upper))
    coords, X, y = jittered_grid(side=20)
    ctx.set_data(coords, X, y)
    bf, tf = ctx.select_bandwidth(capi.GAUSSIAN, False, capi.CRIT_CV, lower=lower, upper=upper)
    bm, tm = ctx.select_bandwidth(capi.GAUSSIAN, False, capi.CRIT_CV, precision=capi.MIXED, lower=lower,
                                  upper=upper)
    o = oracle.gwr_run(coords, X, y, oracle.GAUSSIAN, False, 0, select_bw=True, crit=oracle.CRIT_CV, lower=lower,
                       upper=upper)
    assert bf == o["bandwidth"]
    assert bm == bf
    assert [x for x, _ in tf] == [x for x, _ in o["trace"]]
    for (x, f), (xo, fo) in zip(tf, o["trace"]):
        assert f == pytest.approx(fo, rel=1e-11), (x, f, fo)


def test_isolated_points_fit_and_not_spd_status(ctx, oracle):
    """Status parity where the reference itself fails: bisquare with a
    bandwidth below the outliers' distance leaves them no weight in CV."""
    say("case: isolated points status")
    coords, X, y, nn = with_outliers()
    ctx.set_data(coords, X, y)
    bw = float(nn.min()) * 0.5
    o = oracle.criterion(coords, X, y, capi.BISQUARE, False, bw, oracle.CRIT_CV)
    assert o == oracle.DBL_MAX
    assert ctx.criterion(capi.BISQUARE, False, bw, capi.CRIT_CV) == capi.DBL_MAX
    assert ctx.criterion(capi.BISQUARE, False, bw, capi.CRIT_CV, precision=capi.MIXED) == capi.DBL_MAX


@pytest.mark.parametrize("kernel,bw", [(capi.GAUSSIAN, 2.0), (capi.EXPONENTIAL, 0.4)])
def test_predict_far_locations(ctx, oracle, kernel, bw):
    """Prediction locations 30-40 units from the data (GWRBasic.cpp:167-191):
    15-22 bandwidths for the Gaussian, 77-108 for the exponential, where all
    their weights lie below e^-40, the FP64 cut.  Locations whose local design is well
    conditioned (at most 1e4): the nearer a location's few weighted points
    lie to a line, the more any two FP64 evaluation orders differ."""
    say("case: predict far locations kernel=%d" % kernel)
    d = synth.uniform_square(n=500)
    ctx.set_data(d["coords"], d["X"], d["y"])
    loc = np.array([[50.0, 50.0], [130.0, 50.0], [30.0, 140.0]])
    o = oracle.predict(d["coords"], d["X"], d["y"], loc, kernel, False, bw)
    assert o is not None
    g = ctx.predict(loc, kernel, False, bw)
    assert close(g, o, 1e-9, 1e-12), (g, o)
    gm = ctx.predict(loc, kernel, False, bw, precision=capi.MIXED)
    assert close(gm, g, 1e-6, 1e-9), (gm, g)


def dense_var_cv(coords, x, y, bw):
    """Single-column CV of the multiscale backfit (GWRMultiscale.cpp:560-596)
    with Gaussian weights and no cut: beta_i = sum_{j != i} w x y / sum w x^2."""
    D2 = ((coords[:, None, :] - coords[None, :, :]) ** 2).sum(-1)
    W = np.exp(-0.5 * D2 / (bw * bw))
    np.fill_diagonal(W, 0.0)
    beta = (W @ (x * y)) / (W @ (x * x))
    return float(np.sum((y - x * beta) ** 2))


def test_single_column_cv_isolated(ctx):
    say("case: single-column CV with isolated points")
    coords, X, y, _ = with_outliers(n=1000)
    ctx.set_data(coords, X, y)
    x = X[:, 1]
    ref = dense_var_cv(coords, x, y, 2.0)
    g = ctx.var_criterion(x, y, capi.GAUSSIAN, False, 2.0, capi.CRIT_CV)
    assert g == pytest.approx(ref, rel=1e-10), (g, ref)
    gm = ctx.var_criterion(x, y, capi.GAUSSIAN, False, 2.0, capi.CRIT_CV, precision=capi.MIXED)
    assert gm == g   # taken again in FP64
