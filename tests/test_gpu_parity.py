"""GPU parity of the FP64 HIP path against the CPU oracle (calls go through the
C ABI of include/gwmodel_hip.h).

Tolerances (FP64 mode): bandwidths, adaptive thresholds, Euclidean distances
and golden-section iterates bit-exact; criteria within 1e-11 relative; beta,
SE, qDiag, S within 1e-9 relative; reference KAT diagnostics at the
reference's own tolerances (1e-8 GWRBasic, 1e-6 GWRMultiscale).
"""
import sys

import numpy as np
import pytest

from libgwmodel_amd import capi, synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_err(a, b, floor=1e-300):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def close(a, b, rtol, atol=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(b) + atol))


@pytest.fixture(scope="module")
def ctx():
    say("create context")
    c = capi.Context(0)
    yield c
    c.close()


# ------------------------------------------------------------------ distances

def test_euclidean_distances_bitexact(ctx, oracle, londonhp100):
    say("case: euclidean distances londonhp100")
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    for f in [0, 1, 57, 99]:
        g = ctx.distance_row(f)
        o = oracle.distance_vec(londonhp100["coords"], f)
        assert np.array_equal(g, o)


def test_great_circle_distances(ctx, oracle):
    say("case: great-circle distances")
    d = synth.sphere_box(n=500)
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    for f in [0, 3, 250, 499]:
        g = ctx.distance_row(f)
        o = oracle.distance_vec(d["coords"], f, geographic=True)
        assert close(g, o, 1e-12, 1e-9)
    m = ctx.max_distance()
    assert m == pytest.approx(oracle.max_distance(d["coords"], True), rel=1e-12)


def test_max_distance_euclidean_bitexact(ctx, oracle, londonhp100):
    say("case: max distance")
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    assert ctx.max_distance() == oracle.max_distance(londonhp100["coords"])


def test_max_distance_sliced_and_per_problem(ctx, oracle):
    """n = 5000: the pairs are split over 20 x 20 workgroups and combined by an
    atomic maximum; the value is cached per problem and recomputed for new
    coordinates."""
    rng = np.random.default_rng(21)
    for scale in (100.0, 37.0):
        say("case: max distance n=5000 scale=%g" % scale)
        coords = rng.uniform(0, scale, (5000, 2))
        X = np.column_stack([np.ones(5000), rng.normal(size=5000)])
        ctx.set_data(coords, X, rng.normal(size=5000))
        ref = oracle.max_distance(coords)
        assert ctx.max_distance() == ref
        assert ctx.max_distance() == ref


# ------------------------------------------------------------------ adaptive kNN

@pytest.mark.parametrize("n,seed", [(100, 0), (1000, 1), (3000, 2)])
def test_adaptive_thresholds_bitexact(ctx, oracle, n, seed):
    say("case: adaptive thresholds n=%d" % n)
    rng = np.random.default_rng(seed)
    coords = np.round(rng.uniform(0, 30, (n, 2)), 1)   # many exact ties
    coords[5] = coords[6]                               # coincident points
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = rng.normal(size=n)
    ctx.set_data(coords, X, y)
    for bw in [1.0, 2.0, 20.0, 36.5, n - 1.0, n - 0.25, float(n), 1.5 * n]:
        g = ctx.adaptive_bandwidths(bw)
        o = np.array([oracle.adaptive_threshold(coords, bw, f) for f in range(0, n, max(1, n // 97))])
        assert np.array_equal(g[::max(1, n // 97)], o), "bw=%r" % bw


# ------------------------------------------------------------------ criteria and fit

CASES = [
    # kernel, adaptive, bandwidth
    (capi.GAUSSIAN, False, 15.0),
    (capi.GAUSSIAN, True, 60.0),
    (capi.BISQUARE, True, 45.0),
    (capi.BISQUARE, False, 30.0),
    (capi.TRICUBE, True, 80.5),
    (capi.EXPONENTIAL, False, 10.0),
    (capi.BOXCAR, False, 35.0),
    (capi.BOXCAR, True, 40.0),
]


@pytest.mark.parametrize("kernel,adaptive,bw", CASES)
def test_criteria_match_oracle(ctx, oracle, kernel, adaptive, bw):
    say("case: criteria kernel=%d adaptive=%d bw=%g" % (kernel, adaptive, bw))
    d = synth.uniform_square(n=400)
    ctx.set_data(d["coords"], d["X"], d["y"])
    for crit in [capi.CRIT_CV, capi.CRIT_AIC]:
        g = ctx.criterion(kernel, adaptive, bw, crit)
        o = oracle.criterion(d["coords"], d["X"], d["y"], kernel, adaptive, bw, crit)
        assert g == pytest.approx(o, rel=1e-11), (crit, g, o)


@pytest.mark.parametrize("kernel,adaptive,bw", CASES)
def test_fit_matches_oracle(ctx, oracle, kernel, adaptive, bw):
    say("case: fit kernel=%d adaptive=%d bw=%g" % (kernel, adaptive, bw))
    d = synth.uniform_square(n=300)
    ctx.set_data(d["coords"], d["X"], d["y"])
    g = ctx.fit(kernel, adaptive, bw, want_S=True)
    o = oracle.fit_core(d["coords"], d["X"], d["y"], kernel, adaptive, bw, want_S=True)
    assert close(g["betas"], o["betas"], 1e-9, 1e-12)
    assert close(g["betasSE_raw"], o["betasSE_raw"], 1e-9, 1e-15)
    assert close(g["shat"], o["shat"], 1e-10)
    assert close(g["qDiag"], o["qDiag"], 1e-9, 1e-12)
    assert close(g["S"], o["S"], 1e-9, 1e-12)


@pytest.mark.parametrize("kernel,adaptive,bw", [(capi.BOXCAR, False, 5.0), (capi.BOXCAR, False, 10.0),
                                               (capi.BOXCAR, True, 21.0), (capi.BISQUARE, False, 5.0),
                                               (capi.BISQUARE, True, 13.0)])
def test_boundary_ties_match_oracle(ctx, oracle, kernel, adaptive, bw):
    """Points on an integer grid: many pairs lie exactly at the bandwidth
    (3-4-5 and 6-8-10 triangles for fixed 5 and 10; the floor(bw)-th
    neighbour and its ties for adaptive integer bandwidths).  The reference
    gives them zero weight (d < bw, BandwidthWeight.h:68-121); the FP64 sweeps
    decide the test on the squared distance, so every such pair must fall
    out here too: a boxcar weight of 1 instead of 0 moves the criterion far
    beyond the tolerance."""
    say("case: boundary ties kernel=%d adaptive=%d bw=%g" % (kernel, adaptive, bw))
    rng = np.random.default_rng(31)
    gx, gy = np.meshgrid(np.arange(24.0), np.arange(20.0))
    coords = np.column_stack([gx.ravel(), gy.ravel()])
    n = coords.shape[0]
    X = np.column_stack([np.ones(n), rng.normal(size=n), rng.normal(size=n)])
    y = X @ np.array([1.0, 0.5, -0.25]) + np.sin(coords[:, 0] / 5.0) + rng.normal(0, 0.2, n)
    ctx.set_data(coords, X, y)
    for crit in [capi.CRIT_CV, capi.CRIT_AIC]:
        g = ctx.criterion(kernel, adaptive, bw, crit)
        o = oracle.criterion(coords, X, y, kernel, adaptive, bw, crit)
        assert g == pytest.approx(o, rel=1e-11), (crit, g, o)
    g = ctx.fit(kernel, adaptive, bw)
    o = oracle.fit_core(coords, X, y, kernel, adaptive, bw)
    assert close(g["betas"], o["betas"], 1e-9, 1e-12)
    assert close(g["shat"], o["shat"], 1e-10)


@pytest.mark.parametrize("kernel,bw", [(capi.GAUSSIAN, 4.0), (capi.EXPONENTIAL, 1.0)])
def test_exponent_cut_matches_oracle(ctx, oracle, kernel, bw):
    """The FP64 weights take Gaussian / exponential weights below e^-40 as zero
    (beyond sqrt(80) = 8.9 / 40 bandwidths), where the oracle keeps them down to exp
    underflow.  At these bandwidths a large share of the pairs lies in that
    band, and the results still agree at the FP64 tolerances."""
    say("case: exponent cut kernel=%d bw=%g" % (kernel, bw))
    d = synth.uniform_square(n=1500)
    c = d["coords"]
    cut = (np.sqrt(80.0) if kernel == capi.GAUSSIAN else 40.0) * bw
    dd = np.hypot(c[:300, None, 0] - c[None, :, 0], c[:300, None, 1] - c[None, :, 1])
    assert np.mean(dd > cut) > 0.1   # the band the cut drops is exercised
    ctx.set_data(c, d["X"], d["y"])
    for crit in [capi.CRIT_CV, capi.CRIT_AIC]:
        g = ctx.criterion(kernel, False, bw, crit)
        o = oracle.criterion(c, d["X"], d["y"], kernel, False, bw, crit)
        assert g == pytest.approx(o, rel=1e-11), (crit, g, o)
    g = ctx.fit(kernel, False, bw)
    o = oracle.fit_core(c, d["X"], d["y"], kernel, False, bw)
    assert close(g["betas"], o["betas"], 1e-9, 1e-12)
    assert close(g["shat"], o["shat"], 1e-10)
    assert close(g["qDiag"], o["qDiag"], 1e-9, 1e-12)


@pytest.mark.parametrize("kernel,adaptive,bw", [(capi.GAUSSIAN, False, 2500.0), (capi.GAUSSIAN, False, 1500.0),
                                                 (capi.BISQUARE, True, 60.0), (capi.EXPONENTIAL, False, 800.0)])
def test_great_circle_world_scale(ctx, oracle, kernel, adaptive, bw):
    """Great-circle sweeps over pairs of every separation: the FP64 sweep weighs
    pairs up to about 41 degrees of arc through a polynomial form of the
    distance and the others (antipodal-range pairs, coincident points, points
    within 1e-8 degrees) through the general one.  Criteria and fits against
    the oracle at the FP64 tolerances.  The bandwidths keep the local designs
    well conditioned at this density (points about 900 km apart): at 300 km
    both sides return designs with condition numbers near 1e8, whose betas
    differ at 1e-8 from summation order alone."""
    say("case: world-scale great circle kernel=%d adaptive=%d bw=%g" % (kernel, adaptive, bw))
    rng = np.random.default_rng(77)
    n = 600
    lon = rng.uniform(-180.0, 180.0, n)
    lat = np.degrees(np.arcsin(rng.uniform(-0.95, 0.95, n)))
    lon[10], lat[10] = lon[11], lat[11]                       # coincident
    lon[20], lat[20] = lon[21] + 5e-9, lat[21] - 5e-9         # within the reference's 1e-8 degree test
    lon[30], lat[30] = 180.0, 12.5                            # the antimeridian, both ways
    lon[31], lat[31] = -180.0, 12.5
    coords = np.column_stack([lon, lat])
    X = np.column_stack([np.ones(n), rng.normal(size=n), rng.normal(size=n)])
    y = 1.0 + np.cos(np.radians(lat)) * X[:, 1] + 0.5 * X[:, 2] + 0.2 * rng.normal(size=n)
    ctx.set_data(coords, X, y, capi.GREAT_CIRCLE)
    for crit in [capi.CRIT_CV, capi.CRIT_AIC]:
        g = ctx.criterion(kernel, adaptive, bw, crit)
        o = oracle.criterion(coords, X, y, kernel, adaptive, bw, crit, geographic=True)
        assert g == pytest.approx(o, rel=1e-11), (crit, g, o)
    g = ctx.fit(kernel, adaptive, bw)
    o = oracle.fit_core(coords, X, y, kernel, adaptive, bw, geographic=True)
    assert close(g["betas"], o["betas"], 1e-9, 1e-12)
    assert close(g["betasSE_raw"], o["betasSE_raw"], 1e-9, 1e-15)
    assert close(g["shat"], o["shat"], 1e-10)
    assert close(g["qDiag"], o["qDiag"], 1e-9, 1e-12)


@pytest.mark.parametrize("k", [1, 2, 5, 8, 12, 16])
def test_fit_all_k(ctx, oracle, k):
    say("case: fit k=%d" % k)
    rng = np.random.default_rng(100 + k)
    n = 257
    coords = rng.uniform(0, 50, (n, 2))
    X = np.column_stack([np.ones(n)] + [rng.normal(size=n) for _ in range(k - 1)]) if k > 1 else np.ones((n, 1))
    y = X @ rng.normal(size=k) + rng.normal(size=n)
    ctx.set_data(coords, X, y)
    g = ctx.fit(capi.GAUSSIAN, True, 120.0)
    o = oracle.fit_core(coords, X, y, capi.GAUSSIAN, True, 120.0)
    assert close(g["betas"], o["betas"], 1e-9, 1e-11)
    assert close(g["shat"], o["shat"], 1e-10)
    assert close(g["qDiag"], o["qDiag"], 1e-9, 1e-12)
    cv = ctx.criterion(capi.BISQUARE, True, 90.0, capi.CRIT_CV)
    assert cv == pytest.approx(oracle.criterion(coords, X, y, capi.BISQUARE, True, 90.0, capi.CRIT_CV), rel=1e-11)


# ------------------------------------------------------------------ reference KATs through the C ABI

def test_kat_adaptive_gaussian_36(ctx, londonhp100):
    say("case: KAT adaptive gaussian 36")
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    d = ctx.gwr_run(capi.GAUSSIAN, True, 36)["diagnostic"]
    assert abs(d["AIC"] - 2436.60445730413) < 1e-8
    assert abs(d["AICc"] - 2448.27206524754) < 1e-8
    assert abs(d["RSquare"] - 0.708010632044736) < 1e-8
    assert abs(d["RSquareAdjust"] - 0.674975341723766) < 1e-8


def test_kat_fixed_gaussian_5000(ctx, londonhp100):
    say("case: KAT fixed gaussian 5000")
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    d = ctx.gwr_run(capi.GAUSSIAN, False, 5000)["diagnostic"]
    assert abs(d["AIC"] - 2437.649574267587) < 1e-8
    assert abs(d["AICc"] - 2447.676281164379) < 1e-8
    assert abs(d["RSquare"] - 0.701466295457) < 1e-8
    assert abs(d["RSquareAdjust"] - 0.673691594464) < 1e-8


def test_kat_cv_golden_section(ctx, oracle, londonhp100):
    say("case: KAT CV golden section 67")
    ctx.set_data(londonhp100["coords"], londonhp100["X"], londonhp100["y"])
    r = ctx.gwr_run(capi.GAUSSIAN, True, 0, select_bw=True, crit=capi.CRIT_CV)
    assert r["bandwidth"] == 67
    o = oracle.gwr_run(londonhp100["coords"], londonhp100["X"], londonhp100["y"], capi.GAUSSIAN, True, 0,
                       select_bw=True, crit=oracle.CRIT_CV)
    assert [x for x, _ in r["trace"]] == [x for x, _ in o["trace"]]
    assert close([f for _, f in r["trace"]], [f for _, f in o["trace"]], 1e-11)


def test_kat_mgwr_no_hat(ctx, londonhp100):
    say("case: KAT MGWR adaptive CV no hat")
    ctx.set_data(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"])
    r = ctx.mgwr_run([capi.BISQUARE] * 3, [1] * 3, [0] * 3, [0] * 3, [capi.CRIT_CV] * 3, [0, 1, 1], [1e-5] * 3)
    assert list(r["bandwidths"]) == [35, 98, 98]
    assert abs(r["diagnostic"]["RSquare"] - 0.757377391669) < 1e-6


# ------------------------------------------------------------------ searches vs oracle

def test_config1_cv_search_matches_oracle(ctx, oracle):
    say("case: uniform_square n=600 bisquare adaptive CV search")
    d = synth.uniform_square(n=600)
    ctx.set_data(d["coords"], d["X"], d["y"])
    g = ctx.gwr_run(capi.BISQUARE, True, 0, select_bw=True, crit=capi.CRIT_CV, lower=20, upper=600)
    o = oracle.gwr_run(d["coords"], d["X"], d["y"], oracle.BISQUARE, True, 0, select_bw=True, crit=oracle.CRIT_CV,
                       lower=20, upper=600)
    assert g["bandwidth"] == o["bandwidth"]
    assert [x for x, _ in g["trace"]] == [x for x, _ in o["trace"]]
    for key in ["AICc", "RSquare", "RSquareAdjust"]:
        assert g["diagnostic"][key] == pytest.approx(o["diagnostic"][key], rel=1e-10)
    assert close(g["betas"], o["betas"], 1e-9, 1e-12)
    assert close(g["betasSE"], o["betasSE"], 1e-9, 1e-12)


def test_config2_aicc_search_matches_oracle(ctx, oracle):
    say("case: sphere_box n=800 great-circle gaussian AICc search")
    d = synth.sphere_box(n=800)
    lo, hi = synth.SPHERE_BOX_BW_KM
    ctx.set_data(d["coords"], d["X"], d["y"], capi.GREAT_CIRCLE)
    g = ctx.gwr_run(capi.GAUSSIAN, False, 0, select_bw=True, crit=capi.CRIT_AIC, lower=lo, upper=hi)
    o = oracle.gwr_run(d["coords"], d["X"], d["y"], oracle.GAUSSIAN, False, 0, select_bw=True, crit=oracle.CRIT_AIC,
                       lower=lo, upper=hi, geographic=True)
    assert g["bandwidth"] == o["bandwidth"]
    assert [x for x, _ in g["trace"]] == [x for x, _ in o["trace"]]
    assert close(g["betas"], o["betas"], 1e-9, 1e-11)
    assert g["diagnostic"]["AICc"] == pytest.approx(o["diagnostic"]["AICc"], rel=1e-11)


# ------------------------------------------------------------------ failures

def test_not_spd_and_cancel(ctx):
    say("case: not SPD / cancel")
    rng = np.random.default_rng(9)
    n = 200
    coords = rng.uniform(0, 10, (n, 2))
    x1 = rng.normal(size=n)
    X = np.column_stack([np.ones(n), x1, x1])   # collinear -> singular local designs
    y = rng.normal(size=n)
    ctx.set_data(coords, X, y)
    assert ctx.criterion(capi.GAUSSIAN, False, 3.0, capi.CRIT_CV) == capi.DBL_MAX
    with pytest.raises(capi.GwmError) as e:
        ctx.fit(capi.GAUSSIAN, False, 3.0)
    assert e.value.status == capi.GWM_ERR_NOT_SPD
    X2 = np.column_stack([np.ones(n), x1])
    ctx.set_data(coords, X2, y)
    ctx.set_stop_callback(lambda: True)
    with pytest.raises(capi.GwmError) as e:
        ctx.select_bandwidth(capi.GAUSSIAN, True, capi.CRIT_CV)
    assert e.value.status == capi.GWM_ERR_CANCELLED
    ctx.set_stop_callback(None)
    bw, _ = ctx.select_bandwidth(capi.GAUSSIAN, True, capi.CRIT_CV)
    assert 20 <= bw <= n
