"""Pins of the CPU oracle (oracle/gwr_oracle.cpp) -- runs without a GPU.

- The reference's own known-answer tests on londonhp100
  (test/testGWRBasic.cpp:40-137, test/testGWRMultiscale.cpp:45-257).
- Great-circle distance against closed forms and an independent geodesic
  (Vincenty's inverse formula, 1975) -- the reference has no test for it.
- Kernel closed forms at d = 0 and d = bw.
- The OLS limit of GWR (bandwidth -> infinity) against numpy.linalg.lstsq.
- A brute-force dense-matrix GWR (explicit W_i, C_i, S) on tiny n.
"""
import math
import sys

import numpy as np
import pytest


# --------------------------------------------------------------- KATs

def test_kat_adaptive_gaussian_36(oracle, londonhp100):
    # test/testGWRBasic.cpp:40-68
    r = oracle.gwr_run(londonhp100["coords"], londonhp100["X"], londonhp100["y"], oracle.GAUSSIAN, True, 36)
    d = r["diagnostic"]
    assert abs(d["AIC"] - 2436.60445730413) < 1e-8
    assert abs(d["AICc"] - 2448.27206524754) < 1e-8
    assert abs(d["RSquare"] - 0.708010632044736) < 1e-8
    assert abs(d["RSquareAdjust"] - 0.674975341723766) < 1e-8


def test_kat_fixed_gaussian_5000(oracle, londonhp100):
    # test/testGWRBasic.cpp:70-103
    r = oracle.gwr_run(londonhp100["coords"], londonhp100["X"], londonhp100["y"], oracle.GAUSSIAN, False, 5000)
    d = r["diagnostic"]
    assert abs(d["AIC"] - 2437.649574267587) < 1e-8
    assert abs(d["AICc"] - 2447.676281164379) < 1e-8
    assert abs(d["RSquare"] - 0.701466295457) < 1e-8
    assert abs(d["RSquareAdjust"] - 0.673691594464) < 1e-8


def test_kat_cv_golden_section_67(oracle, londonhp100):
    # test/testGWRBasic.cpp:105-137
    r = oracle.gwr_run(londonhp100["coords"], londonhp100["X"], londonhp100["y"], oracle.GAUSSIAN, True, 0,
                       select_bw=True, crit=oracle.CRIT_CV)
    assert r["bandwidth"] == 67


MGWR_COMMON = dict(centered=[0, 1, 1], bw_threshold=[1e-5] * 3)


def test_kat_mgwr_fixed_bisquare_cv_hat(oracle, londonhp100):
    # test/testGWRMultiscale.cpp:45-88
    r = oracle.mgwr_run(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"], [oracle.BISQUARE] * 3, [0] * 3,
                        [0] * 3, [0] * 3, [oracle.CRIT_CV] * 3, has_hat=True, **MGWR_COMMON)
    bw = r["bandwidths"]
    assert abs(bw[0] - 4623.78) < 0.1 and abs(bw[1] - 12665.70) < 0.1 and abs(bw[2] - 12665.70) < 0.1
    d = r["diagnostic"]
    assert abs(d["AICc"] - 2437.09277417389) < 1e-6
    assert abs(d["RSquare"] - 0.744649364494) < 1e-6
    assert abs(d["RSquareAdjust"] - 0.712344894394) < 1e-6


def test_kat_mgwr_adaptive_aic_initial_hat(oracle, londonhp100):
    # test/testGWRMultiscale.cpp:90-133
    r = oracle.mgwr_run(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"], [oracle.BISQUARE] * 3, [1] * 3,
                        [36] * 3, [1] * 3, [oracle.CRIT_AIC] * 3, has_hat=True, crit_type=1, **MGWR_COMMON)
    assert list(r["bandwidths"]) == [45, 98, 98]
    d = r["diagnostic"]
    assert abs(d["AICc"] - 2437.935218705351) < 1e-6
    assert abs(d["RSquare"] - 0.7486787930045755) < 1e-6
    assert abs(d["RSquareAdjust"] - 0.7118919517893492) < 1e-6


def test_kat_mgwr_adaptive_cv_no_hat(oracle, londonhp100):
    # test/testGWRMultiscale.cpp:135-173
    r = oracle.mgwr_run(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"], [oracle.BISQUARE] * 3, [1] * 3,
                        [0] * 3, [0] * 3, [oracle.CRIT_CV] * 3, has_hat=False, **MGWR_COMMON)
    assert list(r["bandwidths"]) == [35, 98, 98]
    assert abs(r["diagnostic"]["RSquare"] - 0.757377391669) < 1e-6


def test_kat_mgwr_adaptive_cv_cvr_hat(oracle, londonhp100):
    # test/testGWRMultiscale.cpp:175-218
    r = oracle.mgwr_run(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"], [oracle.BISQUARE] * 3, [1] * 3,
                        [36] * 3, [1] * 3, [oracle.CRIT_CV] * 3, has_hat=True, crit_type=0, **MGWR_COMMON)
    assert list(r["bandwidths"]) == [35, 98, 98]
    d = r["diagnostic"]
    assert abs(d["AICc"] - 2438.256543496552) < 1e-6
    assert abs(d["RSquare"] - 0.757377391669) < 1e-6
    assert abs(d["RSquareAdjust"] - 0.715598248225) < 1e-6


def test_kat_mgwr_bounds(oracle, londonhp100):
    # test/testGWRMultiscale.cpp:220-257
    r = oracle.mgwr_run(londonhp100["coords"], londonhp100["Xm"], londonhp100["y"], [oracle.BISQUARE] * 3, [1] * 3,
                        [36] * 3, [0] * 3, [oracle.CRIT_CV] * 3, has_hat=True, crit_type=0, lower=50, upper=100,
                        **MGWR_COMMON)
    assert list(r["bandwidths"]) == [52, 99, 99]


def test_mgwr_per_variable_distance_equator_scaling(oracle):
    """One distance per variable (a SpatialWeight each, GWRMultiscale.h:578).
    On the equator the great-circle distance is a * dlon (radians), so a
    great-circle variable with fixed bandwidth b km weighs like a Euclidean
    (degree) variable with bandwidth b / (a pi / 180): the two runs agree.
    Variable 0 (the initial all-variable search) stays Euclidean in both.  The
    same run with the flag ignored (b km applied to degrees) must differ."""
    rng = np.random.default_rng(11)
    n = 150
    lon = np.sort(rng.uniform(0.0, 20.0, n))
    coords = np.column_stack([lon, np.zeros(n)])
    X = np.column_stack([np.ones(n), rng.normal(size=n), rng.normal(size=n)])
    y = 1.0 + np.sin(lon / 3) * X[:, 1] + 0.5 * X[:, 2] + 0.1 * rng.normal(size=n)
    c = 6378.137 * math.pi / 180
    common = dict(init_type=[2] * 3, approach=[oracle.CRIT_CV] * 3, centered=[0, 1, 1], bw_threshold=[1e-5] * 3,
                  has_hat=True)
    kern = [oracle.BISQUARE] * 3
    g = oracle.mgwr_run(coords, X, y, kern, [0] * 3, [4.0, 6.0 * c, 5.0], geographic=[0, 1, 0], **common)
    e = oracle.mgwr_run(coords, X, y, kern, [0] * 3, [4.0, 6.0, 5.0], geographic=[0, 0, 0], **common)
    assert g["iterations"] == e["iterations"]
    assert np.allclose(g["betas"], e["betas"], rtol=1e-10, atol=1e-12)
    for key in ["RSS", "AICc", "ENP", "RSquare"]:
        assert g["diagnostic"][key] == pytest.approx(e["diagnostic"][key], rel=1e-10), key
    wrong = oracle.mgwr_run(coords, X, y, kern, [0] * 3, [4.0, 6.0 * c, 5.0], geographic=[0, 0, 0], **common)
    assert np.max(np.abs(wrong["betas"][:, 1] - e["betas"][:, 1])) > 1e-2


# --------------------------------------------------------------- great circle

def vincenty_km(lon1, lat1, lon2, lat2):
    """Vincenty (1975) inverse geodesic on WGS-84, km -- an independent method."""
    a, f = 6378.137, 1 / 298.257223563
    b = a * (1 - f)
    L = math.radians(lon2 - lon1)
    U1 = math.atan((1 - f) * math.tan(math.radians(lat1)))
    U2 = math.atan((1 - f) * math.tan(math.radians(lat2)))
    sU1, cU1, sU2, cU2 = math.sin(U1), math.cos(U1), math.sin(U2), math.cos(U2)
    lam = L
    for _ in range(200):
        sl, cl = math.sin(lam), math.cos(lam)
        ss = math.sqrt((cU2 * sl) ** 2 + (cU1 * sU2 - sU1 * cU2 * cl) ** 2)
        if ss == 0:
            return 0.0
        cs = sU1 * sU2 + cU1 * cU2 * cl
        sig = math.atan2(ss, cs)
        sa = cU1 * cU2 * sl / ss
        c2a = 1 - sa * sa
        c2sm = cs - 2 * sU1 * sU2 / c2a if c2a != 0 else 0.0
        C = f / 16 * c2a * (4 + f * (4 - 3 * c2a))
        lp = lam
        lam = L + (1 - C) * f * sa * (sig + C * ss * (c2sm + C * cs * (-1 + 2 * c2sm * c2sm)))
        if abs(lam - lp) < 1e-13:
            break
    u2 = c2a * (a * a - b * b) / (b * b)
    A = 1 + u2 / 16384 * (4096 + u2 * (-768 + u2 * (320 - 175 * u2)))
    B = u2 / 1024 * (256 + u2 * (-128 + u2 * (74 - 47 * u2)))
    ds = B * ss * (c2sm + B / 4 * (cs * (-1 + 2 * c2sm ** 2) - B / 6 * c2sm * (-3 + 4 * ss ** 2) * (-3 + 4 * c2sm ** 2)))
    return b * A * (sig - ds)


def test_gcdist_equator_closed_form(oracle):
    a = 6378.137
    for dlon in [0.001, 0.5, 1.0, 17.0, 120.0]:
        d = oracle.spgcdist(10.0, 10.0 + dlon, 0.0, 0.0)
        assert abs(d - a * math.radians(dlon)) <= 1e-9 * d
    assert abs(oracle.spgcdist(0.0, 1.0, 0.0, 0.0) - 111.3195) < 1e-4


def test_gcdist_coincident_and_antimeridian(oracle):
    assert oracle.spgcdist(12.5, 12.5, 48.0, 48.0) == 0.0
    assert oracle.spgcdist(180.0, -180.0, 30.0, 30.0) == 0.0
    assert oracle.spgcdist(12.5, 12.5 + 5e-9, 48.0, 48.0 + 5e-9) == 0.0


def test_gcdist_vs_vincenty_and_symmetry(oracle):
    rng = np.random.default_rng(7)
    for _ in range(300):
        lon1, lon2 = rng.uniform(-10, 30, 2)
        lat1, lat2 = rng.uniform(35, 60, 2)
        d = oracle.spgcdist(lon1, lon2, lat1, lat2)
        v = vincenty_km(lon1, lat1, lon2, lat2)
        # Andoyer-Lambert is first order in the flattening: ~f^2 relative error
        assert abs(d - v) <= 2e-5 * v + 1e-6
        assert d == pytest.approx(oracle.spgcdist(lon2, lon1, lat2, lat1), rel=1e-14)


# --------------------------------------------------------------- kernels

def test_kernel_closed_forms(oracle):
    d = np.array([0.0, 2.0])
    bw = 2.0
    expect = {oracle.GAUSSIAN: [1.0, math.exp(-0.5)], oracle.EXPONENTIAL: [1.0, math.exp(-1.0)],
              oracle.BISQUARE: [1.0, 0.0], oracle.TRICUBE: [1.0, 0.0], oracle.BOXCAR: [1.0, 0.0]}
    for kern, ex in expect.items():
        w = oracle.weight_vec(d, kern, False, bw)
        assert w[0] == ex[0] and abs(w[1] - ex[1]) < 1e-15
    dd = np.linspace(0, 3, 301)
    for kern in range(5):
        w = oracle.weight_vec(dd, kern, False, 2.5)
        assert np.all(np.diff(w) <= 1e-15)


def test_kernel_midpoint_closed_forms(oracle):
    """d = bw/2 and d = bw/4 separate every kernel's exponent and power: a
    tricube written (t*t) instead of (t*t)*t, an exponential written
    exp(-(d/b)^2) or a bisquare without its square each fail here."""
    bw = 2.0
    d = np.array([bw / 2, bw / 4])
    expect = {oracle.GAUSSIAN: [math.exp(-1 / 8), math.exp(-1 / 32)],
              oracle.EXPONENTIAL: [math.exp(-1 / 2), math.exp(-1 / 4)],
              oracle.BISQUARE: [(1 - 1 / 4) ** 2, (1 - 1 / 16) ** 2],
              oracle.TRICUBE: [(1 - 1 / 8) ** 3, (1 - 1 / 64) ** 3],
              oracle.BOXCAR: [1.0, 1.0]}
    for kern, ex in expect.items():
        w = oracle.weight_vec(d, kern, False, bw)
        assert np.allclose(w, ex, rtol=2e-16, atol=0), (kern, w, ex)
    # adaptive: the threshold replaces bw (BandwidthWeight.cpp:35-61); with
    # the distances 0, 1, 2, ..., bw = 3.5 resolves to 2.5
    dist = np.arange(10, dtype=float)
    w = oracle.weight_vec(dist, oracle.TRICUBE, True, 3.5)
    assert np.allclose(w, np.where(dist < 2.5, (1 - (dist / 2.5) ** 3) ** 3, 0.0), rtol=1e-15, atol=0)


def test_adaptive_threshold_vs_sort(oracle):
    rng = np.random.default_rng(3)
    coords = np.round(rng.uniform(0, 20, (300, 2)), 1)   # rounding creates ties
    for focus in [0, 17, 299]:
        d = oracle.distance_vec(coords, focus)
        s = np.sort(d)
        for bw in [1.0, 2.0, 20.0, 20.4, 150.75, 299.0]:
            b0 = math.floor(bw)
            expect = s[b0 - 1] + (s[b0] - s[b0 - 1]) * (bw - b0)
            assert oracle.adaptive_threshold(coords, bw, focus) == expect
        assert oracle.adaptive_threshold(coords, 450.0, focus) == (450.0 / 300) * d.max()


# --------------------------------------------------------------- OLS limit and brute force

def test_ols_limit(oracle):
    rng = np.random.default_rng(11)
    n = 60
    coords = rng.uniform(0, 10, (n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([1.0, -2.0, 0.5]) + rng.normal(size=n)
    beta, *_ = np.linalg.lstsq(X, y, rcond=None)
    H = X @ np.linalg.solve(X.T @ X, X.T)
    for kern, bw in [(oracle.BOXCAR, 1e6), (oracle.GAUSSIAN, 1e9)]:
        f = oracle.fit_core(coords, X, y, kern, False, bw)
        assert np.allclose(f["betas"], beta[None, :], rtol=1e-9, atol=1e-12)
        assert f["shat"][0] == pytest.approx(3.0, rel=1e-9)
        assert f["shat"][1] == pytest.approx(3.0, rel=1e-9)
        assert np.allclose(f["qDiag"], 1 - np.diag(H), rtol=1e-8, atol=1e-12)


def dense_kernel(kern, d, b):
    """The five kernels from their textbook definitions (numpy, not the
    oracle's code): Gaussian exp(-d^2 / 2b^2), exponential exp(-d/b),
    bisquare (1 - (d/b)^2)^2, tricube (1 - (d/b)^3)^3, boxcar 1, the last
    three for d < b only."""
    q = d / b
    if kern == 0:
        return np.exp(-0.5 * q * q)
    if kern == 1:
        return np.exp(-q)
    if kern == 2:
        return np.where(d < b, (1 - q * q) ** 2, 0.0)
    if kern == 3:
        return np.where(d < b, (1 - q ** 3) ** 3, 0.0)
    return np.where(d < b, 1.0, 0.0)


def dense_gwr(coords, X, y, kern, adaptive, bw, cv=False):
    """Independent dense restatement: explicit W_i, C_i = (X^T W X)^-1 X^T W_i and S
    (cv: the focus point's own weight zeroed, as the CV criterion does)."""
    n, k = X.shape
    D = np.sqrt(((coords[:, None, :] - coords[None, :, :]) ** 2).sum(-1))
    betas = np.zeros((n, k))
    se = np.zeros((n, k))
    S = np.zeros((n, n))
    for i in range(n):
        d = D[i]
        b = bw
        if adaptive:
            s = np.sort(d)
            b0 = int(math.floor(bw))
            b = s[b0 - 1] + (s[b0] - s[b0 - 1]) * (bw - b0)
        w = dense_kernel(kern, d, b)
        if cv:
            w[i] = 0.0
        C = np.linalg.solve(X.T @ (w[:, None] * X), (w[:, None] * X).T)
        betas[i] = C @ y
        se[i] = (C * C).sum(1)
        S[i] = X[i] @ C
    Q = (np.eye(n) - S)
    return betas, se, np.array([np.trace(S), np.trace(S @ S.T)]), (Q * Q).sum(0), S


def test_brute_force_dense(oracle):
    rng = np.random.default_rng(5)
    n = 40
    coords = rng.uniform(0, 10, (n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([0.3, 1.0, -1.0]) + rng.normal(size=n)
    for kern, adaptive, bw in [(oracle.GAUSSIAN, False, 3.0), (oracle.BISQUARE, True, 25.5),
                               (oracle.EXPONENTIAL, False, 2.0), (oracle.TRICUBE, False, 6.0),
                               (oracle.TRICUBE, True, 30.0), (oracle.BOXCAR, False, 7.0),
                               (oracle.EXPONENTIAL, True, 12.0)]:
        b, se, shat, q, S = dense_gwr(coords, X, y, kern, adaptive, bw)
        f = oracle.fit_core(coords, X, y, kern, adaptive, bw, want_S=True)
        assert np.allclose(f["betas"], b, rtol=1e-10, atol=1e-12)
        assert np.allclose(f["betasSE_raw"], se, rtol=1e-10, atol=1e-14)
        assert np.allclose(f["shat"], shat, rtol=1e-10)
        assert np.allclose(f["qDiag"], q, rtol=1e-10)
        assert np.allclose(f["S"], S, rtol=1e-10, atol=1e-13)


def test_brute_force_criteria_all_kernels(oracle):
    """CV (self weight zeroed, GWRBasic.cpp:233-268) and AICc (:270-307) of the
    dense restatement, for all five kernels, fixed and adaptive."""
    rng = np.random.default_rng(9)
    n = 45
    coords = rng.uniform(0, 10, (n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([0.3, 1.0, -1.0]) + rng.normal(size=n)
    cases = [(0, False, 2.5), (1, False, 1.5), (2, False, 6.0), (3, False, 6.5), (4, False, 8.0),
             (0, True, 15.0), (1, True, 12.0), (2, True, 25.5), (3, True, 28.0), (4, True, 30.0)]
    for kern, adaptive, bw in cases:
        b, _, shat, _, _ = dense_gwr(coords, X, y, kern, adaptive, bw)
        rss = float(np.sum((y - np.sum(X * b, 1)) ** 2))
        aicc = n * math.log(rss / n) + n * math.log(2 * math.pi) + n * (n + shat[0]) / (n - 2 - shat[0])
        bcv, _, _, _, _ = dense_gwr(coords, X, y, kern, adaptive, bw, cv=True)
        cv = float(np.sum((y - np.sum(X * bcv, 1)) ** 2))
        assert oracle.criterion(coords, X, y, kern, adaptive, bw, oracle.CRIT_AIC) == pytest.approx(aicc, rel=1e-11)
        assert oracle.criterion(coords, X, y, kern, adaptive, bw, oracle.CRIT_CV) == pytest.approx(cv, rel=1e-11)


# --------------------------------------------------------------- predict (GWRBasic.cpp:167-191)

def dense_predict(coords, X, y, loc, kern, adaptive, bw):
    """Independent dense restatement of predictSerial at locations that are not data points."""
    m, k = loc.shape[0], X.shape[1]
    D = np.sqrt(((loc[:, None, :] - coords[None, :, :]) ** 2).sum(-1))
    betas = np.zeros((m, k))
    for i in range(m):
        d = D[i]
        b = bw
        if adaptive:
            s = np.sort(d)
            b0 = int(math.floor(bw))
            b = s[b0 - 1] + (s[b0] - s[b0 - 1]) * (bw - b0)
        w = dense_kernel(kern, d, b)
        betas[i] = np.linalg.solve(X.T @ (w[:, None] * X), X.T @ (w * y))
    return betas


def test_predict_at_data_points_is_fit(oracle, londonhp100):
    # the KAT configuration (adaptive 36, Gaussian): at the data coordinates the
    # prediction weights are the fit's, so the betas are the fit's betas
    c, X, y = londonhp100["coords"], londonhp100["X"], londonhp100["y"]
    f = oracle.fit_core(c, X, y, oracle.GAUSSIAN, True, 36)
    p = oracle.predict(c, X, y, c, oracle.GAUSSIAN, True, 36)
    assert np.allclose(p, f["betas"], rtol=1e-10, atol=1e-12)
    f = oracle.fit_core(c, X, y, oracle.BISQUARE, False, 5000)
    p = oracle.predict(c, X, y, c, oracle.BISQUARE, False, 5000)
    assert np.allclose(p, f["betas"], rtol=1e-10, atol=1e-12)


def test_predict_at_data_points_is_fit_geographic(oracle):
    rng = np.random.default_rng(3)
    n = 80
    coords = np.column_stack([rng.uniform(100, 120, n), rng.uniform(20, 40, n)])
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([1.0, 0.5, -0.5]) + rng.normal(size=n)
    f = oracle.fit_core(coords, X, y, oracle.GAUSSIAN, True, 30, geographic=True)
    p = oracle.predict(coords, X, y, coords, oracle.GAUSSIAN, True, 30, geographic=True)
    assert np.allclose(p, f["betas"], rtol=1e-10, atol=1e-12)


def test_predict_ols_limit(oracle):
    rng = np.random.default_rng(12)
    n = 60
    coords = rng.uniform(0, 10, (n, 2))
    loc = rng.uniform(-5, 15, (25, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([1.0, -2.0, 0.5]) + rng.normal(size=n)
    beta, *_ = np.linalg.lstsq(X, y, rcond=None)
    for kern, adaptive, bw in [(oracle.BOXCAR, False, 1e6), (oracle.GAUSSIAN, False, 1e9), (oracle.BOXCAR, True, 2 * n)]:
        p = oracle.predict(coords, X, y, loc, kern, adaptive, bw)
        assert np.allclose(p, beta[None, :], rtol=1e-9, atol=1e-12)


def test_predict_brute_force_dense(oracle):
    rng = np.random.default_rng(6)
    n = 50
    coords = rng.uniform(0, 10, (n, 2))
    loc = rng.uniform(0, 10, (30, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=(n, 2))])
    y = X @ np.array([0.3, 1.0, -1.0]) + rng.normal(size=n)
    for kern, adaptive, bw in [(oracle.GAUSSIAN, False, 3.0), (oracle.BISQUARE, True, 25.5),
                               (oracle.GAUSSIAN, True, 7), (oracle.EXPONENTIAL, False, 2.0),
                               (oracle.TRICUBE, True, 20.0), (oracle.BOXCAR, True, 24.0)]:
        p = oracle.predict(coords, X, y, loc, kern, adaptive, bw)
        assert np.allclose(p, dense_predict(coords, X, y, loc, kern, adaptive, bw), rtol=1e-10, atol=1e-12)


def test_predict_not_spd(oracle):
    rng = np.random.default_rng(7)
    n = 40
    coords = rng.uniform(0, 10, (n, 2))
    X = np.column_stack([np.ones(n), rng.normal(size=n)])
    y = rng.normal(size=n)
    loc = np.array([[5.0, 5.0], [500.0, 500.0]])
    p = oracle.predict(coords, X, y, loc, oracle.BISQUARE, False, 20.0)
    assert p is None


# --------------------------------------------------------------- variable selection (testGWRBasic.cpp:139-240)

VARSEL_KAT = [([2], 2551.61359020599), ([3], 2551.30032201349), ([1], 2468.93236280013),
              ([1, 3], 2452.86447942033), ([1, 2], 2450.59642666509), ([1, 2, 3], 2452.80388934625)]


def test_kat_indep_vars_select(oracle, londonhp100):
    sel, models = oracle.indep_vars_select(londonhp100["X"], londonhp100["y"], threshold=3.0)
    assert sel == [1, 3]
    assert len(models) == len(VARSEL_KAT)
    for (v, c), (kv, kc) in zip(models, VARSEL_KAT):
        assert v == kv
        assert abs(c - kc) <= 1e-8


def test_indep_vars_criterion_is_ols_aicc(oracle, londonhp100):
    # unit weights: the criterion is the AICc of the global OLS with tr(S) = p
    X, y = londonhp100["X"], londonhp100["y"]
    n = len(y)
    for cols in ([0, 1], [0, 2, 3], [0, 1, 2, 3]):
        Xs = X[:, cols]
        beta, *_ = np.linalg.lstsq(Xs, y, rcond=None)
        rss = float(((y - Xs @ beta) ** 2).sum())
        p = len(cols)
        aicc = n * math.log(rss / n) + n * math.log(2 * math.pi) + n * (n + p) / (n - 2 - p)
        assert oracle.indep_vars_criterion(X, y, cols[1:]) == pytest.approx(aicc, rel=1e-12)
    # a zero column makes the design singular: DBL_MAX
    Xd = np.column_stack([X, np.zeros(n)])
    assert oracle.indep_vars_criterion(Xd, y, [1, 4]) == sys.float_info.max


def test_multiscale_fixture_inputs_match_generator():
    """tests/golden/cfg3_mgwr_n3000.npz (the oracle's GWRMultiscale result at
    n = 3000, written by tests/golden/make_cfg3_fixture.py) belongs to the
    inputs the seeded generator produces today."""
    import os
    import sys
    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    sys.path.insert(0, golden)
    import make_cfg3_fixture as fx
    ref = np.load(os.path.join(golden, "cfg3_mgwr_n3000.npz"))
    d = fx.inputs()
    assert np.array_equal(fx.fingerprint(d), ref["fingerprint"])
    assert ref["betas"].shape == d["X"].shape and ref["bandwidths"].shape == (d["X"].shape[1],)
