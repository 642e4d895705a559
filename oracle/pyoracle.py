"""ctypes wrapper of the CPU FP64 oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / CPU baseline.  See gwr_oracle.cpp for
the reference file:line each function restates.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libgwr_oracle.so")
_lib = None

GAUSSIAN, EXPONENTIAL, BISQUARE, TRICUBE, BOXCAR = range(5)
CRIT_AIC, CRIT_CV = 0, 1
DBL_MAX = np.finfo(np.float64).max

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_spgcdist.restype = ctypes.c_double
        L.orc_spgcdist.argtypes = [ctypes.c_double] * 4
        L.orc_distance_vec.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_weight_vec.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_adaptive_threshold.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_max_distance.restype = ctypes.c_double
        L.orc_max_distance.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int]
        L.orc_gwr_criterion.restype = ctypes.c_double
        L.orc_gwr_criterion.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_gwr_fit_core.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp,
                                       ctypes.c_int]
        L.orc_gwr_predict.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_indep_vars_criterion.restype = ctypes.c_double
        L.orc_indep_vars_criterion.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _ip, ctypes.c_int]
        L.orc_indep_vars_select.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                            _ip, _ip, _ip, _ip, _dp, _ip]
        L.orc_gwr_run.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, _ip]
        L.orc_mgwr_run.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _ip,
                                   _ip, _ip, _dp, _ip, _ip, _ip, _dp, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_int, _dp, _dp, _dp, _dp, _ip]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _colmajor(a):
    return np.asfortranarray(a, dtype=np.float64)


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _ip_arr(a):
    return a.ctypes.data_as(_ip)


def spgcdist(lon1, lon2, lat1, lat2):
    return lib().orc_spgcdist(lon1, lon2, lat1, lat2)


def distance_vec(coords, focus, geographic=False):
    c = _colmajor(coords)
    out = np.empty(c.shape[0])
    lib().orc_distance_vec(c.shape[0], _p(c), int(geographic), int(focus), _p(out))
    return out


def weight_vec(dist, kernel, adaptive, bw):
    d = _f64(dist)
    out = np.empty_like(d)
    lib().orc_weight_vec(d.shape[0], _p(d), int(kernel), int(adaptive), float(bw), _p(out))
    return out


def adaptive_threshold(coords, bw, focus, geographic=False):
    c = _colmajor(coords)
    out = np.zeros(1)
    lib().orc_adaptive_threshold(c.shape[0], _p(c), int(geographic), float(bw), int(focus), _p(out))
    return out[0]


def max_distance(coords, geographic=False, nthreads=0):
    c = _colmajor(coords)
    return lib().orc_max_distance(c.shape[0], _p(c), int(geographic), int(nthreads))


def criterion(coords, X, y, kernel, adaptive, bw, crit, geographic=False, nthreads=0, focus_count=0):
    c, Xc, yc = _colmajor(coords), _colmajor(X), _f64(y)
    n, k = Xc.shape
    return lib().orc_gwr_criterion(n, k, _p(c), _p(Xc), _p(yc), int(geographic), int(kernel), int(adaptive),
                                   float(bw), int(crit), int(nthreads), int(focus_count))


def fit_core(coords, X, y, kernel, adaptive, bw, geographic=False, nthreads=0, want_S=False, focus_count=0):
    """fitOmp: returns dict(betas, betasSE_raw, shat, qDiag, S) or None on non-SPD."""
    c, Xc, yc = _colmajor(coords), _colmajor(X), _f64(y)
    n, k = Xc.shape
    betas = np.zeros((n, k), order="F")
    se = np.zeros((n, k), order="F")
    shat = np.zeros(2)
    q = np.zeros(n)
    S = np.zeros((n, n), order="F") if want_S else None
    rc = lib().orc_gwr_fit_core(n, k, _p(c), _p(Xc), _p(yc), int(geographic), int(kernel), int(adaptive),
                                float(bw), int(nthreads), _p(betas), _p(se), _p(shat), _p(q), _p(S),
                                int(focus_count))
    if rc != 0:
        return None
    return dict(betas=betas, betasSE_raw=se, shat=shat, qDiag=q, S=S)


def predict(coords, X, y, locations, kernel, adaptive, bw, geographic=False, nthreads=0):
    """predictSerial: betas (m x k) at the locations, or None on non-SPD."""
    c, Xc, yc = _colmajor(coords), _colmajor(X), _f64(y)
    loc = _colmajor(locations)
    n, k = Xc.shape
    m = loc.shape[0]
    betas = np.zeros((m, k), order="F")
    rc = lib().orc_gwr_predict(n, k, _p(c), _p(Xc), _p(yc), int(geographic), m, _p(loc), int(kernel),
                               int(adaptive), float(bw), int(nthreads), _p(betas))
    return None if rc else betas


def indep_vars_criterion(X, y, vars, has_intercept=True):
    """indepVarsSelectionCriterion: AICc of the unit-weight regression on the columns."""
    Xc, yc = _colmajor(X), _f64(y)
    n, k = Xc.shape
    v = np.ascontiguousarray(vars, dtype=np.int32)
    return lib().orc_indep_vars_criterion(n, k, _p(Xc), _p(yc), int(has_intercept), _ip_arr(v), len(v))


def indep_vars_select(X, y, threshold=3.0, has_intercept=True):
    """VariableForwardSelector: (selected columns, [(columns, criterion), ...])."""
    Xc, yc = _colmajor(X), _f64(y)
    n, k = Xc.shape
    sel = np.zeros(k, dtype=np.int32)
    nsel = np.zeros(1, dtype=np.int32)
    cap = k * (k + 1) // 2
    mv = np.zeros(max(cap, 1) * k, dtype=np.int32)
    mn = np.zeros(max(cap, 1), dtype=np.int32)
    mc = np.zeros(max(cap, 1))
    nm = np.zeros(1, dtype=np.int32)
    lib().orc_indep_vars_select(n, k, _p(Xc), _p(yc), int(has_intercept), float(threshold), _ip_arr(sel),
                                _ip_arr(nsel), _ip_arr(mv), _ip_arr(mn), _p(mc), _ip_arr(nm))
    models = [(list(mv[q * k:q * k + mn[q]]), float(mc[q])) for q in range(int(nm[0]))]
    return list(sel[:int(nsel[0])]), models


def gwr_run(coords, X, y, kernel, adaptive, bw, select_bw=False, crit=CRIT_AIC, lower=None, upper=None,
            geographic=False, nthreads=0, want_S=False, trace_cap=256):
    """GWRBasic::fit(): returns dict or raises RuntimeError."""
    c, Xc, yc = _colmajor(coords), _colmajor(X), _f64(y)
    n, k = Xc.shape
    betas = np.zeros((n, k), order="F")
    se = np.zeros((n, k), order="F")
    shat = np.zeros(2)
    q = np.zeros(n)
    S = np.zeros((n, n), order="F") if want_S else None
    diag = np.zeros(7)
    bw_out = np.zeros(1)
    tx = np.zeros(trace_cap)
    tf = np.zeros(trace_cap)
    tlen = np.zeros(1, dtype=np.int32)
    rc = lib().orc_gwr_run(n, k, _p(c), _p(Xc), _p(yc), int(geographic), int(kernel), int(adaptive), float(bw),
                           int(select_bw), int(crit), int(lower is not None), float(lower or 0.0),
                           int(upper is not None), float(upper or 0.0), int(nthreads),
                           _p(betas), _p(se), _p(shat), _p(q), _p(S), _p(diag), _p(bw_out), _p(tx), _p(tf),
                           trace_cap, _ip_arr(tlen))
    if rc == 1:
        raise RuntimeError("inv_sympd(): matrix is singular or not positive definite")
    if rc == 2:
        raise RuntimeError("Invalid initial values.")
    m = min(int(tlen[0]), trace_cap)
    return dict(betas=betas, betasSE=se, shat=shat, qDiag=q, S=S,
                diagnostic=dict(zip(["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"], diag)),
                bandwidth=bw_out[0], trace=list(zip(tx[:m], tf[:m])))


def mgwr_run(coords, X, y, kernels, adaptive, bws, init_type, approach, centered, bw_threshold,
             has_intercept=True, has_hat=True, crit_type=1, crit_threshold=1e-6, max_iter=500,
             retry_times=5, adaptive_lower=10, lower=None, upper=None, geographic=False, nthreads=0):
    c, Xc, yc = _colmajor(coords), _colmajor(X), _f64(y)
    n, k = Xc.shape
    ia = lambda v: np.ascontiguousarray(v, dtype=np.int32)
    kern, ad, it, ap, ce = ia(kernels), ia(adaptive), ia(init_type), ia(approach), ia(centered)
    # one distance per variable (a single flag applies to all)
    geo = ia(geographic if np.ndim(geographic) else [int(bool(geographic))] * k)
    bw = _f64(bws)
    thr = _f64(bw_threshold)
    betas = np.zeros((n, k), order="F")
    bws_out = np.zeros(k)
    diag = np.zeros(7)
    rss0 = np.zeros(1)
    iters = np.zeros(1, dtype=np.int32)
    rc = lib().orc_mgwr_run(n, k, _p(c), _p(Xc), _p(yc), _ip_arr(geo), _ip_arr(kern), _ip_arr(ad), _p(bw),
                            _ip_arr(it), _ip_arr(ap), _ip_arr(ce), _p(thr), int(has_intercept), int(has_hat),
                            int(crit_type), float(crit_threshold), int(max_iter), int(retry_times),
                            int(adaptive_lower), int(lower is not None), float(lower or 0.0),
                            int(upper is not None), float(upper or 0.0), int(nthreads),
                            _p(betas), _p(bws_out), _p(diag), _p(rss0), _ip_arr(iters))
    if rc == 1:
        raise RuntimeError("inv_sympd(): matrix is singular or not positive definite")
    if rc == 2:
        raise RuntimeError("Invalid initial values.")
    return dict(betas=betas, bandwidths=bws_out, rss=rss0[0], iterations=int(iters[0]),
                diagnostic=dict(zip(["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"], diag)))
