"""ctypes binding of the C ABI in include/gwmodel_hip.h.

The shared library is built in-tree (libgwmodel_amd/lib/libgwmodel_hip.so by
libgwmodel_amd/build.py).  There is no fallback: if the library is missing or
no device is present, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgwmodel_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "gwmodel_hip.h")

GWM_OK, GWM_ERR_NOT_SPD, GWM_ERR_CANCELLED, GWM_ERR_HIP, GWM_ERR_OOM, GWM_ERR_INVALID, GWM_ERR_NO_DEVICE, \
    GWM_ERR_INVALID_INITIAL = range(8)
GAUSSIAN, EXPONENTIAL, BISQUARE, TRICUBE, BOXCAR = range(5)
EUCLIDEAN, GREAT_CIRCLE = 0, 1
CRIT_AIC, CRIT_CV = 0, 1
FP64, MIXED = 0, 1
DBL_MAX = np.finfo(np.float64).max

STATUS_NAMES = {0: "OK", 1: "NOT_SPD", 2: "CANCELLED", 3: "HIP_ERROR", 4: "OOM", 5: "INVALID", 6: "NO_DEVICE",
                7: "INVALID_INITIAL"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
STOP_CALLBACK = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p)
PRINT_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p)
PROGRESS_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_double, ctypes.c_double)
LOG_ERR, LOG_INFO = 3, 6

# name -> (restype, argtypes)
SIGNATURES = {
    "gwm_hip_device_count": (_i, [_ip]),
    "gwm_hip_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "gwm_hip_destroy": (None, [_vp]),
    "gwm_hip_last_error": (ctypes.c_char_p, [_vp]),
    "gwm_hip_set_stop_callback": (_i, [_vp, STOP_CALLBACK, _vp]),
    "gwm_hip_set_telegram": (_i, [_vp, PRINT_CALLBACK, PROGRESS_CALLBACK, _vp]),
    "gwm_hip_set_data": (_i, [_vp, _i, _i, _dp, _dp, _dp, _i]),
    "gwm_hip_criterion": (_i, [_vp, _i, _i, _d, _i, _i, _dp]),
    "gwm_hip_fit": (_i, [_vp, _i, _i, _d, _i, _dp, _dp, _dp, _dp, _dp]),
    "gwm_hip_predict": (_i, [_vp, _i, _dp, _i, _i, _d, _i, _dp]),
    "gwm_hip_criterion_noalloc": (_i, [_vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _d, _i, _dp]),
    "gwm_hip_mgwr_set_hat_matrix": (_i, [_vp, _i]),
    "gwm_hip_mgwr_set_distances": (_i, [_vp, _i, _ip]),
    "gwm_hip_last_error_bound": (_i, [_vp, _dp]),
    "gwm_hip_set_autotune": (_i, [_vp, _i]),
    "gwm_hip_set_distance_table": (_i, [_vp, _d]),
    "gwm_hip_set_neighbour_table": (_i, [_vp, _d]),
    "gwm_hip_set_certify_table": (_i, [_vp, _d]),
    "gwm_hip_autotune_report": (_i, [ctypes.c_char_p, _i]),
    "gwm_hip_mgwr_last_shat": (_i, [_vp, _dp]),
    "gwm_hip_indep_vars_criterion": (_i, [_vp, _i, _ip, _i, _dp]),
    "gwm_hip_indep_vars_select": (_i, [_vp, _i, _d, _ip, _ip, _ip, _ip, _dp, _i, _ip]),
    "gwm_hip_select_bandwidth": (_i, [_vp, _i, _i, _i, _i, _i, _d, _i, _d, _dp, _dp, _dp, _i, _ip]),
    "gwm_hip_gwr_run": (_i, [_vp, _i, _i, _d, _i, _i, _i, _i, _d, _i, _d, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                             _dp, _i, _ip]),
    "gwm_hip_max_distance": (_i, [_vp, _dp]),
    "gwm_hip_distance_row": (_i, [_vp, _i, _dp]),
    "gwm_hip_adaptive_bandwidths": (_i, [_vp, _d, _dp]),
    "gwm_hip_sqrt_boundary": (_i, [_d, _dp]),
    "gwm_hip_var_criterion": (_i, [_vp, _dp, _dp, _i, _i, _d, _i, _i, _dp]),
    "gwm_hip_var_fit": (_i, [_vp, _dp, _dp, _i, _i, _d, _i, _dp]),
    "gwm_hip_mgwr_run": (_i, [_vp, _ip, _ip, _dp, _ip, _ip, _ip, _dp, _i, _i, _d, _i, _i, _i, _i, _d, _i, _d, _i,
                              _dp, _dp, _dp, _dp, _ip]),
    "gwm_hip_set_certify_tolerance": (_i, [_vp, _d]),
    "gwm_hip_set_refine_condition": (_i, [_vp, _d, _d]),
    "gwm_hip_set_pruning": (_i, [_vp, _i]),
    "gwm_hip_kernel_pairs": (_i, [_vp, ctypes.c_char_p, _dp]),
    "gwm_hip_last_condition": (_i, [_vp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "gwm_hip_set_timing": (_i, [_vp, _i]),
    "gwm_hip_kernel_time": (_i, [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_longlong), _dp]),
    "gwm_hip_reset_timers": (_i, [_vp]),
    "gwm_hip_solve_count": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "gwm_hip_server_create": (_i, [_i, _i, _d, _i, ctypes.POINTER(_vp)]),
    "gwm_hip_server_set_certify_tolerance": (_i, [_vp, _d]),
    "gwm_hip_server_set_refine_condition": (_i, [_vp, _d, _d]),
    "gwm_hip_server_set_round_ms": (_i, [_vp, _d]),
    "gwm_hip_server_set_partial_budget": (_i, [_vp, _d]),
    "gwm_hip_server_autotune_report": (_i, [_vp, ctypes.c_char_p, _i]),
    "gwm_hip_server_submit": (_i, [_vp, _i, _i, _dp, _dp, _dp, _i, _i, _i, _d, _i, _i, _i, _d, _i, _d, _dp, _dp, _dp,
                                   _dp, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "gwm_hip_server_wait": (_i, [_vp, ctypes.c_longlong, _ip, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "gwm_hip_server_stats": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), _dp, _dp,
                                  _ip]),
    "gwm_hip_server_last_error": (ctypes.c_char_p, [_vp]),
    "gwm_hip_server_destroy": (None, [_vp]),
}

_lib = None


class GwmError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), message))


def load():
    """Load the in-tree HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libgwmodel_hip.so is not built (run __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _ipa(a):
    return a.ctypes.data_as(_ip)


def device_count():
    c = ctypes.c_int(0)
    load().gwm_hip_device_count(ctypes.byref(c))
    return c.value


class Context:
    """One device, one stream, device-resident problem data."""

    def __init__(self, device=0):
        self._lib = load()
        h = _vp()
        rc = self._lib.gwm_hip_create(device, ctypes.byref(h))
        if rc != GWM_OK:
            raise GwmError(rc, "cannot create a HIP context on device %d" % device)
        self.h = h
        self.n = 0
        self.k = 0
        self._stop_cb = None
        self._tele_cb = None

    def close(self):
        if getattr(self, "h", None):
            self._lib.gwm_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != GWM_OK:
            raise GwmError(rc, self._lib.gwm_hip_last_error(self.h).decode())

    def set_stop_callback(self, fn):
        """fn() -> bool; polled between sweeps (ITelegram::stop)."""
        if fn is None:
            self._stop_cb = None
            self._check(self._lib.gwm_hip_set_stop_callback(self.h, STOP_CALLBACK(), None))
            return
        self._stop_cb = STOP_CALLBACK(lambda _u: 1 if fn() else 0)
        self._check(self._lib.gwm_hip_set_stop_callback(self.h, self._stop_cb, None))

    def set_telegram(self, print_fn=None, progress_fn=None):
        """ITelegram: print_fn(level, message) and progress_fn(current, total)
        (total 0: current is a fraction); None switches a channel off."""
        pcb = PRINT_CALLBACK(lambda _u, lv, msg: print_fn(lv, msg.decode())) if print_fn else PRINT_CALLBACK()
        gcb = PROGRESS_CALLBACK(lambda _u, cur, tot: progress_fn(cur, tot)) if progress_fn else PROGRESS_CALLBACK()
        self._tele_cb = (pcb, gcb)
        self._check(self._lib.gwm_hip_set_telegram(self.h, pcb, gcb, None))

    def mgwr_set_distances(self, distances):
        """One distance per variable for mgwr_run (None or []: the data's)."""
        d = np.ascontiguousarray(distances if distances is not None else [], dtype=np.int32)
        self._check(self._lib.gwm_hip_mgwr_set_distances(self.h, len(d), _ipa(d)))

    def set_data(self, coords, X, y, distance=EUCLIDEAN):
        c = np.asfortranarray(coords, dtype=np.float64)
        Xf = np.asfortranarray(X, dtype=np.float64)
        yf = np.ascontiguousarray(y, dtype=np.float64)
        n, k = Xf.shape
        assert c.shape == (n, 2) and yf.shape == (n,)
        self._check(self._lib.gwm_hip_set_data(self.h, n, k, _p(c), _p(Xf), _p(yf), int(distance)))
        self.n, self.k = n, k
        self.X, self.y = Xf, yf

    def criterion(self, kernel, adaptive, bw, crit, precision=FP64):
        """Returns the criterion, or DBL_MAX when a local design is not SPD."""
        v = np.zeros(1)
        rc = self._lib.gwm_hip_criterion(self.h, int(kernel), int(adaptive), float(bw), int(crit), int(precision),
                                         _p(v))
        if rc == GWM_ERR_NOT_SPD:
            return DBL_MAX
        self._check(rc)
        return float(v[0])

    def fit(self, kernel, adaptive, bw, precision=FP64, want_qdiag=True, want_S=False):
        n, k = self.n, self.k
        betas = np.zeros((n, k), order="F")
        se = np.zeros((n, k), order="F")
        shat = np.zeros(2)
        q = np.zeros(n) if want_qdiag else None
        S = np.zeros((n, n), order="F") if want_S else None
        self._check(self._lib.gwm_hip_fit(self.h, int(kernel), int(adaptive), float(bw), int(precision), _p(betas),
                                          _p(se), _p(shat), _p(q), _p(S)))
        return dict(betas=betas, betasSE_raw=se, shat=shat, qDiag=q, S=S)

    def criterion_noalloc(self, n, k, d_coords, d_X, d_y, kernel, adaptive, bw, crit, distance=EUCLIDEAN):
        """gwm_hip_criterion_noalloc on device arrays (integer device
        addresses): returns (status, value); the library allocates nothing."""
        v = np.zeros(1)
        rc = self._lib.gwm_hip_criterion_noalloc(self.h, int(n), int(k), d_coords, d_X, d_y, int(distance),
                                                 int(kernel), int(adaptive), float(bw), int(crit), _p(v))
        return rc, float(v[0])

    def last_error_message(self):
        return (self._lib.gwm_hip_last_error(self.h) or b"").decode()

    def set_autotune(self, on=True):
        """Online autotuning of the sweep launch shapes and the longest-first work
        order of large pruned sweeps (results do not change)."""
        self._check(self._lib.gwm_hip_set_autotune(self.h, int(bool(on))))

    def force_shape(self, shape):
        """Every sweep in launch shape 0..3 (testing: shapes give the same bits)."""
        self._check(self._lib.gwm_hip_set_autotune(self.h, 2 + int(shape)))

    def last_error_bound(self):
        """Error bound the certified search assumes for the last mixed criterion value."""
        out = np.zeros(1)
        self._check(self._lib.gwm_hip_last_error_bound(self.h, _p(out)))
        return float(out[0])

    def mgwr_set_hat_matrix(self, on=True):
        """Keep the multiscale hat matrix (tr S0, tr S0^T S0 in the diagnostics)."""
        self._check(self._lib.gwm_hip_mgwr_set_hat_matrix(self.h, int(on)))

    def mgwr_last_shat(self):
        out = np.zeros(2)
        self._check(self._lib.gwm_hip_mgwr_last_shat(self.h, _p(out)))
        return out

    def predict(self, locations, kernel, adaptive, bw, precision=FP64):
        """Betas (m x k) at the rows of `locations` (m x 2)."""
        loc = np.asfortranarray(np.asarray(locations, dtype=np.float64).reshape(-1, 2))
        m = loc.shape[0]
        betas = np.zeros((m, self.k), order="F")
        self._check(self._lib.gwm_hip_predict(self.h, m, _p(loc), int(kernel), int(adaptive), float(bw),
                                              int(precision), _p(betas)))
        return betas

    def indep_vars_criterion(self, vars, has_intercept=True):
        """AICc of the unit-weight regression on columns {0} + vars (DBL_MAX if singular)."""
        v = np.ascontiguousarray(vars, dtype=np.int32)
        out = np.zeros(1)
        self._check(self._lib.gwm_hip_indep_vars_criterion(self.h, int(has_intercept), _ipa(v), len(v), _p(out)))
        return float(out[0])

    def indep_vars_select(self, threshold=3.0, has_intercept=True):
        """Forward selection: (selected columns, [(columns, criterion), ...])."""
        k = self.k
        cap = k * (k + 1) // 2
        sel = np.zeros(k, dtype=np.int32)
        nsel = np.zeros(1, dtype=np.int32)
        mv = np.zeros(cap * k, dtype=np.int32)
        mn = np.zeros(cap, dtype=np.int32)
        mc = np.zeros(cap)
        nm = np.zeros(1, dtype=np.int32)
        self._check(self._lib.gwm_hip_indep_vars_select(self.h, int(has_intercept), float(threshold), _ipa(sel),
                                                        _ipa(nsel), _ipa(mv), _ipa(mn), _p(mc), cap, _ipa(nm)))
        models = [(list(mv[q * k:q * k + mn[q]]), float(mc[q])) for q in range(min(int(nm[0]), cap))]
        return list(sel[:int(nsel[0])]), models

    def select_bandwidth(self, kernel, adaptive, crit, precision=FP64, lower=None, upper=None, trace_cap=256):
        bw = np.zeros(1)
        tx, tf = np.zeros(trace_cap), np.zeros(trace_cap)
        tl = np.zeros(1, dtype=np.int32)
        self._check(self._lib.gwm_hip_select_bandwidth(
            self.h, int(kernel), int(adaptive), int(crit), int(precision), int(lower is not None), float(lower or 0),
            int(upper is not None), float(upper or 0), _p(bw), _p(tx), _p(tf), trace_cap, _ipa(tl)))
        m = min(int(tl[0]), trace_cap)
        return float(bw[0]), list(zip(tx[:m], tf[:m]))

    def gwr_run(self, kernel, adaptive, bw, select_bw=False, crit=CRIT_AIC, precision=FP64, lower=None, upper=None,
                want_S=False, trace_cap=256):
        n, k = self.n, self.k
        betas = np.zeros((n, k), order="F")
        se = np.zeros((n, k), order="F")
        shat = np.zeros(2)
        q = np.zeros(n)
        S = np.zeros((n, n), order="F") if want_S else None
        diag = np.zeros(7)
        bw_out = np.zeros(1)
        tx, tf = np.zeros(trace_cap), np.zeros(trace_cap)
        tl = np.zeros(1, dtype=np.int32)
        self._check(self._lib.gwm_hip_gwr_run(
            self.h, int(kernel), int(adaptive), float(bw), int(select_bw), int(crit), int(precision),
            int(lower is not None), float(lower or 0), int(upper is not None), float(upper or 0), _p(betas), _p(se),
            _p(shat), _p(q), _p(S), _p(diag), _p(bw_out), _p(tx), _p(tf), trace_cap, _ipa(tl)))
        m = min(int(tl[0]), trace_cap)
        return dict(betas=betas, betasSE=se, shat=shat, qDiag=q, S=S,
                    diagnostic=dict(zip(["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"], diag)),
                    bandwidth=float(bw_out[0]), trace=list(zip(tx[:m], tf[:m])))

    def max_distance(self):
        v = np.zeros(1)
        self._check(self._lib.gwm_hip_max_distance(self.h, _p(v)))
        return float(v[0])

    def distance_row(self, focus):
        out = np.zeros(self.n)
        self._check(self._lib.gwm_hip_distance_row(self.h, int(focus), _p(out)))
        return out

    def adaptive_bandwidths(self, bw):
        out = np.zeros(self.n)
        self._check(self._lib.gwm_hip_adaptive_bandwidths(self.h, float(bw), _p(out)))
        return out

    def var_criterion(self, x, y, kernel, adaptive, bw, crit, precision=FP64):
        xv, yv = np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)
        v = np.zeros(1)
        rc = self._lib.gwm_hip_var_criterion(self.h, _p(xv), _p(yv), int(kernel), int(adaptive), float(bw), int(crit),
                                             int(precision), _p(v))
        if rc == GWM_ERR_NOT_SPD:
            return DBL_MAX
        self._check(rc)
        return float(v[0])

    def var_fit(self, x, y, kernel, adaptive, bw, precision=FP64):
        xv, yv = np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)
        out = np.zeros(self.n)
        self._check(self._lib.gwm_hip_var_fit(self.h, _p(xv), _p(yv), int(kernel), int(adaptive), float(bw),
                                              int(precision), _p(out)))
        return out

    def mgwr_run(self, kernels, adaptive, bws, init_type, approach, centered, bw_threshold, has_intercept=True,
                 crit_type=1, crit_threshold=1e-6, max_iter=500, retry_times=5, adaptive_lower=10, lower=None,
                 upper=None, precision=FP64):
        k = self.k
        ia = lambda v: np.ascontiguousarray(v, dtype=np.int32)
        kern, ad, it, ap, ce = ia(kernels), ia(adaptive), ia(init_type), ia(approach), ia(centered)
        bwv = np.ascontiguousarray(bws, dtype=np.float64)
        thr = np.ascontiguousarray(bw_threshold, dtype=np.float64)
        betas = np.zeros((self.n, k), order="F")
        bws_out = np.zeros(k)
        diag = np.zeros(7)
        rss = np.zeros(1)
        iters = np.zeros(1, dtype=np.int32)
        self._check(self._lib.gwm_hip_mgwr_run(
            self.h, _ipa(kern), _ipa(ad), _p(bwv), _ipa(it), _ipa(ap), _ipa(ce), _p(thr), int(has_intercept),
            int(crit_type), float(crit_threshold), int(max_iter), int(retry_times), int(adaptive_lower),
            int(lower is not None), float(lower or 0), int(upper is not None), float(upper or 0), int(precision),
            _p(betas), _p(bws_out), _p(diag), _p(rss), _ipa(iters)))
        return dict(betas=betas, bandwidths=bws_out, rss=float(rss[0]), iterations=int(iters[0]),
                    diagnostic=dict(zip(["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"], diag)))

    def set_certify_tolerance(self, rel):
        self._check(self._lib.gwm_hip_set_certify_tolerance(self.h, float(rel)))

    def set_refine_condition(self, criterion_limit, fit_limit):
        """Mixed precision: focus points whose conditioning indicator exceeds
        the limit are recomputed in FP64 (criterion evaluations / fits)."""
        self._check(self._lib.gwm_hip_set_refine_condition(self.h, float(criterion_limit), float(fit_limit)))

    def set_pruning(self, on=True):
        """Skip data chunks whose kernel weights are all exactly zero (bit-identical results)."""
        self._check(self._lib.gwm_hip_set_pruning(self.h, int(on)))

    def last_condition(self):
        """(indicator maximum, focus points recomputed in FP64) of the last mixed evaluation."""
        v = np.zeros(1)
        r = ctypes.c_longlong(0)
        self._check(self._lib.gwm_hip_last_condition(self.h, _p(v), ctypes.byref(r)))
        return float(v[0]), r.value

    def set_certify_table(self, max_bytes):
        """Byte cap of the FP64 distance table of certified searches (0: off)."""
        self._check(self._lib.gwm_hip_set_certify_table(self.h, float(max_bytes)))

    def set_neighbour_table(self, max_bytes):
        """Byte cap of the adaptive-bandwidth neighbour table (0: off)."""
        self._check(self._lib.gwm_hip_set_neighbour_table(self.h, float(max_bytes)))

    def set_distance_table(self, max_bytes):
        """Cap of the great-circle distance table of search / fit / backfit calls (0: off)."""
        self._check(self._lib.gwm_hip_set_distance_table(self.h, float(max_bytes)))

    def set_timing(self, on=True):
        self._check(self._lib.gwm_hip_set_timing(self.h, int(on)))

    def kernel_time(self, name):
        n = ctypes.c_longlong(0)
        ms = ctypes.c_double(0)
        self._check(self._lib.gwm_hip_kernel_time(self.h, name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value

    def kernel_pairs(self, name):
        """Focus x data pairs the named sweeps processed while timing was on."""
        v = np.zeros(1)
        self._check(self._lib.gwm_hip_kernel_pairs(self.h, name.encode(), _p(v)))
        return float(v[0])

    def reset_timers(self):
        self._check(self._lib.gwm_hip_reset_timers(self.h))

    def solve_count(self):
        n = ctypes.c_longlong(0)
        self._check(self._lib.gwm_hip_solve_count(self.h, ctypes.byref(n)))
        return n.value


class Server:
    """Continuous-batching server: many GWRBasic requests in flight on one device."""

    def __init__(self, device=0, precision=FP64, hbm_budget_bytes=0.0, max_inflight=64):
        self._lib = load()
        h = _vp()
        rc = self._lib.gwm_hip_server_create(device, int(precision), float(hbm_budget_bytes), int(max_inflight),
                                             ctypes.byref(h))
        if rc != GWM_OK:
            raise GwmError(rc, "cannot create a server on device %d" % device)
        self.h = h
        self._live = {}

    def close(self):
        if getattr(self, "h", None):
            self._lib.gwm_hip_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_certify_tolerance(self, rel):
        rc = self._lib.gwm_hip_server_set_certify_tolerance(self.h, float(rel))
        if rc != GWM_OK:
            raise GwmError(rc)

    def set_refine_condition(self, criterion_limit, fit_limit):
        rc = self._lib.gwm_hip_server_set_refine_condition(self.h, float(criterion_limit), float(fit_limit))
        if rc != GWM_OK:
            raise GwmError(rc)

    def set_partial_budget(self, nbytes):
        """Partial-sum bytes of one request's round slice (later submissions)."""
        rc = self._lib.gwm_hip_server_set_partial_budget(self.h, float(nbytes))
        if rc != GWM_OK:
            raise GwmError(rc)

    def autotune_report(self):
        """Lines of the server's batched-sweep tuning table."""
        buf = ctypes.create_string_buffer(1 << 16)
        self._lib.gwm_hip_server_autotune_report(self.h, buf, len(buf))
        return [l for l in buf.value.decode().splitlines() if l]

    def set_round_ms(self, ms):
        """Work budget of a round in ms of device time (<= 0: whole evaluations)."""
        rc = self._lib.gwm_hip_server_set_round_ms(self.h, float(ms))
        if rc != GWM_OK:
            raise GwmError(rc)

    def submit(self, coords, X, y, kernel, adaptive, bw, select_bw=False, crit=CRIT_CV, distance=EUCLIDEAN,
               lower=None, upper=None):
        """Queue one request; returns its id.  Results are collected by wait()."""
        c = np.asfortranarray(coords, dtype=np.float64)
        Xf = np.asfortranarray(X, dtype=np.float64)
        yf = np.ascontiguousarray(y, dtype=np.float64)
        n, k = Xf.shape
        out = dict(betas=np.zeros((n, k), order="F"), betasSE=np.zeros((n, k), order="F"), shat=np.zeros(2),
                   qDiag=np.zeros(n), diag=np.zeros(7), bandwidth=np.zeros(1))
        rid = ctypes.c_longlong(0)
        rc = self._lib.gwm_hip_server_submit(
            self.h, n, k, _p(c), _p(Xf), _p(yf), int(distance), int(kernel), int(adaptive), float(bw),
            int(select_bw), int(crit), int(lower is not None), float(lower or 0), int(upper is not None),
            float(upper or 0), _p(out["betas"]), _p(out["betasSE"]), _p(out["shat"]), _p(out["qDiag"]),
            _p(out["diag"]), _p(out["bandwidth"]), ctypes.byref(rid))
        if rc != GWM_OK:
            raise GwmError(rc, "submit failed")
        self._live[rid.value] = (out, c, Xf, yf)
        return rid.value

    def wait(self, rid):
        """Blocks until request rid is done; returns its results (raises on error)."""
        st = ctypes.c_int(0)
        lat = ctypes.c_double(0)
        sv = ctypes.c_longlong(0)
        rc = self._lib.gwm_hip_server_wait(self.h, ctypes.c_longlong(rid), ctypes.byref(st), ctypes.byref(lat),
                                           ctypes.byref(sv))
        if rc != GWM_OK:
            raise GwmError(rc, "wait failed")
        out = self._live.pop(rid)[0]
        res = dict(status=st.value, latency_ms=lat.value, solves=sv.value, betas=out["betas"],
                   betasSE=out["betasSE"], shat=out["shat"], qDiag=out["qDiag"], bandwidth=float(out["bandwidth"][0]),
                   diagnostic=dict(zip(["RSS", "AIC", "AICc", "ENP", "EDF", "RSquare", "RSquareAdjust"],
                                       out["diag"])))
        return res

    def stats(self):
        r = ctypes.c_longlong(0)
        s = ctypes.c_longlong(0)
        b = ctypes.c_double(0)
        pb = ctypes.c_double(0)
        pi = ctypes.c_int(0)
        self._lib.gwm_hip_server_stats(self.h, ctypes.byref(r), ctypes.byref(s), ctypes.byref(b), ctypes.byref(pb),
                                       ctypes.byref(pi))
        return dict(rounds=r.value, solves=s.value, busy_ms=b.value, peak_bytes=pb.value, peak_inflight=pi.value)


def autotune_report():
    """Lines of the process-wide autotuning table."""
    buf = ctypes.create_string_buffer(1 << 16)
    load().gwm_hip_autotune_report(buf, len(buf))
    return [l for l in buf.value.decode().splitlines() if l]


def header_symbols(path=HEADER_PATH):
    """Function names declared in include/gwmodel_hip.h."""
    import re
    text = open(path).read()
    return sorted(set(re.findall(r"\b(gwm_hip_[a-z_]+)\s*\(", text)))
