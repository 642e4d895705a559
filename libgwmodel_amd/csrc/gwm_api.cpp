// Host side of the C ABI declared in include/gwmodel_hip.h.
//
// One context = one device, one HIP stream and the device-resident problem
// (coordinates, design, product table Z, workspaces).  Everything n x n stays
// implicit: sweeps stream Z through the fused contraction kernel, so device
// memory is O(n * k^2) per problem.  All HIP failures are reported as status
// codes; the device is never reset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "gwm_internal.h"

using namespace gwm;
using namespace gwm::detail;

namespace gwm {
namespace detail {

int nta_of(int k) { return (k * (k + 1) / 2 + k + 15) / 16; }
int ntb_of(int k) { return (k * (k + 1) / 2 + 15) / 16; }

}  // namespace detail
}  // namespace gwm

namespace {

// ITelegram::LogLevel (include/gwmodelpp/Logger.h:38-48)
constexpr int kLogErr = 3;
constexpr int kLogInfo = 6;
const std::string kStage = "#stage ";               // GWM_LOG_TAG_STAGE (Algorithm.h:102)
const std::string kBackfit = "#back-fitting ";      // GWM_LOG_TAG_MGWR_BACKFITTING (GWRMultiscale.h:20)
const std::string kInitialBw = "#initial-bandwidth ";   // GWM_LOG_TAG_MGWR_INITIAL_BW (GWRMultiscale.h:19)


}  // namespace

namespace gwm {
namespace detail {

void to_sorted(const gwm_hip_ctx* c, const double* src, double* dst, int cols) {
    const size_t n = c->n;
    for (int j = 0; j < cols; ++j)
        for (size_t s = 0; s < n; ++s) dst[s + j * n] = src[(size_t)c->perm[s] + j * n];
}

void to_caller(const gwm_hip_ctx* c, const double* src, double* dst, int cols) {
    const size_t n = c->n;
    for (int j = 0; j < cols; ++j)
        for (size_t s = 0; s < n; ++s) dst[(size_t)c->perm[s] + j * n] = src[s + j * n];
}

void calc_diagnostic(int n_, int k, const double* X, const double* y, const double* betas, const double* shat,
                     double* diag) {
    double rss = 0.0;
    for (int i = 0; i < n_; ++i) {
        double f = 0.0;
        for (int a = 0; a < k; ++a) f += betas[i + (size_t)a * n_] * X[i + (size_t)a * n_];
        double r = y[i] - f;
        rss += r * r;
    }
    const double n = (double)n_;
    double ym = 0.0;
    for (int i = 0; i < n_; ++i) ym += y[i];
    ym /= n;
    double yss = 0.0;
    for (int i = 0; i < n_; ++i) yss += (y[i] - ym) * (y[i] - ym);
    double AIC = n * std::log(rss / n) + n * std::log(2 * kPi) + n + shat[0];
    double AICc = aicc(rss, n, shat[0]);
    double edf = n - 2 * shat[0] + shat[1];
    double enp = 2 * shat[0] - shat[1];
    double r2 = 1 - rss / yss;
    double r2adj = 1 - (1 - r2) * (n - 1) / (edf - 1);
    diag[0] = rss; diag[1] = AIC; diag[2] = AICc; diag[3] = enp; diag[4] = edf; diag[5] = r2; diag[6] = r2adj;
}

double aicc(double rss, double n, double trS) {
    return n * std::log(rss / n) + n * std::log(2 * kPi) + n * ((n + trS) / (n - 2 - trS));
}

// Absolute error bound of a mixed-precision criterion value, for the
// certified golden section: the largest of three terms, each covering one
// error source.  Their constants come from the errors measured against FP64
// at 12,062 criterion evaluations (every golden-section candidate of CV and
// AICc searches over 320 problems of the uneven batch workload, the n = 2000
// workload with three kernels, and the n = 100k great-circle AICc workload
// from 56 to 2200 km): with them no evaluation exceeded 0.69 of its bound.
//  - tol_rel |f| for AICc, 3 tol_rel |f| for CV (whose relative errors run
//    about three times larger at the same bandwidth);
//  - the rounding of the fitted values, random over the focus points: an
//    error of yhat_i relative to itself moves res_i^2 by 2 res_i yhat_i times
//    it, so RSS (or CV) moves by 2 e sqrt(q) with q = sum_i (res_i yhat_i)^2
//    and e = kRoundC 2^-22 kappa, kappa = min(indicator maximum, FP64
//    recomputation limit) >= 1; AICc = n ln(RSS/n) + ... turns it into
//    n 2 e sqrt(q) / RSS;
//  - AICc only: the trace carries the systematic relative error of the mixed
//    weights (it does not cancel as in beta), and AICc amplifies an error of
//    trS by S = n (2n - 2) / (n - 2 - trS)^2: tol_rel S trS.  With the previous
//    term it bounds the error where AICc crosses zero and |f| says nothing
//    of it.
constexpr double kRoundC = 4.0;
constexpr double kCvTol = 3.0;
double crit_error_bound(int crit, double value, double rss, double trS, double q, int n, double cond_max, double limit,
                        double tol_rel) {
    const double nd = n;
    const double kappa = std::max(1.0, std::min(cond_max, limit));
    const double dr = 2.0 * kRoundC * std::ldexp(1.0, -22) * kappa * std::sqrt(std::max(q, 0.0));
    if (crit == GWM_CRIT_CV) return std::max(kCvTol * tol_rel * std::fabs(value), dr);
    const double b = tol_rel * std::fabs(value);
    const double gap = nd - 2 - trS;
    const double S = gap > 0 ? nd * (2 * nd - 2) / (gap * gap) : INFINITY;
    return std::max(b, std::max(rss > 0 ? nd * dr / rss : INFINITY, tol_rel * S * std::fabs(trS)));
}

}  // namespace detail
}  // namespace gwm

namespace {

int check_ready(gwm_hip_ctx* c) {
    if (!c) return GWM_ERR_INVALID;
    if (c->n <= 0) {
        c->err = "no data: call gwm_hip_set_data first";
        return GWM_ERR_INVALID;
    }
    return GWM_OK;
}

int check_kernel(gwm_hip_ctx* c, int kernel, int precision) {
    if (kernel < 0 || kernel > 4) {
        c->err = "unknown kernel";
        return GWM_ERR_INVALID;
    }
    if (precision != GWM_FP64 && precision != GWM_MIXED) {
        c->err = "unknown precision";
        return GWM_ERR_INVALID;
    }
    return GWM_OK;
}

// Criterion of the full design: CV (GWRBasic.cpp:440-485) or AICc (:487-535).
int criterion_impl(gwm_hip_ctx* c, int kernel, int adaptive, double bw, int crit, int precision, double* value,
                   double* bound = nullptr) {
    *value = DBL_MAX;
    if (bound) *bound = 0.0;
    int rc = check_kernel(c, kernel, precision);
    if (rc) return rc;
    if (stopped(c)) return GWM_ERR_CANCELLED;
    const double* bw_pt;
    if ((rc = resolve_bandwidth(c, adaptive, bw, &bw_pt))) return rc;
    // CV zeroes the focus point's own weight: its largest weight is that of
    // its nearest other point
    const double* dnn = nullptr;
    if (crit == GWM_CRIT_CV) {
        if ((rc = ensure_dnn(c))) return rc;
        dnn = c->dnn;
    }
    GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
    int PT, nseg;
    const int mode = crit == GWM_CRIT_CV ? kModeCV : kModePlain;
    if ((rc = sweep(c, mode, c->nta, 0, design_planes(c), bw_pt, bw, kernel, precision, &PT, &nseg, dnn))) return rc;
    SolveArgs s = solve_args(c, PT, nseg, 0, bw_pt, bw, kernel);
    if (precision == GWM_MIXED) s.dnn = dnn;
    const int smode = crit == GWM_CRIT_CV ? kSolveCV : kSolveAIC;
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_solve(c->k, smode, s, c->stream));
    if ((rc = timer_end(c, "solve"))) return rc;
    if (precision == GWM_MIXED) {
        int count = 0;
        if ((rc = refine_fp64(c, mode, smode, 0, 0, bw_pt, bw, kernel, c->refine_crit, &count, nullptr, dnn)))
            return rc;
    }
    // sums, and (mixed) the indicator maximum over the final per-point values
    const bool mixed = precision == GWM_MIXED;
    GWM_HIP_CHECK(c, launch_reduce(c->res2, crit == GWM_CRIT_CV ? nullptr : c->sii, mixed ? c->ssi : nullptr,
                                   mixed ? c->cond : nullptr, c->n, c->red, c->stream));
    double red[4];
    int fail = 0;
    GWM_HIP_CHECK(c, hipMemcpyAsync(red, c->red, sizeof(red), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->solves += c->n;
    if (fail) return GWM_ERR_NOT_SPD;
    double v = crit == GWM_CRIT_CV ? red[0] : aicc(red[0], (double)c->n, red[1]);
    if (!std::isfinite(v)) return GWM_ERR_NOT_SPD;
    *value = v;
    c->last_bound =
        mixed ? crit_error_bound(crit, v, red[0], red[1], red[2], c->n, red[3], c->refine_crit, c->certify_rel) : 0.0;
    if (bound) *bound = c->last_bound;
    return GWM_OK;
}

// Golden-section search, BandwidthSelector::optimize (BandwidthSelector.cpp:6-81).
// f(x, exact, &v, &b) returns a status and the criterion (DBL_MAX when
// invalid); exact = true asks for the FP64 evaluation.  A negative bound b
// reports that v was evaluated in FP64 although exact was not asked for.
//
// With certify set (mixed-precision sweeps), every decision the search takes
// on criterion values -- the branch f1 < f2, the stopping test |f2 - f1| > eps
// and the final choice -- is certified: when a value's mixed-precision error
// bound (tol_rel * |f|) could flip the decision, both points are re-evaluated
// in FP64 and the decision is taken on the FP64 values.  The bound of a mixed
// value is the evaluation's own (crit_error_bound), at least tol_rel |f|.  The iterates are
// therefore those of the FP64 search.  Once a decision has needed the FP64
// values, the search has entered the region where the criterion varies less
// than the mixed-precision error, and later candidates are evaluated in FP64
// directly instead of twice.
}  // namespace

namespace gwm {
namespace detail {

int golden(bool adaptBw, double lower, double upper, const CritFn& f, bool certify, double tol_rel,
           double* xopt_out, std::map<double, double>& log, const EvalReport& report, double* upper_out) {
    double xU = upper, xL = lower;
    const double eps = 1e-4;
    const double R = (std::sqrt(5.0) - 1) / 2;
    int iter = 0;
    double d = R * (xU - xL);
    double x1 = adaptBw ? std::floor(xL + d) : (xL + d);
    double x2 = adaptBw ? std::round(xU - d) : (xU - d);
    double f1 = DBL_MAX, f2 = DBL_MAX;
    double b1 = 0.0, b2 = 0.0;   // error bounds of mixed values (crit_error_bound)
    bool e1 = !certify, e2 = !certify;
    bool sticky = false;   // FP64 from here on
    // evaluate; a failed (DBL_MAX) mixed value is confirmed in FP64
    auto eval = [&](double x, bool& exact, double* fx, double* fb) -> int {
        *fb = 0.0;
        if (upper_out) *upper_out = xU;
        int rc = f(x, exact, fx, fb);
        if (rc == GWM_ERR_NOT_SPD) {
            *fx = DBL_MAX;
            rc = GWM_OK;
        }
        if (*fb < 0.0) {
            exact = true;
            *fb = 0.0;
        }
        if (rc == GWM_OK && !exact && *fx == DBL_MAX) {
            exact = true;
            rc = f(x, true, fx, fb);
            if (rc == GWM_ERR_NOT_SPD) {
                *fx = DBL_MAX;
                rc = GWM_OK;
            }
        }
        if (rc == GWM_OK && report) report(x, *fx);
        return rc;
    };
    auto record = [&](double x, double fx) {
        if (fx < DBL_MAX) log[x] = fx;
    };
    // a mixed value's bound: the evaluation's own, never below tol_rel |f|
    auto bound = [&](double fx, bool exact, double fb) {
        return exact || fx == DBL_MAX ? 0.0 : std::max(fb, tol_rel * std::fabs(fx));
    };
    // upgrade both points to FP64 when a decision on (f1, f2) is not certain;
    // coincident points (rounded adaptive candidates) hold one evaluation, so
    // f2 - f1 = 0 in either precision and every decision is the FP64 one
    auto certify_pair = [&]() -> int {
        if (e1 && e2) return GWM_OK;
        if (x1 == x2) return GWM_OK;
        double m = bound(f1, e1, b1) + bound(f2, e2, b2);
        double diff = std::fabs(f2 - f1);
        bool sure = diff > m && std::fabs(diff - eps) > m;
        if (sure) return GWM_OK;
        sticky = true;
        int rc;
        if (!e1) {
            e1 = true;
            if ((rc = eval(x1, e1, &f1, &b1))) return rc;
            record(x1, f1);
        }
        if (!e2) {
            e2 = true;
            if ((rc = eval(x2, e2, &f2, &b2))) return rc;
            record(x2, f2);
        }
        return GWM_OK;
    };
    int rc;
    if ((rc = eval(x1, e1, &f1, &b1))) return rc;
    if (x2 == x1) {
        f2 = f1;
        e2 = e1;
        b2 = b1;
        if (report) report(x2, f2);
    } else if ((rc = eval(x2, e2, &f2, &b2))) {
        return rc;
    }
    if (f1 == DBL_MAX && f2 == DBL_MAX) return GWM_ERR_INVALID_INITIAL;
    record(x1, f1);
    record(x2, f2);
    if ((rc = certify_pair())) return rc;
    double d1 = f2 - f1;
    double xopt = f1 < f2 ? x1 : x2;
    const double ea = 100;
    while ((std::fabs(d) > eps) && (std::fabs(d1) > eps) && iter < ea) {
        d = R * d;
        if (f1 < f2) {
            xL = x2;
            x2 = x1;
            x1 = adaptBw ? std::round(xL + d) : (xL + d);
            f2 = f1;
            e2 = e1;
            b2 = b1;
            if (x1 == x2) {   // the same candidate: its evaluation, not a second one
                f1 = f2;
                e1 = e2;
                b1 = b2;
                if (report) report(x1, f1);
            } else {
                e1 = !certify || sticky;
                if ((rc = eval(x1, e1, &f1, &b1))) return rc;
                record(x1, f1);
            }
        } else {
            xU = x1;
            x1 = x2;
            x2 = adaptBw ? std::floor(xU - d) : (xU - d);
            f1 = f2;
            e1 = e2;
            b1 = b2;
            if (x2 == x1) {
                f2 = f1;
                e2 = e1;
                b2 = b1;
                if (report) report(x2, f2);
            } else {
                e2 = !certify || sticky;
                if ((rc = eval(x2, e2, &f2, &b2))) return rc;
                record(x2, f2);
            }
        }
        iter = iter + 1;
        if ((rc = certify_pair())) return rc;
        xopt = (f1 < f2) ? x1 : x2;
        d1 = f2 - f1;
    }
    *xopt_out = xopt;
    return GWM_OK;
}

}  // namespace detail
}  // namespace gwm


namespace gwm {
namespace detail {

void write_trace(const std::map<double, double>& log, double* tx, double* tf, int cap, int* len) {
    int m = 0;
    for (const auto& kv : log) {
        if (m < cap) {
            if (tx) tx[m] = kv.first;
            if (tf) tf[m] = kv.second;
        }
        ++m;
    }
    if (len) *len = m;
}

}  // namespace detail
}  // namespace gwm


namespace gwm {
namespace detail {

int max_distance_impl(gwm_hip_ctx* c, double* out) {
    // depends on the coordinates only: computed once per problem
    if (c->max_dist >= 0.0) {
        *out = c->max_dist;
        return GWM_OK;
    }
    int nblocks = (c->n + 255) / 256;
    GWM_HIP_CHECK(c, launch_max_distance(c->dist, c->points(), c->blockmax, nblocks, c->stream));
    std::vector<double> bm(nblocks);
    GWM_HIP_CHECK(c, hipMemcpyAsync(bm.data(), c->blockmax, nblocks * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    double m = 0.0;
    for (double v : bm) m = v > m ? v : m;
    c->max_dist = m;
    *out = m;
    return GWM_OK;
}

// Nearest other point of every point: the adaptive threshold at 2
// neighbours, the point itself being the first (BandwidthWeight.cpp:40-55
// with bw = 2: the second smallest distance, 0 for coincident points).
int ensure_dnn(gwm_hip_ctx* c) {
    if (c->dnn) return GWM_OK;
    double* buf = nullptr;
    if (c->async_mem) GWM_HIP_CHECK(c, hipMallocAsync((void**)&buf, (size_t)c->n * sizeof(double), c->stream));
    else GWM_HIP_CHECK(c, hipMalloc((void**)&buf, (size_t)c->n * sizeof(double)));
    c->dnn = buf;
    hipError_t e = launch_adaptive_bw(c->dist, c->points(), 2.0, buf, c->prune ? c->boxes : nullptr, c->sel_left,
                                      c->sel_nleft, c->stream);
    if (e != hipSuccess) {
        c->err = std::string("nearest-point select: ") + hipGetErrorString(e);
        if (c->async_mem) hipFreeAsync(buf, c->stream);
        else hipFree(buf);
        c->dnn = nullptr;
        return GWM_ERR_HIP;
    }
    return GWM_OK;
}

}  // namespace detail
}  // namespace gwm

namespace {

// Point tables of distance d and its 64-point chunk boxes (into *boxes):
// great circle needs the half unit vectors and sin(lat) of every point (built
// once) and boxes of the half unit vectors; Euclidean boxes of x / y.
int build_tables(gwm_hip_ctx* c, int d, double** boxes) {
    const size_t N = c->n;
    std::vector<double> cs(2 * N);
    to_sorted(c, c->hcoords.data(), cs.data(), 2);
    std::vector<double> box;
    int rc;
    if (d == GWM_GREAT_CIRCLE) {
        std::vector<double> h(4 * N);
        sphere_tables(N, cs.data(), cs.data() + N, h.data(), h.data() + N, h.data() + 2 * N, h.data() + 3 * N);
        if (!c->hx) {
            if ((rc = dalloc(c, &c->hx, N)) || (rc = dalloc(c, &c->hy_, N)) || (rc = dalloc(c, &c->hz, N)) ||
                (rc = dalloc(c, &c->sl, N)))
                return rc;
            GWM_HIP_CHECK(c, hipMemcpyAsync(c->hx, h.data(), N * sizeof(double), hipMemcpyHostToDevice, c->stream));
            GWM_HIP_CHECK(c, hipMemcpyAsync(c->hy_, h.data() + N, N * sizeof(double), hipMemcpyHostToDevice,
                                            c->stream));
            GWM_HIP_CHECK(c, hipMemcpyAsync(c->hz, h.data() + 2 * N, N * sizeof(double), hipMemcpyHostToDevice,
                                            c->stream));
            GWM_HIP_CHECK(c, hipMemcpyAsync(c->sl, h.data() + 3 * N, N * sizeof(double), hipMemcpyHostToDevice,
                                            c->stream));
        }
        const double* xs[3] = {h.data(), h.data() + N, h.data() + 2 * N};
        box = chunk_boxes(c->n, 3, xs);
    } else {
        const double* xs[2] = {cs.data(), cs.data() + N};
        box = chunk_boxes(c->n, 2, xs);
    }
    if ((rc = dalloc(c, boxes, box.size()))) return rc;
    GWM_HIP_CHECK(c, hipMemcpyAsync(*boxes, box.data(), box.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GWM_OK;
}

}  // namespace

namespace gwm {
namespace detail {

int use_distance(gwm_hip_ctx* c, int d) {
    if (d == c->dist) return GWM_OK;
    if (d != GWM_EUCLIDEAN && d != GWM_GREAT_CIRCLE) {
        c->err = "unknown distance";
        return GWM_ERR_INVALID;
    }
    gwm_hip_ctx::DistTables& cur = c->parked[c->dist];
    cur.boxes = c->boxes;
    cur.max_dist = c->max_dist;
    cur.dnn = c->dnn;
    cur.built = true;
    gwm_hip_ctx::DistTables next = c->parked[d];
    if (!next.built) {
        next = gwm_hip_ctx::DistTables{};
        int rc = build_tables(c, d, &next.boxes);
        if (rc) return rc;
    }
    c->parked[d] = gwm_hip_ctx::DistTables{};   // the active tables live in c
    c->boxes = next.boxes;
    c->max_dist = next.max_dist;
    c->dnn = next.dnn;
    c->dist = d;
    return GWM_OK;
}

void tele_print(gwm_hip_ctx* c, int level, const std::string& message) {
    if (c->print) c->print(c->tele_user, level, message.c_str());
}

void tele_progress(gwm_hip_ctx* c, double current, double total) {
    if (c->progress) c->progress(c->tele_user, current, total);
}

// std::to_string formats (int for adaptive bandwidths, "%f" otherwise), as
// IBandwidthSelectable::infoBandwidthCriterion (IBandwidthSelectable.h:30-49).
std::string info_bandwidth_criterion(bool adaptive, double bw, double value) {
    return std::string("#bandwidth-criterion ") + (adaptive ? std::to_string((int)bw) : std::to_string(bw)) + "," +
           std::to_string(value);
}

std::string info_bandwidth_criterion(bool adaptive) {
    return std::string("#bandwidth-criterion ") + (adaptive ? "adaptive" : "fixed") + ",criterion";
}

}  // namespace detail
}  // namespace gwm

namespace {

// Upload the design (caller's order) in sorted order with its product table
// Z = [x_a x_b (a <= b) | x_a y].
int upload_design(gwm_hip_ctx* c, const double* X_caller, const double* y_caller) {
    const int n = c->n, k = c->k, ldz = c->nta * 16;
    std::vector<double> Xs((size_t)n * k), ys(n);
    to_sorted(c, X_caller, Xs.data(), k);
    to_sorted(c, y_caller, ys.data(), 1);
    const double* X = Xs.data();
    const double* y = ys.data();
    std::vector<double> Z((size_t)c->rows_pad * ldz, 0.0);
    for (int j = 0; j < n; ++j) {
        double* z = Z.data() + (size_t)j * ldz;
        int p = 0;
        for (int a = 0; a < k; ++a)
            for (int b = a; b < k; ++b) z[p++] = X[j + (size_t)a * n] * X[j + (size_t)b * n];
        for (int a = 0; a < k; ++a) z[p++] = X[j + (size_t)a * n] * y[j];
    }
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->X, X, (size_t)n * k * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->y, y, (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->Z, Z.data(), Z.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, launch_split_fp16(c->Z, c->rows_pad, ldz, c->colmax, c->Zh, c->Zl, c->zscale, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GWM_OK;
}

// Single-column criterion / fit (GWRMultiscale.cpp:420-482, 560-636).
// The column x and response y of a single-column step on the device, in
// sorted order, with their product table (FP64 and, for mixed precision, its
// FP16 planes).  A golden search over one variable prepares them once.
// The precision a single-column sweep of this kernel runs in.  Bisquare and
// boxcar sweeps run in FP64 in both precisions: without a square root or an
// exponential per pair the FP64 VALU sweep (contract_var_f64) streams pairs
// as fast as the FP16-split MFMA sweep, whose one useful column leaves the
// matrix cores idle, and its criteria need no certification.
static int var_precision(int kernel, int precision) {
    return (kernel == GWM_BISQUARE || kernel == GWM_BOXCAR) ? GWM_FP64 : precision;
}

int var_prepare(gwm_hip_ctx* c, const double* x_caller, const double* y_caller, int precision) {
    const int n = c->n;
    std::vector<double> xs(n), ys(n);
    to_sorted(c, x_caller, xs.data(), 1);
    to_sorted(c, y_caller, ys.data(), 1);
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->vx, xs.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->vy, ys.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, launch_var_table(c->vx, c->vy, n, c->rows_pad, c->vZ, c->stream));
    if (precision == GWM_MIXED)
        GWM_HIP_CHECK(c, launch_split_fp16(c->vZ, c->rows_pad, 16, c->colmax, c->vZh, c->vZl, c->vscale, c->stream));
    // the host copies go out of scope: finish the uploads first
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return GWM_OK;
}

// One single-column criterion (value) or fit (beta) on the column prepared
// by var_prepare (x_caller / y_caller non-null: prepare them first).
int var_impl(gwm_hip_ctx* c, const double* x_caller, const double* y_caller, int kernel, int adaptive, double bw,
             int crit, int precision, double* value, double* beta, double* bound = nullptr) {
    if (bound) *bound = 0.0;
    int rc = check_kernel(c, kernel, precision);
    if (rc) return rc;
    if (value) *value = DBL_MAX;
    if (stopped(c)) return GWM_ERR_CANCELLED;
    const int n = c->n;
    const bool promoted = precision == GWM_MIXED && var_precision(kernel, precision) == GWM_FP64;
    precision = var_precision(kernel, precision);
    if (x_caller && (rc = var_prepare(c, x_caller, y_caller, precision))) return rc;
    const double* bw_pt;
    if ((rc = resolve_bandwidth(c, adaptive, bw, &bw_pt))) return rc;
    const bool cv = value && crit == GWM_CRIT_CV;
    const double* dnn = nullptr;
    if (cv) {
        if ((rc = ensure_dnn(c))) return rc;
        dnn = c->dnn;
    }
    int PT, nseg, fail = 0;
    const int smode = beta ? kSolveBeta : (cv ? kSolveCV : kSolveAIC);
    // a mixed CV evaluation with a point far from every other (its largest
    // weight below kMixedMinWeight: fail bit 2) is taken again in FP64
    // one synchronisation per evaluation: the reduction is queued behind the
    // solve and read back with the failure flags
    double red[3] = {0.0, 0.0, 0.0};
    for (int pass = 0; pass < 2; ++pass) {
        GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
        if ((rc = sweep(c, cv ? kModeCV : kModePlain, 1, 0, var_planes(c), bw_pt, bw, kernel, precision, &PT, &nseg,
                        dnn)))
            return rc;
        GWM_HIP_CHECK(c, launch_solve_var(smode, c->part, nseg, PT, 0, n, c->n_pad, c->vx, c->vy, bw_pt, bw, kernel,
                                          c->res2, c->sii, beta ? c->betas : nullptr, c->fail, c->stream,
                                          precision == GWM_MIXED ? dnn : nullptr));
        if (value) {
            GWM_HIP_CHECK(c, launch_reduce(c->res2, cv ? nullptr : c->sii, nullptr, nullptr, n, c->red, c->stream));
            GWM_HIP_CHECK(c, hipMemcpyAsync(red, c->red, sizeof(red), hipMemcpyDeviceToHost, c->stream));
            GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        } else if (precision == GWM_MIXED && cv) {
            GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        }
        if (precision != GWM_MIXED || !cv || !(fail & 2)) break;
        precision = GWM_FP64;
        c->solves += n;
    }
    if (value) {
        c->solves += n;
        if (fail & 1) return GWM_ERR_NOT_SPD;
        double v = cv ? red[0] : aicc(red[0], (double)n, red[1]);
        if (!std::isfinite(v)) return GWM_ERR_NOT_SPD;
        *value = v;
        // single column: no conditioning indicator and no (res yhat)^2 sums,
        // so the bound is tol_rel |f| and, for AICc, the trace term
        c->last_bound =
            precision == GWM_MIXED ? crit_error_bound(crit, v, red[0], red[1], 0.0, n, 1.0, 1.0, c->certify_rel) : 0.0;
        // an FP64 value of a mixed-precision request: reported by a negative bound
        if (bound) *bound = promoted ? -1.0 : c->last_bound;
        return GWM_OK;
    }
    GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((rc = download_caller(c, c->betas, beta, 1))) return rc;
    c->solves += n;
    return fail ? GWM_ERR_NOT_SPD : GWM_OK;
}

// Full-design fit returning betas only (fitAll without the hat matrix,
// GWRMultiscale.cpp:394-417).
int fit_betas_impl(gwm_hip_ctx* c, int kernel, int adaptive, double bw, int precision, double* betas,
                   double* ainv = nullptr) {
    int rc;
    const double* bw_pt;
    if ((rc = resolve_bandwidth(c, adaptive, bw, &bw_pt))) return rc;
    GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
    int PT, nseg;
    if ((rc = sweep(c, kModePlain, c->nta, 0, design_planes(c), bw_pt, bw, kernel, precision, &PT, &nseg))) return rc;
    SolveArgs s = solve_args(c, PT, nseg, 0, bw_pt, bw, kernel);
    s.ainv = ainv;
    GWM_HIP_CHECK(c, launch_solve(c->k, kSolveBeta, s, c->stream));
    if (precision == GWM_MIXED) {
        int count = 0;
        if ((rc = refine_fp64(c, kModePlain, kSolveBeta, 0, 0, bw_pt, bw, kernel, c->refine_fit, &count))) return rc;
    }
    int fail = 0;
    GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if ((rc = download_caller(c, c->betas, betas, c->k))) return rc;
    c->solves += c->n;
    return fail ? GWM_ERR_NOT_SPD : GWM_OK;
}

// Betas at prediction locations (GWRBasic::predictSerial, GWRBasic.cpp:167-191):
// beta(u) = (X^T W(u) X)^-1 X^T W(u) y with W(u) the weights of the distances
// from location u to every data point; adaptive bandwidths are the bw-th
// nearest data point of u (BandwidthWeight::weight on that distance vector).
// The locations are swept in their own spatial order; results come back in
// the caller's order.
int predict_impl(gwm_hip_ctx* c, int m, const double* loc, int kernel, int adaptive, double bw, int precision,
                 double* betas_out) {
    int rc;
    if ((rc = check_kernel(c, kernel, precision))) return rc;
    if (adaptive && !(bw >= 1.0)) {
        c->err = "adaptive bandwidth must be >= 1 neighbour";
        return GWM_ERR_INVALID;
    }
    if (!adaptive && !(bw > 0.0)) {
        c->err = "fixed bandwidth must be > 0";
        return GWM_ERR_INVALID;
    }
    if (stopped(c)) return GWM_ERR_CANCELLED;
    const size_t M = m;
    const int k = c->k;
    const bool gc = c->dist == GWM_GREAT_CIRCLE;
    // spatial order of the locations and their sorted tables
    const std::vector<int> qperm = spatial_order(m, loc, loc + M);
    const int ntab = gc ? 6 : 2;
    std::vector<double> tab(ntab * M);
    for (size_t s = 0; s < M; ++s) {
        tab[s] = loc[qperm[s]];
        tab[M + s] = loc[M + qperm[s]];
    }
    if (gc) sphere_tables(M, tab.data(), tab.data() + M, &tab[2 * M], &tab[3 * M], &tab[4 * M], &tab[5 * M]);
    // device arrays: tables, bandwidths, betas, indicator; lists
    DevBuf dbuf, ibuf;
    const size_t nd = M * (ntab + 3 + k);
    GWM_HIP_CHECK(c, hipMalloc(&dbuf.p, nd * sizeof(double)));
    GWM_HIP_CHECK(c, hipMalloc(&ibuf.p, (2 * M + 1) * sizeof(int)));
    double* d = static_cast<double*>(dbuf.p);
    int* left = static_cast<int*>(ibuf.p);
    int* rlist = left + M;
    unsigned int* nleft = reinterpret_cast<unsigned int*>(left + 2 * M);
    GWM_HIP_CHECK(c, hipMemcpyAsync(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    gwm::Points Q{};
    Q.px = d;
    Q.py = d + M;
    if (gc) {
        Q.hx = d + 2 * M;
        Q.hy = d + 3 * M;
        Q.hz = d + 4 * M;
        Q.sl = d + 5 * M;
    }
    Q.n = m;
    double* qbw = d + ntab * M;
    double* qcond = qbw + M;
    double* qdnn = qcond + M;
    double* qbetas = qdnn + M;
    const double* bw_pt = nullptr;
    if (adaptive) {
        timer_begin(c);
        GWM_HIP_CHECK(c, launch_adaptive_bw(c->dist, c->points(), bw, qbw, c->prune ? c->boxes : nullptr, left, nleft,
                                            c->stream, &Q));
        if ((rc = timer_end(c, "select"))) return rc;
        bw_pt = qbw;
    }
    // A location is not a data point, so its largest weight is that of its
    // nearest data point (the adaptive threshold of 1 neighbour): the FP64
    // exponent cut is relative to it, and mixed precision recomputes
    // locations whose largest weight is tiny.
    const bool mixed = precision == GWM_MIXED;
    const double* dnn = nullptr;
    if (mixed || kernel == GWM_GAUSSIAN || kernel == GWM_EXPONENTIAL) {
        GWM_HIP_CHECK(c, launch_adaptive_bw(c->dist, c->points(), 1.0, qdnn, c->prune ? c->boxes : nullptr, left,
                                            nleft, c->stream, &Q));
        dnn = qdnn;
    }
    // one sweep over the locations
    int nseg, seg_len;
    data_segments(c->n, nseg, seg_len);
    const int PT = c->nta * 16;
    const int mpad = (m + 127) / 128 * 128;
    if ((rc = ensure_part(c, (size_t)nseg * PT * mpad * sizeof(double)))) return rc;
    SweepLaunch L;
    SweepJob& j = L.job0;
    j.P = c->points();
    j.Q = Q;
    j.Z = c->Z;
    j.Zh = c->Zh;
    j.Zl = c->Zl;
    j.colscale = c->zscale;
    j.bw_pt = bw_pt;
    j.bw = bw;
    j.kernel = kernel;
    j.n = c->n;
    j.n_pad = mpad;
    j.seg_len = seg_len;
    j.nseg = nseg;
    j.nta = c->nta;
    j.ntb = 0;
    j.out = c->part;
    j.dnn = dnn;
    if (c->prune) j.boxes = c->boxes;
    L.jobs = nullptr;
    L.items = nullptr;
    L.nitems = 0;
    GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
    // large pruned sweeps: items longest first, as for the fits (the
    // locations are in their own spatial order)
    {
        const int r = contract_r(c->nta, 0, mixed ? 1 : 0, kModePlain, c->dist);
        const int tiles = (m + 64 * r - 1) / (64 * r);
        if (c->prune && c->autotune && tiles * nseg >= kTuneGridMax && nseg <= kPlanMaxSeg) {
            if ((rc = ensure_plan(c, (size_t)tiles * nseg))) return rc;
            GWM_HIP_CHECK(c, launch_sweep_plan(c->dist, Q, c->n, c->boxes, bw_pt, bw, nullptr, kernel, mixed ? 1 : 0,
                                               64 * r, seg_len, nseg, tiles, c->plan_counts, c->plan_order,
                                               c->stream));
            L.order = c->plan_order;
            L.norder = tiles * nseg;
        }
    }
    timer_begin(c);
    if (mixed) GWM_HIP_CHECK(c, launch_contract_mixed(c->dist, kModePlain, c->nta, 0, L, c->stream));
    else GWM_HIP_CHECK(c, launch_contract_f64(c->dist, kModePlain, c->nta, 0, L, c->stream));
    if ((rc = timer_end(c, mixed ? "contract_mixed_predict" : "contract_fp64_predict"))) return rc;
    SolveArgs sa = solve_args(c, PT, nseg, 0, bw_pt, bw, kernel);
    sa.n = m;
    sa.n_pad = mpad;
    sa.X = nullptr;
    sa.y = nullptr;
    sa.betas = qbetas;
    sa.cond = qcond;
    if (mixed) sa.dnn = dnn;
    GWM_HIP_CHECK(c, launch_solve(k, kSolvePredict, sa, c->stream));
    if (mixed) {
        const PredictSet q{Q, qcond, rlist, qbetas};
        int count = 0;
        if ((rc = refine_fp64(c, kModePlain, kSolvePredict, 0, 0, bw_pt, bw, kernel, c->refine_fit, &count, &q, dnn)))
            return rc;
    }
    int fail = 0;
    std::vector<double> b(M * k);
    GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(b.data(), qbetas, b.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (int p = 0; p < k; ++p)
        for (size_t s = 0; s < M; ++s) betas_out[(size_t)qperm[s] + p * M] = b[s + p * M];
    c->solves += m;
    if (fail) {
        c->err = "inv_sympd(): matrix is singular or not positive definite";
        return GWM_ERR_NOT_SPD;
    }
    return GWM_OK;
}

// Hat matrix of a multiscale run (GWRMultiscale.cpp:163-168, 176-179,
// 215-225, 296-301), resident on the device: R (k slices) and S0 from fitAll,
// then one update per backfitting step.  Memory: (k + 3) n^2 doubles.
struct MgwrHat {
    DevBuf mats, ainv, scratch;
    double *R = nullptr, *S0 = nullptr, *M = nullptr, *S = nullptr, *dinv = nullptr, *part = nullptr;
    size_t nn = 0;
};

int hat_alloc(gwm_hip_ctx* c, MgwrHat& h) {
    const size_t n = c->n, k = c->k;
    h.nn = n * n;
    const size_t bytes = (k + 3) * h.nn * sizeof(double);
    size_t free_b = 0, total_b = 0;
    GWM_HIP_CHECK(c, hipMemGetInfo(&free_b, &total_b));
    if (bytes > free_b / 10 * 9) {
        c->err = "the multiscale hat matrix needs (k + 3) n^2 doubles (" + std::to_string(bytes >> 20) +
                 " MiB), more than the device has free; run without it";
        return GWM_ERR_OOM;
    }
    GWM_HIP_CHECK(c, hipMalloc(&h.mats.p, bytes));
    GWM_HIP_CHECK(c, hipMalloc(&h.ainv.p, n * k * k * sizeof(double)));
    GWM_HIP_CHECK(c, hipMalloc(&h.scratch.p, (n + 2048 + 2) * sizeof(double)));
    h.R = static_cast<double*>(h.mats.p);
    h.S0 = h.R + k * h.nn;
    h.M = h.S0 + h.nn;
    h.S = h.M + h.nn;
    h.dinv = static_cast<double*>(h.scratch.p);
    h.part = h.dinv + n;
    return GWM_OK;
}

// Backfitting step of variable i at bandwidth bw: R_i <- S_i (R_i - S0) + S_i,
// S0 <- R_i(new) - (R_i(old) - S0).
int hat_step(gwm_hip_ctx* c, MgwrHat& h, int i, int kernel, int adaptive, double bw) {
    const int n = c->n;
    const double* bw_pt;
    int rc;
    if ((rc = resolve_bandwidth(c, adaptive, bw, &bw_pt))) return rc;
    double* Ri = h.R + (size_t)i * h.nn;
    GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_var_smatrix(c->dist, c->points(), c->X + (size_t)i * n, bw_pt, bw, kernel, h.dinv, c->fail,
                                        h.S, c->stream));
    GWM_HIP_CHECK(c, launch_sub(Ri, h.S0, h.M, h.nn, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(Ri, h.S, h.nn * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    GWM_HIP_CHECK(c, launch_dgemm_acc(n, h.S, h.M, Ri, c->stream));   // R_i = S_i M + S_i
    GWM_HIP_CHECK(c, launch_sub(Ri, h.M, h.S0, h.nn, c->stream));
    if ((rc = timer_end(c, "mgwr_hat"))) return rc;
    int fail = 0;
    GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (fail) {
        c->err = "inv_sympd(): matrix is singular or not positive definite";
        return GWM_ERR_NOT_SPD;
    }
    return GWM_OK;
}

// Criteria of candidate column sets (GWRBasic::indepVarsSelectionCriterion,
// GWRBasic.cpp:309-346): columns {0 if has_intercept} + vars; RSS and tr(S)
// on the device from the Gram matrix G of [X y] (launch_gram), AICc here;
// DBL_MAX when the design is not SPD or the value is not finite
// (GWRBasic.cpp:893).
int varsel_criteria(gwm_hip_ctx* c, const double* G, int has_intercept, const std::vector<std::vector<int>>& sets,
                    std::vector<double>& crit) {
    const int k = c->k, ns = (int)sets.size();
    crit.assign(ns, DBL_MAX);
    if (ns == 0) return GWM_OK;
    std::vector<int> h(ns * (k + 1), 0);
    for (int q = 0; q < ns; ++q) {
        int p = 0;
        if (has_intercept) h[q * k + p++] = 0;
        for (int v : sets[q]) h[q * k + p++] = v;
        h[ns * k + q] = p;
    }
    DevBuf ib, ob;
    GWM_HIP_CHECK(c, hipMalloc(&ib.p, h.size() * sizeof(int)));
    const size_t nblk = varsel_blocks(c->n);
    GWM_HIP_CHECK(c, hipMalloc(&ob.p, 2 * (size_t)ns * (1 + nblk) * sizeof(double)));
    int* dsets = static_cast<int*>(ib.p);
    double* dout = static_cast<double*>(ob.p);
    double* dpart = dout + 2 * (size_t)ns;
    GWM_HIP_CHECK(c, hipMemcpyAsync(dsets, h.data(), h.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_varsel(c->X, c->y, c->n, k, G, dsets, dsets + ns * k, ns, dpart, dout, c->stream));
    int rc = timer_end(c, "varsel");
    if (rc) return rc;
    std::vector<double> o(2 * ns);
    GWM_HIP_CHECK(c, hipMemcpyAsync(o.data(), dout, o.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    for (int q = 0; q < ns; ++q) {
        const double v = aicc(o[2 * q], (double)c->n, o[2 * q + 1]);
        crit[q] = std::isfinite(v) ? v : DBL_MAX;
    }
    c->solves += ns;
    return GWM_OK;
}

int gram(gwm_hip_ctx* c, DevBuf& g) {
    const int k = c->k;
    GWM_HIP_CHECK(c, hipMalloc(&g.p, (size_t)(k + 1) * (k + 2) / 2 * sizeof(double)));
    GWM_HIP_CHECK(c, launch_gram(c->X, c->y, c->n, k, static_cast<double*>(g.p), c->stream));
    return GWM_OK;
}

// Criterion evaluations reported as the reference's criterion functions do
// (GWRBasic.cpp:477-484, GWRMultiscale.cpp:588-594): a finite value gives the
// "#bandwidth-criterion <bw>,<value>" message and the progress fraction
// exp(-|last - value|); a failed evaluation (a local design that is not SPD)
// an error message.  golden() reports every evaluation, and a candidate
// that coincides with the previous one (rounded adaptive candidates) once
// more, as the reference evaluates it again.
EvalReport telegram_report(gwm_hip_ctx* c, bool adaptive) {
    if (!c->print && !c->progress) return EvalReport();
    return [c, adaptive](double x, double fx) {
        if (fx < DBL_MAX) {
            tele_print(c, kLogInfo, info_bandwidth_criterion(adaptive, x, fx));
            tele_progress(c, std::exp(-std::fabs(c->last_crit - fx)), 0.0);
            c->last_crit = fx;
        } else {
            tele_print(c, kLogErr, "inv_sympd(): matrix is singular or not positive definite");
        }
    };
}

int check_vars(gwm_hip_ctx* c, int has_intercept, const int* vars, int nvars) {
    const int first = has_intercept ? 1 : 0;
    if (nvars < 0 || (nvars > 0 && !vars) || nvars + first < 1 || nvars + first > c->k) {
        c->err = "variables: need 1 <= columns <= k";
        return GWM_ERR_INVALID;
    }
    for (int i = 0; i < nvars; ++i)
        if (vars[i] < first || vars[i] >= c->k) {
            c->err = "variables: column index out of range";
            return GWM_ERR_INVALID;
        }
    return GWM_OK;
}

}  // namespace

extern "C" {

int gwm_hip_device_count(int* count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (count) *count = e == hipSuccess ? c : 0;
    return e == hipSuccess ? GWM_OK : GWM_ERR_NO_DEVICE;
}

int gwm_hip_create(int device, gwm_hip_ctx** out) {
    if (!out) return GWM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0 || device < 0 || device >= count)
        return GWM_ERR_NO_DEVICE;
    gwm_hip_ctx* c = new gwm_hip_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        noalloc_init(device) != GWM_OK) {
        gwm_hip_destroy(c);
        return GWM_ERR_HIP;
    }
    *out = c;
    return GWM_OK;
}

void gwm_hip_destroy(gwm_hip_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    free_problem(c);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->tev0) hipEventDestroy(c->tev0);
    if (c->tev1) hipEventDestroy(c->tev1);
    if (c->pairs_dev) hipFree(c->pairs_dev);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* gwm_hip_last_error(const gwm_hip_ctx* c) { return c ? c->err.c_str() : "null context"; }

int gwm_hip_set_stop_callback(gwm_hip_ctx* c, int (*stop)(void*), void* user) {
    if (!c) return GWM_ERR_INVALID;
    c->stop = stop;
    c->stop_user = user;
    return GWM_OK;
}

int gwm_hip_set_telegram(gwm_hip_ctx* c, void (*print)(void*, int, const char*),
                         void (*progress)(void*, double, double), void* user) {
    if (!c) return GWM_ERR_INVALID;
    c->print = print;
    c->progress = progress;
    c->tele_user = user;
    c->last_crit = DBL_MAX;
    return GWM_OK;
}

int gwm_hip_set_data(gwm_hip_ctx* c, int n, int k, const double* coords, const double* X, const double* y,
                     int distance) {
    if (!c) return GWM_ERR_INVALID;
    if (n < 2 || k < 1 || k > 16 || !coords || !X || !y || (distance != GWM_EUCLIDEAN && distance != GWM_GREAT_CIRCLE)) {
        c->err = "invalid data: need n >= 2, 1 <= k <= 16, coords/X/y, a known distance";
        return GWM_ERR_INVALID;
    }
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    free_problem(c);
    c->n = n;
    c->k = k;
    c->dist = distance;
    c->nta = nta_of(k);
    c->ntb = ntb_of(k);
    c->m = k * (k + 1) / 2;
    c->n_pad = (n + 127) / 128 * 128;
    c->rows_pad = (n + 63) / 64 * 64;
    c->hcoords.assign(coords, coords + 2 * (size_t)n);
    c->hX.assign(X, X + (size_t)n * k);
    c->hy.assign(y, y + n);
    int rc;
    const size_t N = n;
    if ((rc = dalloc(c, &c->px, N)) || (rc = dalloc(c, &c->py, N)) || (rc = dalloc(c, &c->X, N * k)) ||
        (rc = dalloc(c, &c->y, N)) || (rc = dalloc(c, &c->Z, (size_t)c->rows_pad * c->nta * 16)) ||
        (rc = dalloc(c, &c->G, (size_t)c->rows_pad * c->ntb * 16)) || (rc = dalloc(c, &c->bwi, N)) ||
        (rc = dalloc(c, &c->res2, N)) || (rc = dalloc(c, &c->cond, N)) ||
        (rc = dalloc(c, &c->rlist, N)) || (rc = dalloc(c, &c->rcount, 1)) || (rc = dalloc(c, &c->sel_left, N)) ||
        (rc = dalloc(c, &c->sel_nleft, 1)) || (rc = dalloc(c, &c->sii, N)) || (rc = dalloc(c, &c->ssi, N)) ||
        (rc = dalloc(c, &c->betas, N * k)) || (rc = dalloc(c, &c->se, N * k)) || (rc = dalloc(c, &c->gvec, N * k)) ||
        (rc = dalloc(c, &c->qd, N)) || (rc = dalloc(c, &c->red, 8)) || (rc = dalloc(c, &c->blockmax, (N + 255) / 256)) ||
        (rc = dalloc(c, &c->vx, N)) || (rc = dalloc(c, &c->vy, N)) || (rc = dalloc(c, &c->vZ, (size_t)c->rows_pad * 16)) ||
        (rc = dalloc(c, &c->fail, 1)) || (rc = dalloc(c, &c->Zh, (size_t)c->rows_pad * c->nta * 16)) ||
        (rc = dalloc(c, &c->Zl, (size_t)c->rows_pad * c->nta * 16)) ||
        (rc = dalloc(c, &c->Gh, (size_t)c->rows_pad * c->ntb * 16)) ||
        (rc = dalloc(c, &c->Gl, (size_t)c->rows_pad * c->ntb * 16)) ||
        (rc = dalloc(c, &c->vZh, (size_t)c->rows_pad * 16)) || (rc = dalloc(c, &c->vZl, (size_t)c->rows_pad * 16)) ||
        (rc = dalloc(c, &c->zscale, (size_t)c->nta * 16)) || (rc = dalloc(c, &c->gscale, (size_t)c->ntb * 16)) ||
        (rc = dalloc(c, &c->vscale, 16)) || (rc = dalloc(c, &c->colmax, (size_t)c->nta * 16))) {
        free_problem(c);
        return rc;
    }
    GWM_HIP_CHECK(c, hipMemsetAsync(c->G, 0, (size_t)c->rows_pad * c->ntb * 16 * sizeof(double), c->stream));
    // spatial order, sorted coordinates, then the distance's point tables and
    // the 64-point chunk boxes
    c->perm = spatial_order(n, coords, coords + N);
    std::vector<double> cs(2 * N);
    to_sorted(c, coords, cs.data(), 2);
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->px, cs.data(), N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(c->py, cs.data() + N, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if ((rc = dalloc(c, &c->chunk_bw, (N + 63) / 64)) || (rc = build_tables(c, distance, &c->boxes))) {
        free_problem(c);
        return rc;
    }
    c->dist = c->dist_data = distance;
    return upload_design(c, X, y);
}

int gwm_hip_criterion(gwm_hip_ctx* c, int kernel, int adaptive, double bw, int crit, int precision, double* value) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!value) return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    return criterion_impl(c, kernel, adaptive, bw, crit, precision, value);
}

int gwm_hip_fit(gwm_hip_ctx* c, int kernel, int adaptive, double bw, int precision, double* betas, double* se_raw,
                double* shat, double* qDiag, double* S) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_kernel(c, kernel, precision))) return rc;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    if (stopped(c)) return GWM_ERR_CANCELLED;
    const double* bw_pt;
    if ((rc = resolve_bandwidth(c, adaptive, bw, &bw_pt))) return rc;
    GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
    int PT, nseg;
    if ((rc = sweep(c, kModeHat, c->nta, c->ntb, design_planes(c), bw_pt, bw, kernel, precision, &PT, &nseg)))
        return rc;
    SolveArgs s = solve_args(c, PT, nseg, c->nta * 16, bw_pt, bw, kernel);
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_solve(c->k, kSolveFit, s, c->stream));
    if ((rc = timer_end(c, "solve"))) return rc;
    if (precision == GWM_MIXED) {
        int count = 0;
        if ((rc = refine_fp64(c, kModeHat, kSolveFit, c->ntb, c->nta * 16, bw_pt, bw, kernel, c->refine_fit, &count)))
            return rc;
    }
    int fail = 0;
    GWM_HIP_CHECK(c, hipMemcpyAsync(&fail, c->fail, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->solves += c->n;
    if (fail) {
        c->err = "inv_sympd(): matrix is singular or not positive definite";
        return GWM_ERR_NOT_SPD;
    }
    if (qDiag) {
        // second sweep: M_j = sum_i w_ij^2 vech(g_i g_i^T), bandwidth of i
        if (precision == GWM_MIXED)
            GWM_HIP_CHECK(c, launch_split_fp16(c->G, c->rows_pad, c->ntb * 16, c->colmax, c->Gh, c->Gl, c->gscale,
                                               c->stream));
        if ((rc = sweep(c, kModeQDiag, c->ntb, 0, g_planes(c), bw_pt, bw, kernel, precision, &PT, &nseg))) return rc;
        GWM_HIP_CHECK(c, launch_qdiag(c->k, c->part, nseg, PT, c->n, c->n_pad, c->X, c->sii, c->qd, c->stream));
        if ((rc = download_caller(c, c->qd, qDiag, 1))) return rc;
    }
    if (S) {
        if (c->n > 65535) {
            c->err = "the S matrix is only formed for n <= 65535";
            return GWM_ERR_INVALID;
        }
        if (!c->S) {
            int rc = dalloc(c, &c->S, (size_t)c->n * c->n);
            if (rc) return rc;
        }
        GWM_HIP_CHECK(c, launch_smatrix(c->dist, c->k, c->points(), c->X, c->gvec, bw_pt, bw, kernel, c->S, c->stream));
        std::vector<double> Ss((size_t)c->n * c->n);
        GWM_HIP_CHECK(c, hipMemcpyAsync(Ss.data(), c->S, Ss.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        const size_t nn = c->n;
        for (size_t j = 0; j < nn; ++j)
            for (size_t i = 0; i < nn; ++i) S[(size_t)c->perm[i] + (size_t)c->perm[j] * nn] = Ss[i + j * nn];
    }
    GWM_HIP_CHECK(c, launch_reduce(c->sii, c->ssi, c->res2, nullptr, c->n, c->red, c->stream));
    double red[3];
    GWM_HIP_CHECK(c, hipMemcpyAsync(red, c->red, sizeof(red), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (betas && (rc = download_caller(c, c->betas, betas, c->k))) return rc;
    if (se_raw && (rc = download_caller(c, c->se, se_raw, c->k))) return rc;
    if (shat) {
        shat[0] = red[0];
        shat[1] = red[1];
    }
    return GWM_OK;
}

int gwm_hip_select_bandwidth(gwm_hip_ctx* c, int kernel, int adaptive, int crit, int precision, int has_lower,
                             double lower, int has_upper, double upper, double* bw_out, double* tx, double* tf,
                             int cap, int* len) {
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_kernel(c, kernel, precision))) return rc;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    D2Scope scope(c);
    CertScope cert(c, precision == GWM_MIXED && !adaptive);
    if (len) *len = 0;
    double lo = has_lower ? lower : (adaptive ? 20 : 0.0);
    double hi;
    if (has_upper) hi = upper;
    else if (adaptive) hi = (double)c->n;
    else if ((rc = max_distance_impl(c, &hi))) return rc;
    std::map<double, double> log;
    double xopt = 0.0;
    rc = golden(adaptive != 0, lo, hi,
                [&](double x, bool exact, double* fx, double* fb) {
                    return criterion_impl(c, kernel, adaptive, x, crit, exact ? GWM_FP64 : precision, fx, fb);
                },
                precision == GWM_MIXED, c->certify_rel, &xopt, log, telegram_report(c, adaptive != 0),
                &c->cert_hi);
    write_trace(log, tx, tf, cap, len);
    if (rc) return rc;
    if (bw_out) *bw_out = xopt;
    return GWM_OK;
}

int gwm_hip_gwr_run(gwm_hip_ctx* c, int kernel, int adaptive, double bw, int select_bw, int crit, int precision,
                    int has_lower, double lower, int has_upper, double upper, double* betas, double* betasSE,
                    double* shat, double* qDiag, double* S, double* diag, double* bw_out, double* tx, double* tf,
                    int cap, int* len) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!betas || !betasSE || !shat || !diag) return GWM_ERR_INVALID;
    D2Scope scope(c);
    if (len) *len = 0;
    if (select_bw) {
        tele_print(c, kLogInfo, kStage + "Bandwidth selection");
        tele_print(c, kLogInfo, info_bandwidth_criterion(adaptive != 0));
        if ((rc = gwm_hip_select_bandwidth(c, kernel, adaptive, crit, precision, has_lower, lower, has_upper, upper,
                                           &bw, tx, tf, cap, len)))
            return rc;
    }
    if (bw_out) *bw_out = bw;
    tele_print(c, kLogInfo, kStage + "Model fitting");
    if ((rc = gwm_hip_fit(c, kernel, adaptive, bw, precision, betas, betasSE, shat, qDiag, S))) {
        if (rc == GWM_ERR_NOT_SPD) tele_print(c, kLogErr, "inv_sympd(): matrix is singular or not positive definite");
        return rc;
    }
    tele_progress(c, c->n, c->n);
    tele_print(c, kLogInfo, kStage + "Model Diagnostic");
    // CalcDiagnostic (GWRBasic.cpp:21-34) and SE scaling (:100-103)
    const int n_ = c->n, k = c->k;
    calc_diagnostic(n_, k, c->hX.data(), c->hy.data(), betas, shat, diag);
    const double n = (double)n_, rss = diag[0];
    double sigma = rss / (n - 2 * shat[0] + shat[1]);
    for (size_t i = 0; i < (size_t)n_ * k; ++i) betasSE[i] = std::sqrt(sigma * betasSE[i]);
    return GWM_OK;
}

int gwm_hip_max_distance(gwm_hip_ctx* c, double* out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    return max_distance_impl(c, out);
}

int gwm_hip_distance_row(gwm_hip_ctx* c, int focus, double* out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out || focus < 0 || focus >= c->n) return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    int pos = 0;   // sorted position of the caller's point `focus`
    while (c->perm[pos] != focus) ++pos;
    GWM_HIP_CHECK(c, launch_distance_row(c->dist, c->points(), pos, c->res2, c->stream));
    return download_caller(c, c->res2, out, 1);
}

int gwm_hip_sqrt_boundary(double bw, double* out) {
    if (!out || std::isnan(bw)) return GWM_ERR_INVALID;
    *out = sqrt_boundary(bw);
    return GWM_OK;
}

int gwm_hip_adaptive_bandwidths(gwm_hip_ctx* c, double bw, double* out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!out) return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    const double* bw_pt;
    if ((rc = resolve_bandwidth(c, 1, bw, &bw_pt))) return rc;
    return download_caller(c, bw_pt, out, 1);
}

int gwm_hip_var_criterion(gwm_hip_ctx* c, const double* x, const double* y, int kernel, int adaptive, double bw,
                          int crit, int precision, double* value) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!x || !y || !value) return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    return var_impl(c, x, y, kernel, adaptive, bw, crit, precision, value, nullptr);
}

int gwm_hip_var_fit(gwm_hip_ctx* c, const double* x, const double* y, int kernel, int adaptive, double bw,
                    int precision, double* beta) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!x || !y || !beta) return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    return var_impl(c, x, y, kernel, adaptive, bw, GWM_CRIT_AIC, precision, nullptr, beta);
}

int gwm_hip_mgwr_run(gwm_hip_ctx* c, const int* kernels, const int* adaptive, const double* bws, const int* init_in,
                     const int* approach, const int* centered, const double* bw_threshold, int has_intercept,
                     int crit_type, double crit_threshold, int max_iter, int retry_times, int adaptive_lower,
                     int has_lower, double lower, int has_upper, double upper, int precision, double* betas_out,
                     double* bws_out, double* diag, double* rss_out, int* iterations_out) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!kernels || !adaptive || !bws || !init_in || !approach || !centered || !bw_threshold || !betas_out ||
        !bws_out || !diag)
        return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    D2Scope scope(c);
    const int n = c->n, k = c->k;
    if (!c->mgwr_dists.empty() && (int)c->mgwr_dists.size() != k) {
        c->err = "mgwr: the per-variable distances were set for " + std::to_string(c->mgwr_dists.size()) +
                 " variables, the data has " + std::to_string(k);
        return GWM_ERR_INVALID;
    }
    // one distance per variable (GWRMultiscale.h:578, a SpatialWeight each);
    // the initial all-variable search and fit use variable 0's
    // (GWRMultiscale.cpp:1476-1483)
    auto dist_of = [&](int v) { return c->mgwr_dists.empty() ? c->dist_data : c->mgwr_dists[v]; };
    const std::vector<double> X0 = c->hX, y = c->hy;
    // GWRMultiscale.cpp:88-99: centre predictors
    std::vector<double> Xc = X0;
    for (int i = has_intercept ? 1 : 0; i < k; ++i)
        if (centered[i]) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += Xc[j + (size_t)i * n];
            double mu = s / n;
            for (int j = 0; j < n; ++j) Xc[j + (size_t)i * n] -= mu;
        }
    if ((rc = upload_design(c, Xc.data(), y.data()))) return rc;
    auto restore = [&]() {
        upload_design(c, X0.data(), y.data());
        use_distance(c, c->dist_data);
    };
    std::vector<int> init(init_in, init_in + k);
    std::vector<double> bw(bws, bws + k), maxD(k, 0.0);
    auto col = [&](int v) { return Xc.data() + (size_t)v * n; };
    auto fail_out = [&](int code) {
        restore();
        return code;
    };
    // a variable's bandwidth search, reported as GWRMultiscale.cpp:588-594
    // (on the column and response var_prepare has put on the device)
    auto search_var = [&](int i, double lo, double hi, double* out) {
        std::map<double, double> log;
        return golden(adaptive[i] != 0, lo, hi,
                      [&](double x, bool exact, double* fx, double* fb) {
                          return var_impl(c, nullptr, nullptr, kernels[i], adaptive[i], x, approach[i],
                                          exact ? GWM_FP64 : precision, fx, nullptr, fb);
                      },
                      precision == GWM_MIXED, c->certify_rel, out, log, telegram_report(c, adaptive[i] != 0));
    };
    // initial per-variable bandwidths (GWRMultiscale.cpp:104-135)
    tele_print(c, kLogInfo, kStage + "Calculating initial bandwidths");
    for (int i = 0; i < k; ++i) {
        if (stopped(c)) return fail_out(GWM_ERR_CANCELLED);
        if (init[i] != 0) continue;
        bool ad = adaptive[i] != 0;
        if ((rc = use_distance(c, dist_of(i))) || (rc = max_distance_impl(c, &maxD[i]))) return fail_out(rc);
        double lo = has_lower ? lower : (ad ? (double)adaptive_lower : maxD[i] / 5000.0);
        double hi = has_upper ? upper : (ad ? (double)n : maxD[i]);
        tele_print(c, kLogInfo, kInitialBw + std::to_string(i));
        if ((rc = var_prepare(c, col(i), y.data(), var_precision(kernels[i], precision))) ||
            (rc = search_var(i, lo, hi, &bw[i])))
            return fail_out(rc);
        tele_print(c, kLogInfo, kInitialBw + std::to_string(i) + "," + std::to_string(bw[i]));
    }
    // initial all-variable bandwidth (GWRMultiscale.cpp:137-160)
    tele_print(c, kLogInfo, kStage + "Calculating initial bandwidth");
    if ((rc = use_distance(c, dist_of(0)))) return fail_out(rc);
    double maxDist;
    if ((rc = max_distance_impl(c, &maxDist))) return fail_out(rc);
    bool ad0 = adaptive[0] != 0;
    double lo0 = has_lower ? lower : (ad0 ? (double)adaptive_lower : maxDist / 5000.0);
    double hi0 = has_upper ? upper : (ad0 ? (double)n : maxDist);
    std::map<double, double> log0;
    double initBw = 0.0;
    tele_print(c, kLogInfo, info_bandwidth_criterion(ad0));
    rc = golden(ad0, lo0, hi0,
                [&](double x, bool exact, double* fx, double* fb) {
                    return criterion_impl(c, kernels[0], adaptive[0], x, approach[0], exact ? GWM_FP64 : precision,
                                          fx, fb);
                },
                precision == GWM_MIXED, c->certify_rel, &initBw, log0, telegram_report(c, ad0));
    if (rc) return fail_out(rc);
    if (stopped(c)) return fail_out(GWM_ERR_CANCELLED);
    tele_print(c, kLogInfo, kStage + "Model fitting");
    tele_print(c, kLogInfo, kBackfit + "Model fitting with inital bandwidth");
    // backfitting (GWRMultiscale.cpp:208-317)
    std::vector<double> betas((size_t)n * k), resid(n), yi(n), colv(n);
    const bool hat = c->mgwr_hat;
    MgwrHat H;
    if (hat) {
        // fitAll in FP64 with the per-point A^-1 (GWRMultiscale.cpp:362-395), then R and S0 (:215-225)
        if ((rc = hat_alloc(c, H))) return fail_out(rc);
        if ((rc = fit_betas_impl(c, kernels[0], adaptive[0], initBw, GWM_FP64, betas.data(),
                                 static_cast<double*>(H.ainv.p))))
            return fail_out(rc);
        const double* bw_pt;
        if ((rc = resolve_bandwidth(c, adaptive[0], initBw, &bw_pt))) return fail_out(rc);
        GWM_HIP_CHECK(c, launch_hat_init(c->dist, c->points(), k, c->X, static_cast<double*>(H.ainv.p), bw_pt, initBw,
                                         kernels[0], H.R, H.S0, c->stream));
    } else if ((rc = fit_betas_impl(c, kernels[0], adaptive[0], initBw, precision, betas.data()))) {
        return fail_out(rc);
    }
    auto update_resid = [&]() {
        for (int j = 0; j < n; ++j) {
            double f = 0.0;
            for (int a = 0; a < k; ++a) f += betas[j + (size_t)a * n] * Xc[j + (size_t)a * n];
            resid[j] = y[j] - f;
        }
    };
    update_resid();
    double RSS0 = 0.0;
    for (int j = 0; j < n; ++j) RSS0 += resid[j] * resid[j];
    double RSS1 = DBL_MAX, crit = DBL_MAX;
    std::vector<int> bwChangeNo(k, 0);
    int iteration = 1;
    tele_print(c, kLogInfo, kBackfit + "Selecting the optimum bandwidths for each independent variable");
    for (; iteration <= max_iter && crit > crit_threshold; ++iteration) {
        if (stopped(c)) return fail_out(GWM_ERR_CANCELLED);
        tele_print(c, kLogInfo, kBackfit + "#iteration " + std::to_string(iteration));
        for (int i = 0; i < k; ++i) {
            const double* xi = col(i);
            for (int j = 0; j < n; ++j) yi[j] = resid[j] + betas[j + (size_t)i * n] * xi[j];
            if ((rc = use_distance(c, dist_of(i))) ||
                (rc = var_prepare(c, xi, yi.data(), var_precision(kernels[i], precision))))
                return fail_out(rc);
            if (init[i] != 2) {
                tele_print(c, kLogInfo, kBackfit + "#variable-bandwidth-selection " + std::to_string(i));
                bool ad = adaptive[i] != 0;
                // maxD[i] is set for variables searched at the start; a
                // variable given an initial bandwidth searches from 0
                // (mMaxDistances is zero-filled there, GWRMultiscale.cpp:74)
                double lo = has_lower ? lower : (ad ? (double)adaptive_lower : maxD[i] / 5000.0);
                double hi = has_upper ? upper : (ad ? (double)n : maxD[i]);
                double bw1 = 0.0;
                if ((rc = search_var(i, lo, hi, &bw1))) return fail_out(rc);
                // GWRMultiscale.cpp:263-290
                std::string vbs = std::to_string(i) + "," + std::to_string(bw[i]) + "," + std::to_string(bw1) + "," +
                                  std::to_string(std::fabs(bw1 - bw[i]));
                if (std::fabs(bw1 - bw[i]) > bw_threshold[i]) {
                    bwChangeNo[i] = 0;
                    vbs += ",false";
                } else {
                    bwChangeNo[i] += 1;
                    if (bwChangeNo[i] >= retry_times) {
                        init[i] = 2;
                        vbs += ",true";
                    } else {
                        vbs += "," + std::to_string(bwChangeNo[i]) + "," + std::to_string(retry_times - bwChangeNo[i]);
                    }
                }
                bw[i] = bw1;
                tele_print(c, kLogInfo, kBackfit + "#variable-bandwidth-selection " + vbs);
            }
            if ((rc = var_impl(c, nullptr, nullptr, kernels[i], adaptive[i], bw[i], GWM_CRIT_AIC, precision, nullptr,
                               colv.data())))
                return fail_out(rc);
            std::copy(colv.begin(), colv.end(), betas.begin() + (size_t)i * n);
            if (hat && (rc = hat_step(c, H, i, kernels[i], adaptive[i], bw[i]))) return fail_out(rc);
            update_resid();
        }
        RSS1 = 0.0;
        for (int j = 0; j < n; ++j) RSS1 += resid[j] * resid[j];
        crit = crit_type == 0 ? std::fabs(RSS1 - RSS0) : std::sqrt(std::fabs(RSS1 - RSS0) / RSS1);
        tele_print(c, kLogInfo, kBackfit + "#backfitting-criterion " + std::to_string(crit));
        tele_progress(c, std::exp(-std::fabs(crit - crit_threshold)), 0.0);
        RSS0 = RSS1;
    }
    tele_print(c, kLogInfo, kBackfit + "Finished");
    tele_print(c, kLogInfo, kStage + "Model Diagnostic");
    if (iterations_out) *iterations_out = iteration - 1;
    // CalcDiagnostic (GWRMultiscale.cpp:41-66) with shat = {tr S0, tr S0^T S0}
    // ({0, 0} without the hat matrix, :176-179)
    double shat[2] = {0.0, 0.0};
    if (hat) {
        GWM_HIP_CHECK(c, launch_trace(H.S0, n, H.part, H.part + 2048, c->stream));
        GWM_HIP_CHECK(c, hipMemcpyAsync(shat, H.part + 2048, sizeof(shat), hipMemcpyDeviceToHost, c->stream));
        GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    }
    c->mgwr_shat[0] = shat[0];
    c->mgwr_shat[1] = shat[1];
    const double nd = n;
    double ym = 0.0;
    for (int j = 0; j < n; ++j) ym += y[j];
    ym /= nd;
    double TSS = 0.0;
    for (int j = 0; j < n; ++j) TSS += (y[j] - ym) * (y[j] - ym);
    double r2 = 1 - RSS0 / TSS;
    const double trS = shat[0], trStS = shat[1];
    double edf = nd - 2 * trS + trStS;
    diag[0] = RSS0;
    diag[1] = 0.0;
    diag[2] = nd * std::log(RSS0 / nd) + nd * std::log(2 * kPi) + nd * ((nd + trS) / (nd - 2 - trS));
    diag[3] = 2 * trS - trStS;
    diag[4] = edf;
    diag[5] = r2;
    diag[6] = 1 - (1 - r2) * (nd - 1) / (edf - 1);
    if (rss_out) *rss_out = RSS0;
    std::copy(betas.begin(), betas.end(), betas_out);
    std::copy(bw.begin(), bw.end(), bws_out);
    restore();
    return GWM_OK;
}

int gwm_hip_predict(gwm_hip_ctx* c, int m, const double* locations, int kernel, int adaptive, double bw,
                    int precision, double* betas) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (m < 1 || !locations || !betas) {
        c->err = "predict: need m >= 1 locations and an output buffer";
        return GWM_ERR_INVALID;
    }
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    return predict_impl(c, m, locations, kernel, adaptive, bw, precision, betas);
}

int gwm_hip_indep_vars_criterion(gwm_hip_ctx* c, int has_intercept, const int* vars, int nvars, double* value) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!value) return GWM_ERR_INVALID;
    *value = DBL_MAX;
    if ((rc = check_vars(c, has_intercept, vars, nvars))) return rc;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    if (stopped(c)) return GWM_ERR_CANCELLED;
    DevBuf g;
    if ((rc = gram(c, g))) return rc;
    std::vector<double> crit;
    if ((rc = varsel_criteria(c, static_cast<double*>(g.p), has_intercept, {std::vector<int>(vars, vars + nvars)},
                              crit)))
        return rc;
    *value = crit[0];
    return GWM_OK;
}

// VariableForwardSelector::optimize / sort / select
// (src/gwmodelpp/VariableForwardSelector.cpp:9-77): each step adds the
// remaining column with the smallest criterion (first on ties, like
// arma::index_min); the model list is then ordered by size and, within a size,
// by descending criterion (a stable sort: the reference's std::sort leaves
// equal values in an unspecified order); the selection is the last model
// whose predecessor's criterion is larger by at least the threshold, else the
// last model (the reference's loop reads before the first element there).
int gwm_hip_indep_vars_select(gwm_hip_ctx* c, int has_intercept, double threshold, int* selected, int* nselected,
                              int* model_vars, int* model_nvars, double* model_crit, int cap_models, int* nmodels) {
    int rc = check_ready(c);
    if (rc) return rc;
    if (!selected || !nselected || !nmodels || (cap_models > 0 && (!model_vars || !model_nvars || !model_crit)))
        return GWM_ERR_INVALID;
    GWM_HIP_CHECK(c, hipSetDevice(c->device));
    const int k = c->k;
    std::vector<int> vars;
    for (int col = has_intercept ? 1 : 0; col < k; ++col) vars.push_back(col);
    const int m = (int)vars.size();
    std::vector<int> cur, rest(vars);
    std::vector<std::pair<std::vector<int>, double>> models;
    double progress_cur = 0.0;
    tele_print(c, kLogInfo, kStage + "Independent variable selection");
    tele_print(c, kLogInfo, "#variable-criterion variables,criterion");   // IVarialbeSelectable.h:27-30
    DevBuf g;
    if (m > 0 && (rc = gram(c, g))) return rc;
    for (int step = 0; step < m; ++step) {
        if (stopped(c)) return GWM_ERR_CANCELLED;
        std::vector<std::vector<int>> sets;
        for (int v : rest) {
            sets.push_back(cur);
            sets.back().push_back(v);
        }
        std::vector<double> crit;
        if ((rc = varsel_criteria(c, static_cast<double*>(g.p), has_intercept, sets, crit))) return rc;
        for (size_t j = 0; j < sets.size(); ++j) models.emplace_back(sets[j], crit[j]);
        // GWRBasic::indepVarsSelectionCriterion (GWRBasic.cpp:338-342) reports
        // each candidate set in evaluation order
        for (size_t j = 0; j < sets.size(); ++j) {
            if (crit[j] == DBL_MAX) {
                tele_print(c, kLogErr, "inv_sympd(): matrix is singular or not positive definite");
                continue;
            }
            tele_progress(c, ++progress_cur, (double)m * (m + 1) / 2);
            std::string vars_s;
            for (size_t q = 0; q < sets[j].size(); ++q) vars_s += (q ? "+" : "") + std::to_string(sets[j][q]);
            tele_print(c, kLogInfo, "#variable-criterion " + vars_s + "," + std::to_string(crit[j]));
        }
        const size_t best = std::min_element(crit.begin(), crit.end()) - crit.begin();
        cur.push_back(rest[best]);
        rest.erase(rest.begin() + best);
    }
    size_t tag = 0;
    for (int size = m; size > 0; --size) {
        std::stable_sort(models.begin() + tag, models.begin() + tag + size,
                         [](const std::pair<std::vector<int>, double>& l,
                            const std::pair<std::vector<int>, double>& r) { return l.second > r.second; });
        tag += size;
    }
    *nselected = 0;
    if (!models.empty()) {
        size_t pick = models.size() - 1;
        for (size_t i = models.size() - 1; i >= 1; --i)
            if (models[i - 1].second - models[i].second >= threshold) {
                pick = i;
                break;
            }
        for (int v : models[pick].first) selected[(*nselected)++] = v;
    }
    *nmodels = (int)models.size();
    for (int q = 0; q < std::min(cap_models, (int)models.size()); ++q) {
        model_nvars[q] = (int)models[q].first.size();
        for (int a = 0; a < k; ++a)
            model_vars[(size_t)q * k + a] = a < model_nvars[q] ? models[q].first[a] : -1;
        model_crit[q] = models[q].second;
    }
    return GWM_OK;
}

int gwm_hip_mgwr_set_distances(gwm_hip_ctx* c, int k, const int* distances) {
    if (!c || k < 0 || (k > 0 && !distances)) return GWM_ERR_INVALID;
    for (int i = 0; i < k; ++i)
        if (distances[i] != GWM_EUCLIDEAN && distances[i] != GWM_GREAT_CIRCLE) {
            c->err = "mgwr distances: variable " + std::to_string(i) + " has an unknown distance";
            return GWM_ERR_INVALID;
        }
    c->mgwr_dists.assign(distances, distances + k);
    return GWM_OK;
}

int gwm_hip_mgwr_set_hat_matrix(gwm_hip_ctx* c, int enabled) {
    if (!c) return GWM_ERR_INVALID;
    c->mgwr_hat = enabled != 0;
    return GWM_OK;
}

int gwm_hip_mgwr_last_shat(gwm_hip_ctx* c, double* shat) {
    if (!c || !shat) return GWM_ERR_INVALID;
    shat[0] = c->mgwr_shat[0];
    shat[1] = c->mgwr_shat[1];
    return GWM_OK;
}

int gwm_hip_set_certify_tolerance(gwm_hip_ctx* c, double rel) {
    if (!c || !(rel >= 0.0)) return GWM_ERR_INVALID;
    c->certify_rel = rel;
    return GWM_OK;
}

int gwm_hip_set_pruning(gwm_hip_ctx* c, int enabled) {
    if (!c) return GWM_ERR_INVALID;
    c->prune = enabled != 0;
    return GWM_OK;
}

int gwm_hip_set_refine_condition(gwm_hip_ctx* c, double criterion_limit, double fit_limit) {
    if (!c || std::isnan(criterion_limit) || std::isnan(fit_limit)) return GWM_ERR_INVALID;
    c->refine_crit = criterion_limit;
    c->refine_fit = fit_limit;
    return GWM_OK;
}

int gwm_hip_last_condition(gwm_hip_ctx* c, double* cond, long long* refined) {
    if (!c) return GWM_ERR_INVALID;
    if (cond) *cond = c->last_cond;
    if (refined) *refined = c->last_refined;
    return GWM_OK;
}

int gwm_hip_last_error_bound(gwm_hip_ctx* c, double* bound) {
    if (!c || !bound) return GWM_ERR_INVALID;
    *bound = c->last_bound;
    return GWM_OK;
}

int gwm_hip_set_autotune(gwm_hip_ctx* c, int enabled) {
    if (!c || enabled < 0 || enabled > 5) return GWM_ERR_INVALID;
    c->autotune = enabled == 1;
    c->force_shape = enabled >= 2 ? enabled - 2 : -1;
    return GWM_OK;
}

int gwm_hip_set_distance_table(gwm_hip_ctx* c, double max_bytes) {
    if (!c || !(max_bytes >= 0.0)) return GWM_ERR_INVALID;
    c->d2_cap = max_bytes >= 1.8e19 ? SIZE_MAX : (size_t)max_bytes;
    return GWM_OK;
}

int gwm_hip_set_certify_table(gwm_hip_ctx* c, double max_bytes) {
    if (!c || !(max_bytes >= 0.0)) return GWM_ERR_INVALID;
    c->cert_cap = max_bytes >= 1.8e19 ? SIZE_MAX : (size_t)max_bytes;
    return GWM_OK;
}

int gwm_hip_set_neighbour_table(gwm_hip_ctx* c, double max_bytes) {
    if (!c || !(max_bytes >= 0.0)) return GWM_ERR_INVALID;
    c->nbr_cap = max_bytes >= 1.8e19 ? SIZE_MAX : (size_t)max_bytes;
    // tables already built stay until the data change; a lower cap drops them
    for (auto& t : c->nbr) {
        if (t.keys && (size_t)c->n * (size_t)t.K * sizeof(unsigned long long) > c->nbr_cap) {
            hipFree(t.keys);
            t = gwm_hip_ctx::NbrTable{};
        }
    }
    return GWM_OK;
}


int gwm_hip_set_timing(gwm_hip_ctx* c, int enabled) {
    if (!c) return GWM_ERR_INVALID;
    c->timing = enabled != 0;
    return GWM_OK;
}

int gwm_hip_kernel_time(gwm_hip_ctx* c, const char* name, long long* launches, double* total_ms) {
    if (!c || !name) return GWM_ERR_INVALID;
    auto it = c->stats.find(name);
    if (launches) *launches = it == c->stats.end() ? 0 : it->second.launches;
    if (total_ms) *total_ms = it == c->stats.end() ? 0.0 : it->second.ms;
    return GWM_OK;
}

int gwm_hip_kernel_pairs(gwm_hip_ctx* c, const char* name, double* pairs) {
    if (!c || !name || !pairs) return GWM_ERR_INVALID;
    auto it = c->stats.find(name);
    *pairs = it == c->stats.end() ? 0.0 : it->second.pairs;
    return GWM_OK;
}

int gwm_hip_reset_timers(gwm_hip_ctx* c) {
    if (!c) return GWM_ERR_INVALID;
    c->stats.clear();
    c->solves = 0;
    return GWM_OK;
}

int gwm_hip_solve_count(gwm_hip_ctx* c, long long* solves) {
    if (!c || !solves) return GWM_ERR_INVALID;
    *solves = c->solves;
    return GWM_OK;
}

}  // extern "C"
