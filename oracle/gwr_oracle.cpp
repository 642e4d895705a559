/*
 * CPU FP64 oracle for the GWR hot path -- TEST INFRASTRUCTURE ONLY (the
 * checker and the CPU baseline; never shipped or called by the product).
 *
 * Restates libgwmodel v0.9.4's Serial/OpenMP path without Armadillo/LAPACK:
 *   - CRSDistance        src/gwmodelpp/spatialweight/CRSDistance.cpp:15-81,122-143
 *                        include/gwmodelpp/spatialweight/CRSDistance.h:95-99
 *   - BandwidthWeight    include/gwmodelpp/spatialweight/BandwidthWeight.h:68-121
 *                        src/gwmodelpp/spatialweight/BandwidthWeight.cpp:35-61
 *   - GWRBasic           src/gwmodelpp/GWRBasic.cpp:21-128,193-307,386-535
 *   - BandwidthSelector  src/gwmodelpp/BandwidthSelector.cpp:6-95
 *   - GWRMultiscale      src/gwmodelpp/GWRMultiscale.cpp:41-317,362-636
 * Build with -ffp-contract=off so that no a*b+c is fused (the reference's
 * Armadillo expressions evaluate each operation separately).
 *
 * Third-party arithmetic the reference delegates and this file restates:
 *   - arma::inv_sympd -> LAPACK dpotrf('L') + dpotri('L'): Cholesky, then the
 *     inverse from the triangular factor (Golub & Van Loan Alg. 4.2.1).
 *   - BLAS dgemm/dgemv summation order: sequential over the data index.
 *   - std::sort of the distance vector (same as arma::sort for finite data).
 * Parity pins: reference KATs on londonhp100 (test/testGWRBasic.cpp:40-240,
 * test/testGWRMultiscale.cpp:45-257), closed forms and the OLS limit
 * (tests/test_oracle.py).
 */
#include "gwr_oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const double kPi = 3.14159265358979323846;   // src/CMakeLists.txt:33 (M_PI)
const double kDoubleEps = 1e-8;              // src/CMakeLists.txt:33 (DOUBLE_EPS)

struct NotSpd : std::runtime_error {
    NotSpd() : std::runtime_error("inv_sympd(): matrix is singular or not positive definite") {}
};

int thread_id() {
#ifdef _OPENMP
    return omp_get_thread_num();
#else
    return 0;
#endif
}

/* ---------------- distances ---------------- */

/* CRSDistance::SpGcdist, CRSDistance.cpp:15-69 (WGS-84, km). */
double spgcdist(double lon1, double lon2, double lat1, double lat2) {
    double DE2RA = kPi / 180;
    double a = 6378.137;
    double f = 1.0 / 298.257223563;
    if (std::fabs(lat1 - lat2) < kDoubleEps) {
        if (std::fabs(lon1 - lon2) < kDoubleEps) return 0.0;
        else if (std::fabs((std::fabs(lon1) + std::fabs(lon2)) - 360.0) < kDoubleEps) return 0.0;
    }
    double lat1R = lat1 * DE2RA, lat2R = lat2 * DE2RA;
    double lon1R = lon1 * DE2RA, lon2R = lon2 * DE2RA;
    double F = (lat1R + lat2R) / 2.0;
    double G = (lat1R - lat2R) / 2.0;
    double L = (lon1R - lon2R) / 2.0;
    double sinG2 = std::pow(std::sin(G), 2);
    double cosG2 = std::pow(std::cos(G), 2);
    double sinF2 = std::pow(std::sin(F), 2);
    double cosF2 = std::pow(std::cos(F), 2);
    double sinL2 = std::pow(std::sin(L), 2);
    double cosL2 = std::pow(std::cos(L), 2);
    double S = sinG2 * cosL2 + cosF2 * sinL2;
    double C = cosG2 * cosL2 + sinF2 * sinL2;
    double w = std::atan(std::sqrt(S / C));
    double R = std::sqrt(S * C) / w;
    double D = 2 * w * a;
    double H1 = (3 * R - 1) / (2 * C);
    double H2 = (3 * R + 1) / (2 * S);
    return D * (1 + f * H1 * sinF2 * cosG2 - f * H2 * cosF2 * sinG2);
}

/* CRSDistance::distance(focus) (CRSDistance.cpp:122-131) with focus = data =
 * coords (SpatialMonoscaleAlgorithm.cpp:5-12). Euclidean: CRSDistance.h:95-99,
 * sqrt(dx*dx + dy*dy). Geographic: SpatialDistance, CRSDistance.cpp:71-81,
 * SpGcdist(lon_j, lon_f, lat_j, lat_f). */
void distance_vec(int n, const double* coords, bool geo, int focus, double* out) {
    const double fx = coords[focus], fy = coords[n + focus];
    if (geo) {
        for (int j = 0; j < n; ++j) out[j] = spgcdist(coords[j], fx, coords[n + j], fy);
    } else {
        for (int j = 0; j < n; ++j) {
            double dx = coords[j] - fx, dy = coords[n + j] - fy;
            double s = dx * dx;
            s = s + dy * dy;
            out[j] = std::sqrt(s);
        }
    }
}

/* Distance from a point (fx, fy) that is not a data point to every data
 * point: CRSDistance::distance(focus) with the prediction parameter
 * {locations, coords} (GWRBasic.cpp:158-165, CRSDistance.cpp:122-131). */
void distance_from(int n, const double* coords, bool geo, double fx, double fy, double* out) {
    if (geo) {
        for (int j = 0; j < n; ++j) out[j] = spgcdist(coords[j], fx, coords[n + j], fy);
    } else {
        for (int j = 0; j < n; ++j) {
            double dx = coords[j] - fx, dy = coords[n + j] - fy;
            double s = dx * dx;
            s = s + dy * dy;
            out[j] = std::sqrt(s);
        }
    }
}

/* ---------------- kernels ---------------- */

/* BandwidthWeight.h:68-121; the compact kernels use the CPU boundary (d < bw). */
inline double kernel_value(int kernel, double d, double bw) {
    switch (kernel) {
    case ORC_GAUSSIAN: return std::exp((d * d) / ((-2.0) * (bw * bw)));
    case ORC_EXPONENTIAL: return std::exp(-d / bw);
    case ORC_BISQUARE: {
        double t = 1.0 - (d * d) / (bw * bw);
        return (double)(d < bw) * (t * t);
    }
    case ORC_TRICUBE: {
        double t = 1.0 - (d * d * d) / (bw * bw * bw);
        return (double)(d < bw) * (t * t * t);
    }
    case ORC_BOXCAR: return (double)(d < bw) * 1.0;
    }
    return NAN;
}

/* BandwidthWeight::weight adaptive branch (BandwidthWeight.cpp:40-55):
 * bw/n < 1 -> linear interpolation between the floor(bw)-th and
 * (floor(bw)+1)-th smallest distances (the focus itself is rank 1);
 * else (bw/n)*max(dist). */
double adaptive_threshold(const double* dist, int n, double bw, std::vector<double>& scratch) {
    double dn = bw / n;
    if (dn < 1) {
        scratch.assign(dist, dist + n);
        std::sort(scratch.begin(), scratch.end());
        double b0 = std::floor(bw), bx = bw - b0;
        double d0 = scratch[(int)b0 - 1], d1 = scratch[(int)b0];
        return d0 + (d1 - d0) * bx;
    }
    double mx = dist[0];
    for (int j = 1; j < n; ++j) mx = std::max(mx, dist[j]);
    return dn * mx;
}

void weight_vec(const double* dist, int n, int kernel, bool adaptive, double bw, double* w,
                std::vector<double>& scratch) {
    double b = adaptive ? adaptive_threshold(dist, n, bw, scratch) : bw;
    for (int j = 0; j < n; ++j) w[j] = kernel_value(kernel, dist[j], b);
}

/* ---------------- dense helpers ---------------- */

/* arma::inv_sympd via dpotrf('L') + dpotri('L'); throws NotSpd like
 * Armadillo when a pivot is not positive. A is k*k column-major, in place. */
void inv_sympd(double* A, int k) {
    std::vector<double> L(k * k, 0.0);
    for (int j = 0; j < k; ++j) {
        double ajj = A[j + j * k];
        for (int p = 0; p < j; ++p) ajj -= L[j + p * k] * L[j + p * k];
        if (!(ajj > 0.0)) throw NotSpd();
        double ljj = std::sqrt(ajj);
        L[j + j * k] = ljj;
        for (int i = j + 1; i < k; ++i) {
            double s = A[i + j * k];
            for (int p = 0; p < j; ++p) s -= L[i + p * k] * L[j + p * k];
            L[i + j * k] = s / ljj;
        }
    }
    /* dtrtri: Linv lower triangular. */
    std::vector<double> Li(k * k, 0.0);
    for (int j = 0; j < k; ++j) {
        Li[j + j * k] = 1.0 / L[j + j * k];
        for (int i = j + 1; i < k; ++i) {
            double s = 0.0;
            for (int p = j; p < i; ++p) s -= L[i + p * k] * Li[p + j * k];
            Li[i + j * k] = s / L[i + i * k];
        }
    }
    /* dlauum: inv = Linv^T * Linv, then symmatl. */
    for (int j = 0; j < k; ++j)
        for (int i = j; i < k; ++i) {
            double s = 0.0;
            for (int p = i; p < k; ++p) s += Li[p + i * k] * Li[p + j * k];
            A[i + j * k] = s;
            A[j + i * k] = s;
        }
}

/* Per-location normal equations: xtw = (X .* w)^T, A = xtw * X, b = xtw * y
 * (GWRBasic.cpp:204-207). */
void normal_equations(int n, int k, const double* X, const double* y, const double* w,
                      double* A, double* b) {
    std::fill(A, A + k * k, 0.0);
    std::fill(b, b + k, 0.0);
    for (int j = 0; j < n; ++j) {
        double wj = w[j];
        if (wj == 0.0) continue;   // exact: adds +0 to every entry
        for (int a = 0; a < k; ++a) {
            double xa = X[j + a * n] * wj;
            for (int c = 0; c < k; ++c) A[a + c * k] += xa * X[j + c * n];
            b[a] += xa * y[j];
        }
    }
}

double mean_of(const double* v, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += v[i];
    return s / n;
}

/* GWRBase::AICc (GWRBase.h:90-94) with RSS/Fitted (GWRBase.h:41-72). */
double rss_of(int n, int k, const double* X, const double* y, const double* betas) {
    double rss = 0.0;
    for (int i = 0; i < n; ++i) {
        double f = 0.0;
        for (int a = 0; a < k; ++a) f += betas[i + a * n] * X[i + a * n];
        double r = y[i] - f;
        rss += r * r;
    }
    return rss;
}

double aicc_of(double rss, double n, double trS) {
    return n * std::log(rss / n) + n * std::log(2 * kPi) + n * ((n + trS) / (n - 2 - trS));
}

struct Problem {
    int n, k;
    const double* coords;
    const double* X;
    const double* y;
    bool geo;
};

int nthreads_or_default(int t) {
#ifdef _OPENMP
    return t > 0 ? t : omp_get_max_threads();
#else
    (void)t;
    return 1;
#endif
}

/* bandwidthSizeCriterionCVOmp (GWRBasic.cpp:440-485). */
double criterion_cv(const Problem& P, int kernel, bool adaptive, double bw, int nthreads, int focus_count) {
    const int n = P.n, k = P.k;
    const int nf = focus_count > 0 ? std::min(focus_count, n) : n;
    int T = nthreads_or_default(nthreads);
    std::vector<double> cv_all(T, 0.0);
    bool flag = true;
#pragma omp parallel num_threads(T)
    {
        std::vector<double> d(n), w(n), scratch, A(k * k), b(k), beta(k);
#pragma omp for schedule(static)
        for (int i = 0; i < nf; ++i) {
            if (!flag) continue;
            distance_vec(n, P.coords, P.geo, i, d.data());
            weight_vec(d.data(), n, kernel, adaptive, bw, w.data(), scratch);
            w[i] = 0.0;
            normal_equations(n, k, P.X, P.y, w.data(), A.data(), b.data());
            try {
                inv_sympd(A.data(), k);
                double yhat = 0.0;
                for (int a = 0; a < k; ++a) {
                    double s = 0.0;
                    for (int c = 0; c < k; ++c) s += A[a + c * k] * b[c];
                    beta[a] = s;
                }
                for (int a = 0; a < k; ++a) yhat += P.X[i + a * n] * beta[a];
                double res = P.y[i] - yhat;
                if (std::isfinite(res)) cv_all[thread_id()] += res * res;
                else flag = false;
            } catch (const NotSpd&) {
                flag = false;
            }
        }
    }
    if (!flag) return DBL_MAX;
    double cv = 0.0;
    for (double v : cv_all) cv += v;
    return cv;
}

/* Per-location hat quantities shared by the AIC criterion and the fit
 * (GWRBasic.cpp:409-419): C_i = A^-1 xtw, s_i = x_i C_i. */
struct FitOut {
    double* betas;     // n*k or nullptr
    double* betasSE;   // n*k or nullptr
    double* qDiag;     // n or nullptr
    double* S;         // n*n or nullptr
};

/* bandwidthSizeCriterionAICOmp (GWRBasic.cpp:487-535) and fitOmp
 * (GWRBasic.cpp:386-438) share this sweep. Returns false on non-SPD. */
bool hat_sweep(const Problem& P, int kernel, bool adaptive, double bw, int nthreads, int focus_count,
               double* betas, double* betasSE, double* qDiag, double* S, double shat[2]) {
    const int n = P.n, k = P.k;
    const int nf = focus_count > 0 ? std::min(focus_count, n) : n;
    int T = nthreads_or_default(nthreads);
    std::vector<double> shat_all(2 * T, 0.0);
    std::vector<std::vector<double>> q_all;
    if (qDiag) q_all.assign(T, std::vector<double>(n, 0.0));
    bool success = true;
#pragma omp parallel num_threads(T)
    {
        std::vector<double> d(n), w(n), scratch, A(k * k), b(k), ci(k * n), si(n);
#pragma omp for schedule(static)
        for (int i = 0; i < nf; ++i) {
            if (!success) continue;
            int t = thread_id();
            distance_vec(n, P.coords, P.geo, i, d.data());
            weight_vec(d.data(), n, kernel, adaptive, bw, w.data(), scratch);
            normal_equations(n, k, P.X, P.y, w.data(), A.data(), b.data());
            try {
                inv_sympd(A.data(), k);
            } catch (const NotSpd&) {
                success = false;
                continue;
            }
            for (int a = 0; a < k; ++a) {
                double s = 0.0;
                for (int c = 0; c < k; ++c) s += A[a + c * k] * b[c];
                betas[i + a * n] = s;
            }
            /* ci = inv * xtw (k x n), xtw(c, j) = X(j, c) * w(j) */
            for (int j = 0; j < n; ++j) {
                for (int a = 0; a < k; ++a) {
                    double s = 0.0;
                    for (int c = 0; c < k; ++c) s += A[a + c * k] * (P.X[j + c * n] * w[j]);
                    ci[a + j * k] = s;
                }
            }
            if (betasSE) {
                for (int a = 0; a < k; ++a) {
                    double s = 0.0;
                    for (int j = 0; j < n; ++j) s += ci[a + j * k] * ci[a + j * k];
                    betasSE[i + a * n] = s;
                }
            }
            double ss = 0.0;
            for (int j = 0; j < n; ++j) {
                double s = 0.0;
                for (int a = 0; a < k; ++a) s += P.X[i + a * n] * ci[a + j * k];
                si[j] = s;
                ss += s * s;
            }
            shat_all[2 * t] += si[i];
            shat_all[2 * t + 1] += ss;
            if (qDiag) {
                std::vector<double>& q = q_all[t];
                for (int j = 0; j < n; ++j) {
                    double p = -si[j];
                    if (j == i) p += 1.0;
                    q[j] += p * p;
                }
            }
            if (S) for (int j = 0; j < n; ++j) S[i + (size_t)j * n] = si[j];
        }
    }
    if (!success) return false;
    shat[0] = shat[1] = 0.0;
    for (int t = 0; t < T; ++t) {
        shat[0] += shat_all[2 * t];
        shat[1] += shat_all[2 * t + 1];
    }
    if (qDiag) {
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int t = 0; t < T; ++t) s += q_all[t][j];
            qDiag[j] = s;
        }
    }
    return true;
}

double criterion_aic(const Problem& P, int kernel, bool adaptive, double bw, int nthreads, int focus_count) {
    std::vector<double> betas((size_t)P.n * P.k, 0.0);
    double shat[2];
    if (!hat_sweep(P, kernel, adaptive, bw, nthreads, focus_count, betas.data(), nullptr, nullptr,
                   nullptr, shat))
        return DBL_MAX;
    double value = aicc_of(rss_of(P.n, P.k, P.X, P.y, betas.data()), (double)P.n, shat[0]);
    return std::isfinite(value) ? value : DBL_MAX;
}

double criterion(const Problem& P, int crit, int kernel, bool adaptive, double bw, int nthreads,
                 int focus_count) {
    return crit == ORC_CRIT_CV ? criterion_cv(P, kernel, adaptive, bw, nthreads, focus_count)
                               : criterion_aic(P, kernel, adaptive, bw, nthreads, focus_count);
}

/* BandwidthSelector::optimize (BandwidthSelector.cpp:6-81). f(x) returns the
 * criterion or DBL_MAX. The criterion map is keyed by bandwidth, so the list
 * comes back sorted (BandwidthSelector.cpp:83-95). */
template <class F>
double golden_section(bool adaptBw, double lower, double upper, F f, std::map<double, double>& log) {
    double xU = upper, xL = lower;
    const double eps = 1e-4;
    const double R = (std::sqrt(5) - 1) / 2;
    int iter = 0;
    double d = R * (xU - xL);
    double x1 = adaptBw ? std::floor(xL + d) : (xL + d);
    double x2 = adaptBw ? std::round(xU - d) : (xU - d);
    double f1 = f(x1);
    double f2 = f(x2);
    if (f1 == DBL_MAX && f2 == DBL_MAX) throw std::invalid_argument("Invalid initial values.");
    if (f1 < DBL_MAX) log[x1] = f1;
    if (f2 < DBL_MAX) log[x2] = f2;
    double d1 = f2 - f1;
    double xopt = f1 < f2 ? x1 : x2;
    double ea = 100;
    while ((std::fabs(d) > eps) && (std::fabs(d1) > eps) && iter < ea) {
        d = R * d;
        if (f1 < f2) {
            xL = x2;
            x2 = x1;
            x1 = adaptBw ? std::round(xL + d) : (xL + d);
            f2 = f1;
            f1 = f(x1);
            if (f1 < DBL_MAX) log[x1] = f1;
        } else {
            xU = x1;
            x1 = x2;
            x2 = adaptBw ? std::floor(xU - d) : (xU - d);
            f1 = f2;
            f2 = f(x2);
            if (f2 < DBL_MAX) log[x2] = f2;
        }
        iter = iter + 1;
        xopt = (f1 < f2) ? x1 : x2;
        d1 = f2 - f1;
    }
    return xopt;
}

void write_trace(const std::map<double, double>& log, double* tx, double* tf, int cap, int* len) {
    int m = 0;
    for (auto& kv : log) {
        if (m < cap) {
            if (tx) tx[m] = kv.first;
            if (tf) tf[m] = kv.second;
        }
        ++m;
    }
    if (len) *len = m;
}

/* GWRBasic::CalcDiagnostic (GWRBasic.cpp:21-34). */
void calc_diagnostic(int n_, int k, const double* X, const double* y, const double* betas,
                     const double* shat, double* diag) {
    double rss = rss_of(n_, k, X, y, betas);
    double n = (double)n_;
    double AIC = n * std::log(rss / n) + n * std::log(2 * kPi) + n + shat[0];
    double AICc = n * std::log(rss / n) + n * std::log(2 * kPi) + n * ((n + shat[0]) / (n - 2 - shat[0]));
    double edf = n - 2 * shat[0] + shat[1];
    double enp = 2 * shat[0] - shat[1];
    double my = mean_of(y, n_);
    double yss = 0.0;
    for (int i = 0; i < n_; ++i) yss += (y[i] - my) * (y[i] - my);
    double r2 = 1 - rss / yss;
    double r2_adj = 1 - (1 - r2) * (n - 1) / (edf - 1);
    diag[0] = rss; diag[1] = AIC; diag[2] = AICc; diag[3] = enp; diag[4] = edf; diag[5] = r2; diag[6] = r2_adj;
}

/* ---------------- GWRMultiscale ---------------- */

/* Single-column local solve (GWRMultiscale.cpp:420-482, 560-636): 1x1
 * inv_sympd fails when the scalar is not positive. */
inline bool inv1(double a, double& inv) {
    if (!(a > 0.0)) return false;
    double l = std::sqrt(a);
    double li = 1.0 / l;
    inv = li * li;
    return true;
}

struct Mgwr {
    const Problem* P;      // X here is the centred design
    std::vector<double> Xc;
    int kernel_of(int v) const { return kernels[v]; }
    std::vector<int> kernels;
    std::vector<int> adaptive;
    std::vector<double> bws;
    std::vector<int> geo;   // per variable: great-circle distance (GWRMultiscale.h:578, one SpatialWeight each)
    int nthreads;
    bool has_hat;
    std::vector<double> S0;       // n*n
    std::vector<double> SArray;   // n*n*k
    std::vector<double> C;        // k*n*n (slice i: k x n)
    std::vector<double> betasSE;

    const double* col(int v) const { return Xc.data() + (size_t)v * P->n; }

    /* bandwidthSizeCriterionAll{CV,AIC}Omp (GWRMultiscale.cpp:484-558), the
     * all-variable criterion with the init spatial weight's distance. */
    double all_criterion(int crit, int kernel, bool adaptive_, double bw) {
        Problem Q = *P;
        Q.X = Xc.data();
        Q.geo = geo[0] != 0;   // the init spatial weight is spatial weight 0 (GWRMultiscale.cpp:1476-1483)
        return criterion(Q, crit, kernel, adaptive_, bw, nthreads, 0);
    }

    /* bandwidthSizeCriterionVar{CV,AIC}Omp (GWRMultiscale.cpp:560-636). */
    double var_criterion(int v, int crit, int kernel, bool adaptive_, double bw, const double* xi, const double* yi) {
        const int n = P->n;
        const bool g = geo[v] != 0;   // mSpatialWeights[v].distance() (GWRMultiscale.cpp:560-636)
        int T = nthreads_or_default(nthreads);
        std::vector<double> acc(2 * T, 0.0);
        std::vector<double> betas(n, 0.0);
        bool flag = true;
#pragma omp parallel num_threads(T)
        {
            std::vector<double> d(n), w(n), scratch;
#pragma omp for schedule(static)
            for (int i = 0; i < n; ++i) {
                if (!flag) continue;
                int t = thread_id();
                distance_vec(n, P->coords, g, i, d.data());
                weight_vec(d.data(), n, kernel, adaptive_, bw, w.data(), scratch);
                if (crit == ORC_CRIT_CV) w[i] = 0.0;
                double a = 0.0, b = 0.0;
                for (int j = 0; j < n; ++j) {
                    double xw = xi[j] * w[j];
                    a += xw * xi[j];
                    b += xw * yi[j];
                }
                double inv;
                if (!inv1(a, inv)) { flag = false; continue; }
                double beta = inv * b;
                if (crit == ORC_CRIT_CV) {
                    double res = yi[i] - xi[i] * beta;
                    if (std::isfinite(res)) acc[2 * t] += res * res;
                    else flag = false;
                } else {
                    betas[i] = beta;
                    double ss = 0.0, sii = 0.0;
                    for (int j = 0; j < n; ++j) {
                        double s = xi[i] * (inv * (xi[j] * w[j]));
                        if (j == i) sii = s;
                        ss += s * s;
                    }
                    acc[2 * t] += sii;
                    acc[2 * t + 1] += ss;
                }
            }
        }
        if (!flag) return DBL_MAX;
        if (crit == ORC_CRIT_CV) {
            double cv = 0.0;
            for (int t = 0; t < T; ++t) cv += acc[2 * t];
            return std::isfinite(cv) ? cv : DBL_MAX;
        }
        double trS = 0.0;
        for (int t = 0; t < T; ++t) trS += acc[2 * t];
        double value = aicc_of(rss_of(n, 1, xi, yi, betas.data()), (double)n, trS);
        return std::isfinite(value) ? value : DBL_MAX;
    }

    /* fitAll (GWRMultiscale.cpp:362-418) with the init spatial weight. */
    void fit_all(int kernel, bool adaptive_, double bw, const double* y, double* betas) {
        const int n = P->n, k = P->k;
        bool success = true;
        int T = nthreads_or_default(nthreads);
        if (has_hat) betasSE.assign((size_t)n * k, 0.0);
#pragma omp parallel num_threads(T)
        {
            std::vector<double> d(n), w(n), scratch, A(k * k), b(k);
#pragma omp for schedule(static)
            for (int i = 0; i < n; ++i) {
                if (!success) continue;
                distance_vec(n, P->coords, geo[0] != 0, i, d.data());
                weight_vec(d.data(), n, kernel, adaptive_, bw, w.data(), scratch);
                normal_equations(n, k, Xc.data(), y, w.data(), A.data(), b.data());
                try { inv_sympd(A.data(), k); } catch (const NotSpd&) { success = false; continue; }
                for (int a = 0; a < k; ++a) {
                    double s = 0.0;
                    for (int c = 0; c < k; ++c) s += A[a + c * k] * b[c];
                    betas[i + a * n] = s;
                }
                if (has_hat) {
                    double* ci = C.data() + (size_t)i * k * n;
                    for (int j = 0; j < n; ++j)
                        for (int a = 0; a < k; ++a) {
                            double s = 0.0;
                            for (int c = 0; c < k; ++c) s += A[a + c * k] * (Xc[j + (size_t)c * n] * w[j]);
                            ci[a + (size_t)j * k] = s;
                        }
                    for (int a = 0; a < k; ++a) {
                        double s = 0.0;
                        for (int j = 0; j < n; ++j) s += ci[a + (size_t)j * k] * ci[a + (size_t)j * k];
                        betasSE[i + (size_t)a * n] = s;
                    }
                    for (int j = 0; j < n; ++j) {
                        double s = 0.0;
                        for (int a = 0; a < k; ++a) s += Xc[i + (size_t)a * n] * ci[a + (size_t)j * k];
                        S0[i + (size_t)j * n] = s;
                    }
                }
            }
        }
        if (!success) throw NotSpd();
    }

    /* fitVar (GWRMultiscale.cpp:420-482); S (n*n) filled when has_hat. */
    void fit_var(int v, const double* xi, const double* yi, double* beta_col, double* S) {
        const int n = P->n;
        bool success = true;
        int T = nthreads_or_default(nthreads);
        int kernel = kernels[v];
        bool adapt = adaptive[v] != 0;
        double bw = bws[v];
#pragma omp parallel num_threads(T)
        {
            std::vector<double> d(n), w(n), scratch;
#pragma omp for schedule(static)
            for (int i = 0; i < n; ++i) {
                if (!success) continue;
                distance_vec(n, P->coords, geo[v] != 0, i, d.data());
                weight_vec(d.data(), n, kernel, adapt, bw, w.data(), scratch);
                double a = 0.0, b = 0.0;
                for (int j = 0; j < n; ++j) {
                    double xw = xi[j] * w[j];
                    a += xw * xi[j];
                    b += xw * yi[j];
                }
                double inv;
                if (!inv1(a, inv)) { success = false; continue; }
                beta_col[i] = inv * b;
                if (S)
                    for (int j = 0; j < n; ++j) S[i + (size_t)j * n] = xi[i] * (inv * (xi[j] * w[j]));
            }
        }
        if (!success) throw NotSpd();
    }
};

/* dense n x n product helper for the hat-matrix branch: out = A * B */
void matmul_nn(int n, const double* A, const double* B, double* out) {
    std::fill(out, out + (size_t)n * n, 0.0);
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; ++j)
        for (int p = 0; p < n; ++p) {
            double b = B[p + (size_t)j * n];
            if (b == 0.0) continue;
            for (int i = 0; i < n; ++i) out[i + (size_t)j * n] += A[i + (size_t)p * n] * b;
        }
}

}  // namespace

extern "C" {

double orc_spgcdist(double lon1, double lon2, double lat1, double lat2) {
    return spgcdist(lon1, lon2, lat1, lat2);
}

int orc_distance_vec(int n, const double* coords, int geographic, int focus, double* out) {
    if (n <= 0 || focus < 0 || focus >= n) return -1;
    distance_vec(n, coords, geographic != 0, focus, out);
    return 0;
}

int orc_weight_vec(int n, const double* dist, int kernel, int adaptive, double bw, double* out) {
    std::vector<double> scratch;
    weight_vec(dist, n, kernel, adaptive != 0, bw, out, scratch);
    return 0;
}

int orc_adaptive_threshold(int n, const double* coords, int geographic, double bw, int focus, double* out) {
    std::vector<double> d(n), scratch;
    distance_vec(n, coords, geographic != 0, focus, d.data());
    *out = adaptive_threshold(d.data(), n, bw, scratch);
    return 0;
}

double orc_max_distance(int n, const double* coords, int geographic, int nthreads) {
    int T = nthreads_or_default(nthreads);
    double maxD = 0.0;
#pragma omp parallel num_threads(T)
    {
        std::vector<double> d(n);
        double local = 0.0;
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            distance_vec(n, coords, geographic != 0, i, d.data());
            double m = d[0];
            for (int j = 1; j < n; ++j) m = std::max(m, d[j]);
            local = m > local ? m : local;
        }
#pragma omp critical
        maxD = local > maxD ? local : maxD;
    }
    return maxD;
}

double orc_gwr_criterion(int n, int k, const double* coords, const double* X, const double* y,
                         int geographic, int kernel, int adaptive, double bw, int crit,
                         int nthreads, int focus_count) {
    Problem P{n, k, coords, X, y, geographic != 0};
    return criterion(P, crit, kernel, adaptive != 0, bw, nthreads, focus_count);
}

int orc_gwr_fit_core(int n, int k, const double* coords, const double* X, const double* y,
                     int geographic, int kernel, int adaptive, double bw, int nthreads,
                     double* betas, double* betasSE, double* shat, double* qDiag, double* S,
                     int focus_count) {
    Problem P{n, k, coords, X, y, geographic != 0};
    return hat_sweep(P, kernel, adaptive != 0, bw, nthreads, focus_count, betas, betasSE, qDiag, S, shat) ? 0 : 1;
}

int orc_gwr_predict(int n, int k, const double* coords, const double* X, const double* y, int geographic,
                    int m, const double* locations, int kernel, int adaptive, double bw, int nthreads,
                    double* betas) {
    const int T = nthreads_or_default(nthreads);
    int fail = 0;
#pragma omp parallel num_threads(T)
    {
        std::vector<double> d(n), w(n), scratch, A(k * k), b(k);
#pragma omp for schedule(static)
        for (int i = 0; i < m; ++i) {
            distance_from(n, coords, geographic != 0, locations[i], locations[m + i], d.data());
            weight_vec(d.data(), n, kernel, adaptive != 0, bw, w.data(), scratch);
            normal_equations(n, k, X, y, w.data(), A.data(), b.data());
            try {
                inv_sympd(A.data(), k);
                for (int a = 0; a < k; ++a) {
                    double s = 0.0;
                    for (int c = 0; c < k; ++c) s += A[a + c * k] * b[c];
                    betas[i + (size_t)a * m] = s;
                }
            } catch (const NotSpd&) {
                for (int a = 0; a < k; ++a) betas[i + (size_t)a * m] = NAN;
#pragma omp atomic write
                fail = 1;
            }
        }
    }
    return fail;
}

double orc_indep_vars_criterion(int n, int k, const double* X, const double* y, int has_intercept,
                                const int* vars, int nvars) {
    // x = mX.cols(index2uvec(vars, hasIntercept)) (VariableForwardSelector.h:26-35)
    std::vector<int> cols;
    if (has_intercept) cols.push_back(0);
    for (int v = 0; v < nvars; ++v) cols.push_back(vars[v]);
    const int p = (int)cols.size();
    std::vector<double> x((size_t)n * p), w(n, 1.0), A(p * p), b(p), beta(p), ci(p);
    for (int a = 0; a < p; ++a)
        for (int i = 0; i < n; ++i) x[i + (size_t)a * n] = X[i + (size_t)cols[a] * n];
    (void)k;
    // unit weights (GWRBasic.cpp:319): every location has the same xtwx, xtwy
    normal_equations(n, p, x.data(), y, w.data(), A.data(), b.data());
    try {
        inv_sympd(A.data(), p);
    } catch (const NotSpd&) {
        return DBL_MAX;
    }
    for (int a = 0; a < p; ++a) {
        double t = 0.0;
        for (int c = 0; c < p; ++c) t += A[a + c * p] * b[c];
        beta[a] = t;
    }
    // shat(0) = sum_i si(0, i) with si = x_i^T inv(xtwx) xtw, i.e. x_i^T inv x_i
    double tr = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < p; ++a) {
            double t = 0.0;
            for (int c = 0; c < p; ++c) t += A[a + c * p] * x[i + (size_t)c * n];
            ci[a] = t;
        }
        double sii = 0.0;
        for (int a = 0; a < p; ++a) sii += x[i + (size_t)a * n] * ci[a];
        tr += sii;
    }
    std::vector<double> betas((size_t)n * p);
    for (int a = 0; a < p; ++a)
        for (int i = 0; i < n; ++i) betas[i + (size_t)a * n] = beta[a];
    const double v = aicc_of(rss_of(n, p, x.data(), y, betas.data()), n, tr);
    return std::isfinite(v) ? v : DBL_MAX;
}

int orc_indep_vars_select(int n, int k, const double* X, const double* y, int has_intercept, double threshold,
                          int* selected, int* nselected, int* model_vars, int* model_nvars, double* model_crit,
                          int* nmodels) {
    // VariableForwardSelector::optimize (VariableForwardSelector.cpp:9-42)
    std::vector<int> vars;
    for (int c = has_intercept ? 1 : 0; c < k; ++c) vars.push_back(c);
    const int m = (int)vars.size();
    std::vector<int> cur, rest;
    for (int i = 0; i < m; ++i) rest.push_back(i);
    std::vector<std::pair<std::vector<int>, double>> models;
    auto to_vars = [&](const std::vector<int>& idx) {
        std::vector<int> v;
        for (int i : idx) v.push_back(vars[i]);
        return v;
    };
    for (int i = 0; i < m; ++i) {
        std::vector<double> crit(rest.size());
        for (size_t j = 0; j < rest.size(); ++j) {
            cur.push_back(rest[j]);
            const std::vector<int> v = to_vars(cur);
            crit[j] = orc_indep_vars_criterion(n, k, X, y, has_intercept, v.data(), (int)v.size());
            models.emplace_back(cur, crit[j]);
            cur.pop_back();
        }
        const size_t best = std::min_element(crit.begin(), crit.end()) - crit.begin();
        cur.push_back(rest[best]);
        rest.erase(rest.begin() + best);
    }
    // sort (:53-65): each group of equal size in descending criterion
    size_t tag = 0;
    for (int g = m; g > 0; --g) {
        std::stable_sort(models.begin() + tag, models.begin() + tag + g,
                         [](const std::pair<std::vector<int>, double>& l, const std::pair<std::vector<int>, double>& r) {
                             return l.second > r.second;
                         });
        tag += g;
    }
    // select (:67-77): the last model after which the criterion drops by >= threshold
    *nselected = 0;
    if (!models.empty()) {
        size_t pick = models.size() - 1;
        bool found = false;
        for (size_t i = models.size() - 1; i >= 1 && !found; --i)
            if (models[i - 1].second - models[i].second >= threshold) {
                pick = i;
                found = true;
            }
        for (int v : to_vars(models[pick].first)) selected[(*nselected)++] = v;
    }
    *nmodels = (int)models.size();
    for (size_t q = 0; q < models.size(); ++q) {
        const std::vector<int> v = to_vars(models[q].first);
        model_nvars[q] = (int)v.size();
        for (int a = 0; a < k; ++a) model_vars[q * k + a] = a < (int)v.size() ? v[a] : -1;
        model_crit[q] = models[q].second;
    }
    return 0;
}

int orc_gwr_run(int n, int k, const double* coords, const double* X, const double* y,
                int geographic, int kernel, int adaptive, double bw,
                int select_bw, int crit, int has_lower, double lower, int has_upper, double upper,
                int nthreads,
                double* betas, double* betasSE, double* shat, double* qDiag, double* S,
                double* diag, double* bw_out, double* trace_x, double* trace_f, int trace_cap,
                int* trace_len) {
    Problem P{n, k, coords, X, y, geographic != 0};
    bool adapt = adaptive != 0;
    if (trace_len) *trace_len = 0;
    if (select_bw) {
        /* GWRBasic.cpp:74-96: lower = value_or(adaptive ? 20 : 0), upper =
         * value_or(adaptive ? n : maxDistance()). */
        double lo = has_lower ? lower : (adapt ? 20 : 0.0);
        double hi = has_upper ? upper : (adapt ? (double)n : orc_max_distance(n, coords, geographic, nthreads));
        std::map<double, double> log;
        try {
            bw = golden_section(adapt, lo, hi,
                                [&](double x) { return criterion(P, crit, kernel, adapt, x, nthreads, 0); }, log);
        } catch (const std::invalid_argument&) {
            return 2;
        }
        write_trace(log, trace_x, trace_f, trace_cap, trace_len);
    }
    if (bw_out) *bw_out = bw;
    if (!hat_sweep(P, kernel, adapt, bw, nthreads, 0, betas, betasSE, qDiag, S, shat)) return 1;
    calc_diagnostic(n, k, X, y, betas, shat, diag);
    /* GWRBasic.cpp:100-103: sigma = RSS / (n - 2 trS + trStS); SE = sqrt(sigma * SE) */
    double sigmaHat = diag[0] / (n - 2 * shat[0] + shat[1]);
    for (size_t i = 0; i < (size_t)n * k; ++i) betasSE[i] = std::sqrt(sigmaHat * betasSE[i]);
    return 0;
}

int orc_mgwr_run(int n, int k, const double* coords, const double* X, const double* y,
                 const int* geographic, const int* kernels, const int* adaptive, const double* bws,
                 const int* init_type_in, const int* approach, const int* centered,
                 const double* bw_threshold, int has_intercept, int has_hat,
                 int crit_type, double crit_threshold, int max_iter, int retry_times,
                 int adaptive_lower, int has_lower, double lower, int has_upper, double upper,
                 int nthreads,
                 double* betas_out, double* bws_out, double* diag, double* rss0_out,
                 int* iterations_out) {
    Problem P{n, k, coords, X, y, geographic[0] != 0};
    Mgwr M;
    M.P = &P;
    M.geo.assign(geographic, geographic + k);
    M.kernels.assign(kernels, kernels + k);
    M.adaptive.assign(adaptive, adaptive + k);
    M.bws.assign(bws, bws + k);
    M.nthreads = nthreads;
    M.has_hat = has_hat != 0;
    std::vector<int> init_type(init_type_in, init_type_in + k);
    /* GWRMultiscale.cpp:88-99: centre predictors. */
    M.Xc.assign(X, X + (size_t)n * k);
    for (int i = has_intercept ? 1 : 0; i < k; ++i)
        if (centered[i]) {
            double m = mean_of(M.col(i), n);
            for (int j = 0; j < n; ++j) M.Xc[j + (size_t)i * n] -= m;
        }
    /* The init spatial weight is a copy of spatial weight 0 at set time
     * (GWRMultiscale.cpp:1476-1483). */
    int init_kernel = kernels[0];
    bool init_adaptive = adaptive[0] != 0;
    std::vector<double> maxDistances(k, 0.0);
    try {
        /* GWRMultiscale.cpp:104-135: initial per-variable bandwidths for Null vars. */
        for (int i = 0; i < k; ++i) {
            if (init_type[i] != 0) continue;
            bool adapt = adaptive[i] != 0;
            maxDistances[i] = orc_max_distance(n, coords, geographic[i], nthreads);   // GWRMultiscale.cpp:117
            double lo = has_lower ? lower : (adapt ? (double)adaptive_lower : maxDistances[i] / 5000.0);
            double hi = has_upper ? upper : (adapt ? (double)n : maxDistances[i]);
            std::map<double, double> log;
            const double* xi = M.col(i);
            M.bws[i] = golden_section(adapt, lo, hi, [&](double x) {
                return M.var_criterion(i, approach[i], M.kernels[i], adapt, x, xi, y);
            }, log);
        }
        /* GWRMultiscale.cpp:137-160: initial all-variable bandwidth. */
        double maxDist = orc_max_distance(n, coords, geographic[0], nthreads);   // :147, spatial weight 0
        bool adapt0 = adaptive[0] != 0;
        double lo = has_lower ? lower : (adapt0 ? (double)adaptive_lower : maxDist / 5000.0);
        double hi = has_upper ? upper : (adapt0 ? (double)n : maxDist);
        std::map<double, double> log0;
        int k0 = M.kernels[0];
        double initBw = golden_section(adapt0, lo, hi, [&](double x) {
            return M.all_criterion(approach[0], k0, adapt0, x);
        }, log0);
        (void)init_kernel;
        (void)init_adaptive;

        if (M.has_hat) {
            M.S0.assign((size_t)n * n, 0.0);
            M.SArray.assign((size_t)n * n * k, 0.0);
            M.C.assign((size_t)k * n * n, 0.0);
        }
        /* backfitting, GWRMultiscale.cpp:208-317 */
        std::vector<double> betas((size_t)n * k, 0.0);
        M.fit_all(k0, adapt0, initBw, y, betas.data());
        if (M.has_hat) {
            for (int i = 0; i < k; ++i)
                for (int j = 0; j < n; ++j) {
                    const double* cj = M.C.data() + (size_t)j * k * n;
                    double xji = M.Xc[j + (size_t)i * n];
                    double* slice = M.SArray.data() + (size_t)i * n * n;
                    for (int c = 0; c < n; ++c) slice[j + (size_t)c * n] = xji * cj[i + (size_t)c * k];
                }
        }
        std::vector<int> bwChangeNo(k, 0);
        std::vector<double> resid(n), yi(n), Svar, tmp;
        auto fitted_resid = [&]() {
            for (int j = 0; j < n; ++j) {
                double f = 0.0;
                for (int a = 0; a < k; ++a) f += betas[j + (size_t)a * n] * M.Xc[j + (size_t)a * n];
                resid[j] = y[j] - f;
            }
        };
        fitted_resid();
        double RSS0 = 0.0;
        for (int j = 0; j < n; ++j) RSS0 += resid[j] * resid[j];
        double RSS1 = DBL_MAX, crit = DBL_MAX;
        int iteration = 1;
        for (; iteration <= max_iter && crit > crit_threshold; iteration++) {
            for (int i = 0; i < k; ++i) {
                const double* xi = M.col(i);
                for (int j = 0; j < n; ++j) yi[j] = resid[j] + betas[j + (size_t)i * n] * xi[j];
                if (init_type[i] != 2) {
                    bool adapt = adaptive[i] != 0;
                    double maxD = maxDistances[i];
                    double lo2 = has_lower ? lower : (adapt ? (double)adaptive_lower : maxD / 5000.0);
                    double hi2 = has_upper ? upper : (adapt ? (double)n : maxD);
                    std::map<double, double> log;
                    double bw0 = M.bws[i];
                    double bw1 = golden_section(adapt, lo2, hi2, [&](double x) {
                        return M.var_criterion(i, approach[i], M.kernels[i], adapt, x, xi, yi.data());
                    }, log);
                    if (std::fabs(bw1 - bw0) > bw_threshold[i]) bwChangeNo[i] = 0;
                    else {
                        bwChangeNo[i] += 1;
                        if (bwChangeNo[i] >= retry_times) init_type[i] = 2;
                    }
                    M.bws[i] = bw1;
                }
                std::vector<double> col(n);
                if (M.has_hat) Svar.assign((size_t)n * n, 0.0);
                M.fit_var(i, xi, yi.data(), col.data(), M.has_hat ? Svar.data() : nullptr);
                for (int j = 0; j < n; ++j) betas[j + (size_t)i * n] = col[j];
                if (M.has_hat) {
                    /* SArrayi = SArray_i - S0; SArray_i = S*SArrayi + S; S0 = SArray_i - SArrayi */
                    double* Si = M.SArray.data() + (size_t)i * n * n;
                    std::vector<double> SArrayi((size_t)n * n);
                    for (size_t q = 0; q < (size_t)n * n; ++q) SArrayi[q] = Si[q] - M.S0[q];
                    tmp.assign((size_t)n * n, 0.0);
                    matmul_nn(n, Svar.data(), SArrayi.data(), tmp.data());
                    for (size_t q = 0; q < (size_t)n * n; ++q) Si[q] = tmp[q] + Svar[q];
                    for (size_t q = 0; q < (size_t)n * n; ++q) M.S0[q] = Si[q] - SArrayi[q];
                }
                fitted_resid();
            }
            RSS1 = rss_of(n, k, M.Xc.data(), y, betas.data());
            crit = (crit_type == 0) ? std::fabs(RSS1 - RSS0) : std::sqrt(std::fabs(RSS1 - RSS0) / RSS1);
            RSS0 = RSS1;
        }
        if (iterations_out) *iterations_out = iteration - 1;
        /* GWRMultiscale.cpp:170-180 + CalcDiagnostic :41-66 */
        double shat[2] = {0.0, 0.0};
        if (M.has_hat) {
            for (int j = 0; j < n; ++j) shat[0] += M.S0[j + (size_t)j * n];
            /* trace(S0^T S0) = sum of squares of all entries */
            double s = 0.0;
            for (size_t q = 0; q < (size_t)n * n; ++q) s += M.S0[q] * M.S0[q];
            shat[1] = s;
        }
        double nd = (double)n;
        double sigmaHat21 = RSS0 / nd;
        double my = mean_of(y, n);
        double TSS = 0.0;
        for (int j = 0; j < n; ++j) TSS += (y[j] - my) * (y[j] - my);
        double r2 = 1 - RSS0 / TSS;
        double enp = 2 * shat[0] - shat[1];
        double edf = nd - 2 * shat[0] + shat[1];
        double AICc = nd * std::log(sigmaHat21) + nd * std::log(2 * kPi) + nd * ((nd + shat[0]) / (nd - 2 - shat[0]));
        double adj = 1 - (1 - r2) * (nd - 1) / (edf - 1);
        diag[0] = RSS0; diag[1] = 0.0; diag[2] = AICc; diag[3] = enp; diag[4] = edf; diag[5] = r2; diag[6] = adj;
        if (rss0_out) *rss0_out = RSS0;
        std::copy(betas.begin(), betas.end(), betas_out);
        std::copy(M.bws.begin(), M.bws.end(), bws_out);
    } catch (const NotSpd&) {
        return 1;
    } catch (const std::invalid_argument&) {
        return 2;
    }
    return 0;
}

}  // extern "C"
