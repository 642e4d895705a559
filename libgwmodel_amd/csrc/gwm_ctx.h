// Private host-side definitions shared by the single-problem API
// (gwm_api.cpp) and the batch/serving scheduler (gwm_batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/gwmodel_hip.h"
#include "gwm_contract.h"

namespace gwm {
namespace detail {

struct KernelStat {
    long long launches = 0;
    double ms = 0.0;
};

}  // namespace detail
}  // namespace gwm

struct gwm_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int (*stop)(void*) = nullptr;
    void* stop_user = nullptr;

    // problem
    int n = 0, k = 0, dist = 0;
    int nta = 0, ntb = 0, m = 0;
    int n_pad = 0, rows_pad = 0;
    std::vector<double> hcoords, hX, hy;

    // device buffers
    double *px = nullptr, *py = nullptr, *hx = nullptr, *hy_ = nullptr, *hz = nullptr, *sl = nullptr;
    double *X = nullptr, *y = nullptr, *Z = nullptr, *G = nullptr;
    double* part = nullptr;
    size_t part_bytes = 0;
    double *bwi = nullptr, *res2 = nullptr, *sii = nullptr, *ssi = nullptr;
    double *betas = nullptr, *se = nullptr, *gvec = nullptr, *qd = nullptr, *red = nullptr, *blockmax = nullptr;
    double *vx = nullptr, *vy = nullptr, *vZ = nullptr;
    double* S = nullptr;
    int* fail = nullptr;
    // FP16 hi/lo planes for the mixed-precision sweeps (Z, G and the
    // single-column table) with their per-column unscale factors
    _Float16 *Zh = nullptr, *Zl = nullptr, *Gh = nullptr, *Gl = nullptr, *vZh = nullptr, *vZl = nullptr;
    double *zscale = nullptr, *gscale = nullptr, *vscale = nullptr;
    unsigned long long* colmax = nullptr;
    double certify_rel = 1e-7;   // relative bound used to certify mixed comparisons

    // timing
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::map<std::string, gwm::detail::KernelStat> stats;
    long long solves = 0;

    gwm::Points points() const {
        gwm::Points P;
        P.px = px;
        P.py = py;
        P.hx = hx;
        P.hy = hy_;
        P.hz = hz;
        P.sl = sl;
        P.n = n;
        return P;
    }
};


namespace gwm {
namespace detail {

int nta_of(int k);
int ntb_of(int k);
double aicc(double rss, double n, double trS);
// GWRBasic::CalcDiagnostic (src/gwmodelpp/GWRBasic.cpp:21-34): diag[7].
void calc_diagnostic(int n, int k, const double* X, const double* y, const double* betas, const double* shat,
                     double* diag);
typedef std::function<int(double, bool, double*)> CritFn;
int golden(bool adaptBw, double lower, double upper, const CritFn& f, bool certify, double tol_rel, double* xopt_out,
           std::map<double, double>& log);
void write_trace(const std::map<double, double>& log, double* tx, double* tf, int cap, int* len);
int max_distance_impl(gwm_hip_ctx* c, double* out);

}  // namespace detail
}  // namespace gwm
