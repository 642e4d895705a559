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
    double pairs = 0.0;   // sweeps: focus x data pairs actually swept
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
    std::vector<double> hcoords, hX, hy;   // the caller's order
    // Spatial order: every per-point device array holds the points sorted
    // along a Hilbert curve, so that runs of 64 points are compact; perm[s] is
    // the caller's index of sorted position s.  Host copies in and out of the
    // device are permuted at the API boundary.
    std::vector<int> perm;
    bool prune = true;             // skip data chunks whose weights are all exactly zero
    double* boxes = nullptr;       // bounding boxes of the 64-point chunks (sorted order)
    double* chunk_bw = nullptr;    // largest per-point bandwidth of each chunk (qDiag sweeps)
    // work order of large pruned sweeps (launch_sweep_plan): per-item chunk
    // counts and the item order, plan_items ints each
    int* plan_counts = nullptr;
    int* plan_order = nullptr;
    size_t plan_items = 0;
    double max_dist = -1.0;        // largest pairwise distance, once computed (< 0: not yet)
    // adaptive bandwidth vectors of recently used bandwidths (least recently
    // used replaced): searches revisit bandwidths (FP64 confirmations, the
    // fixed first candidates of every multiscale search)
    struct BwEntry {
        double bw;
        int dist;   // the distance the thresholds were selected with
        double* buf;
        unsigned long long used;
    };
    std::vector<BwEntry> bw_cache;
    unsigned long long bw_clock = 0;
    // Neighbour tables, one per distance (launch_neighbour_table): the K
    // smallest selection keys of every point, sorted; built on the second
    // select of a problem that they would answer, when n K 8 bytes fit
    // nbr_cap (gwm_hip_set_neighbour_table) and the device beside 8 GiB
    struct NbrTable {
        unsigned long long* keys = nullptr;
        int K = 0;
        int misses = 0;   // selects the table would have answered before it was built
    };
    NbrTable nbr[2];
    size_t nbr_cap = size_t(16) << 30;
    int* sel_left = nullptr;       // [n] points the one-wave select leaves to the workgroup select
    unsigned int* sel_nleft = nullptr;
    // [n] distance from each point to its nearest other point (the adaptive
    // threshold of 2 neighbours, self included), once computed (ensure_dnn):
    // CV sweeps weigh relative to that point's weight
    double* dnn = nullptr;
    // Distance-dependent tables.  `dist`, `boxes`, `max_dist` and `dnn` are
    // those of the active distance; a multiscale run whose variables use
    // different distances (gwm_hip_mgwr_set_distances) switches between
    // them with use_distance, which parks the inactive distance's tables here.
    int dist_data = 0;                  // the distance of gwm_hip_set_data
    struct DistTables {
        double* boxes = nullptr;
        double max_dist = -1.0;
        double* dnn = nullptr;
        bool built = false;
    };
    DistTables parked[2];
    std::vector<int> mgwr_dists;        // per-variable distances of the next multiscale run
    // ITelegram (include/gwmodelpp/Logger.h:31-83): messages and progress
    void (*print)(void*, int, const char*) = nullptr;
    void (*progress)(void*, double, double) = nullptr;
    void* tele_user = nullptr;
    double last_crit = 1.7976931348623157e308;   // mBandwidthLastCriterion (DBL_MAX at start)

    // device buffers
    double *px = nullptr, *py = nullptr, *hx = nullptr, *hy_ = nullptr, *hz = nullptr, *sl = nullptr;
    double *X = nullptr, *y = nullptr, *Z = nullptr, *G = nullptr;
    double* part = nullptr;
    size_t part_bytes = 0;
    double *bwi = nullptr, *res2 = nullptr, *sii = nullptr, *ssi = nullptr, *cond = nullptr;
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
    // Mixed precision: focus points whose local design has a conditioning
    // indicator above the limit (criterion evaluations / fits) are recomputed
    // in FP64 (gwm_hip_set_refine_condition).
    double refine_crit = 10.0, refine_fit = 2.0;
    double last_cond = 0.0;      // indicator maximum of the last mixed evaluation
    long long last_refined = 0;  // focus points it recomputed in FP64
    double last_bound = 0.0;     // error bound of the last mixed criterion value (crit_error_bound)
    int* rlist = nullptr;        // [n] focus list of the refinement
    int* rcount = nullptr;
    double* rpart = nullptr;     // partial sums of the refinement sweep
    size_t rpart_bytes = 0;

    // timing
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    unsigned long long* pairs_dev = nullptr;   // swept-pair counter (timing only)
    // Multiscale runs: keep the hat matrix (gwm_hip_mgwr_set_hat_matrix); last {tr S, tr S^T S}
    bool mgwr_hat = false;
    // problem buffers from the stream-ordered allocator on this context's
    // stream (server-owned contexts: no device-wide synchronisation at free)
    bool async_mem = false;
    // online autotuning of single-problem sweep shapes (gwm_hip_set_autotune)
    bool autotune = true;
    // a launch shape forced for every single-problem sweep (0..3, see the
    // autotuner), or -1
    int force_shape = -1;
    hipEvent_t tev0 = nullptr, tev1 = nullptr;
    double mgwr_shat[2] = {0.0, 0.0};
    std::map<std::string, gwm::detail::KernelStat> stats;
    long long solves = 0;
    // Squared great-circle distances of every pair (launch_d2_table), read by
    // the mixed sweeps of one search / fit / backfit call instead of forming
    // the distances again in every sweep.  Built at the first such sweep of a
    // call (d2_scope > 0) when it fits d2_cap and the free device memory, and
    // released when the outermost call returns (gwm_hip_set_distance_table).
    float* d2tab = nullptr;
    size_t d2tab_bytes = 0;
    bool d2_valid = false;
    int d2_scope = 0;
    size_t d2_cap = size_t(96) << 30;
    // FP64 distance table of the certification sweeps of a mixed-precision
    // fixed-bandwidth great-circle search (launch_contract_f64_table): built
    // at the search's first FP64 CV / AICc sweep for the widest bandwidth the
    // search may still evaluate (cert_hi, the bracket's upper end), released
    // when the search returns
    struct CertTable {
        double* tab = nullptr;
        size_t bytes = 0;
        long long* off = nullptr;   // [items] first block of each (tile, segment) item
        int* cnt = nullptr;         // [items] its blocks
        size_t items_cap = 0;
        double bw = 0.0;            // the table's bandwidth
        int mode = -1, kernel = -1, nta = 0;
        bool cv_dnn = false;
        bool valid = false;
        double failed_bw = 0.0;     // a table this wide did not fit: only narrower ones are tried
    } cert;
    int cert_scope = 0;             // > 0 inside a mixed-precision fixed-bandwidth search
    double cert_hi = 0.0;
    size_t cert_cap = size_t(48) << 30;

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

// Data segments of a problem's sweeps: a function of n alone, so that every
// path (single problem, batched, mixed, FP64, FP64 recomputation) sums a focus
// point's partial sums over the same segments in the same order.  Up to 16
// segments (consecutive workgroups take consecutive segments, so each of the
// 8 XCDs streams two slices of the table), each 64..8192 points.
inline void data_segments(int n, int& nseg, int& seg_len) {
    seg_len = ((n + 15) / 16 + 63) / 64 * 64;
    seg_len = seg_len < 64 ? 64 : (seg_len > 8192 ? 8192 : seg_len);
    nseg = (n + seg_len - 1) / seg_len;
}
// Column-major n x cols arrays from the caller's order to the context's
// sorted order and back.
void to_sorted(const gwm_hip_ctx* c, const double* src, double* dst, int cols);
void to_caller(const gwm_hip_ctx* c, const double* src, double* dst, int cols);
int nta_of(int k);
int ntb_of(int k);
double aicc(double rss, double n, double trS);
// GWRBasic::CalcDiagnostic (src/gwmodelpp/GWRBasic.cpp:21-34): diag[7].
void calc_diagnostic(int n, int k, const double* X, const double* y, const double* betas, const double* shat,
                     double* diag);
// f(x, exact, &value, &bound): bound receives the absolute error bound of a
// mixed-precision value (crit_error_bound), ignored for exact values.
typedef std::function<int(double, bool, double*, double*)> CritFn;
// Error bound of a mixed-precision criterion value (absolute); see gwm_api.cpp.
double crit_error_bound(int crit, double value, double rss, double trS, double q, int n, double cond_max, double limit,
                        double tol_rel);
// report(x, value) after each evaluation golden() uses (value DBL_MAX: failed)
typedef std::function<void(double, double)> EvalReport;
// upper_out (optional) holds the bracket's upper end at every evaluation:
// no later candidate lies above it.
int golden(bool adaptBw, double lower, double upper, const CritFn& f, bool certify, double tol_rel, double* xopt_out,
           std::map<double, double>& log, const EvalReport& report = EvalReport(), double* upper_out = nullptr);
void write_trace(const std::map<double, double>& log, double* tx, double* tf, int cap, int* len);
int max_distance_impl(gwm_hip_ctx* c, double* out);
// c->dnn, computed on the context stream on first use.
int ensure_dnn(gwm_hip_ctx* c);
// Make distance d the active one (building its point tables and boxes on
// first use).
int use_distance(gwm_hip_ctx* c, int d);
// ITelegram::print (level: ITelegram::LogLevel) and progress (total > 0:
// current of total; total = 0: current is a fraction), when a telegram is set.
void tele_print(gwm_hip_ctx* c, int level, const std::string& message);
void tele_progress(gwm_hip_ctx* c, double current, double total);
// IBandwidthSelectable::infoBandwidthCriterion (IBandwidthSelectable.h:30-49)
std::string info_bandwidth_criterion(bool adaptive, double bw, double value);
std::string info_bandwidth_criterion(bool adaptive);
// Looks up the device address of the allocation-free entry point's scratch
// array on `device` (the current device) once per process (gwm_noalloc.hip).
int noalloc_init(int device);


}  // namespace detail
}  // namespace gwm
