/*
 * gwmodel_hip.h -- C ABI of the MI355X (gfx950) GWR back-end.
 *
 * This is the boundary the HIP back-end sits behind.  Each entry point
 * replaces one of libgwmodel's CUDA back-end member functions, which the
 * reference dispatches through member-function pointers selected by
 * setParallelType (src/gwmodelpp/GWRBasic.cpp:904-975,
 * src/gwmodelpp/GWRMultiscale.cpp:1396-1474).  Plain C: opaque handle,
 * caller-owned column-major FP64 buffers (copied in), integer status codes,
 * no exceptions across the ABI and no device reset on error (the reference's
 * checkCudaErrors resets the device, include/gwmodelpp/utils/CudaUtils.h:53-66).
 *
 * Matrices are column-major FP64 with n rows, exactly arma::mat's memory.
 * Kernel ids: BandwidthWeight::KernelFunctionType
 *             (include/gwmodelpp/spatialweight/BandwidthWeight.h:46-53).
 * Criterion ids: GWRBasic::BandwidthSelectionCriterionType (GWRBasic.h:41-45).
 */
#ifndef GWMODEL_HIP_H
#define GWMODEL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define GWM_HIP_ABI_VERSION 1

typedef struct gwm_hip_ctx gwm_hip_ctx;

/* status codes */
enum {
    GWM_OK = 0,
    GWM_ERR_NOT_SPD = 1,          /* a local design is not SPD (inv_sympd failure) */
    GWM_ERR_CANCELLED = 2,        /* the stop callback asked to stop (Status::Terminated) */
    GWM_ERR_HIP = 3,              /* HIP runtime error; see gwm_hip_last_error */
    GWM_ERR_OOM = 4,              /* device allocation failed */
    GWM_ERR_INVALID = 5,          /* invalid argument or state */
    GWM_ERR_NO_DEVICE = 6,        /* no usable gfx950 device */
    GWM_ERR_INVALID_INITIAL = 7   /* golden section: both initial values invalid */
};

/* kernels: BandwidthWeight::KernelFunctionType */
enum { GWM_GAUSSIAN = 0, GWM_EXPONENTIAL = 1, GWM_BISQUARE = 2, GWM_TRICUBE = 3, GWM_BOXCAR = 4 };
/* distances: CRSDistance(false) / CRSDistance(true) */
enum { GWM_EUCLIDEAN = 0, GWM_GREAT_CIRCLE = 1 };
/* criteria: GWRBasic::BandwidthSelectionCriterionType */
enum { GWM_CRIT_AIC = 0, GWM_CRIT_CV = 1 };
/* arithmetic of the sweep: FP64 everywhere, or FP32 distances/weights with an
 * FP16 hi+lo split contraction on the matrix cores (FP32 accumulate, FP64
 * partial sums, FP64 solves). */
enum { GWM_FP64 = 0, GWM_MIXED = 1 };

/* Number of visible devices (hipGetDeviceCount). */
int gwm_hip_device_count(int* count);

/* Context: one device, one stream, device-resident problem data.
 * Replaces GWRBasic::fit's cublasCreate + SpatialWeight::prepareCuda(gpuId)
 * (src/gwmodelpp/GWRBasic.cpp:41-47) and IParallelCudaEnabled::setGPUId. */
int gwm_hip_create(int device, gwm_hip_ctx** out);
void gwm_hip_destroy(gwm_hip_ctx* ctx);
const char* gwm_hip_last_error(const gwm_hip_ctx* ctx);

/* Cancellation: polled between sweeps (ITelegram::stop, include/gwmodelpp/Logger.h:76-81). */
int gwm_hip_set_stop_callback(gwm_hip_ctx* ctx, int (*stop)(void* user), void* user);

/* Messages and progress: the ITelegram interface (include/gwmodelpp/Logger.h:31-83)
 * that Algorithm::mTelegram serves (GWM_LOG_INFO / GWM_LOG_ERROR /
 * GWM_LOG_PROGRESS / GWM_LOG_PROGRESS_PERCENT, include/gwmodelpp/Algorithm.h:21-100).
 * print(user, level, message): level is ITelegram::LogLevel (3 error, 6 info);
 * messages carry the reference's tags: "#stage ", "#bandwidth-criterion ",
 * "#variable-criterion ", "#initial-bandwidth ", "#back-fitting ".
 * progress(user, current, total): total > 0 for "current of total"
 * (variable selection, the end of a fit: one report per fit, not one per
 * focus point, since the fit is one sweep); total = 0 for a fraction in
 * current (criterion evaluations: exp(-|last - value|), back-fitting
 * iterations: exp(-|criterion - threshold|)).  Calls come from the calling
 * thread, inside the entry point.  NULL callbacks switch reporting off. */
int gwm_hip_set_telegram(gwm_hip_ctx* ctx, void (*print)(void* user, int level, const char* message),
                         void (*progress)(void* user, double current, double total), void* user);

/* Problem data (copied to the device).  coords n x 2 (x,y or lon,lat in
 * degrees); X n x k; y n.  1 <= k <= 16: the reference takes any k; here the
 * per-focus k x k solve lives in registers and the product table in at most
 * 16 MFMA column tiles, so k > 16 is rejected with GWM_ERR_INVALID.  Replaces the H2D uploads of
 * CRSDistance::prepareCuda (src/gwmodelpp/spatialweight/CRSDistance.cpp:158-168)
 * and of fitCuda / bandwidthSizeCriterion*Cuda (GWRBasic.cpp:585-600, 709-720). */
int gwm_hip_set_data(gwm_hip_ctx* ctx, int n, int k, const double* coords, const double* X, const double* y,
                     int distance);

/* Criterion at one bandwidth: replaces GWRBasic::bandwidthSizeCriterionCVCuda /
 * bandwidthSizeCriterionAICCuda (src/gwmodelpp/GWRBasic.cpp:709-836) and their
 * Serial/OpenMP twins (:233-307, :440-535).  *value = DBL_MAX with status
 * GWM_ERR_NOT_SPD when a local design is not SPD or the value is not finite. */
int gwm_hip_criterion(gwm_hip_ctx* ctx, int kernel, int adaptive, double bw, int criterion, int precision,
                      double* value);

/* Fit at one bandwidth: replaces GWRBasic::fitCuda (src/gwmodelpp/GWRBasic.cpp:585-663)
 * / fitOmp (:386-438).  Outputs (nullable): betas n x k, betasSE_raw n x k (row
 * sums of C_i.^2 before the sigma scaling of fit()), shat[2] = {tr S, tr S^T S},
 * qDiag n, S n x n (only when S != NULL; the reference stores S when n < 8192,
 * GWRBasic.h:708).  Returns GWM_ERR_NOT_SPD like fitOmp's rethrow. */
int gwm_hip_fit(gwm_hip_ctx* ctx, int kernel, int adaptive, double bw, int precision, double* betas,
                double* betasSE_raw, double* shat, double* qDiag, double* S);

/* Betas at m prediction locations: replaces GWRBasic::predictSerial /
 * predictOmp / predictCuda (src/gwmodelpp/GWRBasic.cpp:167-191, :349-384,
 * :665-707) behind GWRBasic::predict (:130-156).  locations is m x 2
 * column-major in the data's coordinate system; betas receives m x k
 * column-major.  Adaptive bandwidths count the bw nearest data points of each
 * location.  Returns GWM_ERR_NOT_SPD (betas NaN at those locations) where
 * predictSerial's inv_sympd throws. */
int gwm_hip_predict(gwm_hip_ctx* ctx, int m, const double* locations, int kernel, int adaptive, double bw,
                    int precision, double* betas);

/* Variable-selection criterion of one column set: replaces
 * GWRBasic::indepVarsSelectionCriterionSerial / Omp / Cuda
 * (src/gwmodelpp/GWRBasic.cpp:309-346, :537-580, :838-900), the
 * IndepVarsSelectCriterionCalculator of GWRBasic.h:53.  Columns are
 * {0 if has_intercept} + vars (VariableForwardSelector::index2uvec,
 * VariableForwardSelector.h:26-35); the regression has unit weights (:319), so
 * *value is the AICc of the global OLS with tr(S) = sum_i x_i^T (X^T X)^-1 x_i,
 * or DBL_MAX when X^T X is not SPD or the value is not finite. */
int gwm_hip_indep_vars_criterion(gwm_hip_ctx* ctx, int has_intercept, const int* vars, int nvars, double* value);

/* Forward selection of independent variables: replaces
 * VariableForwardSelector::optimize / sort / select
 * (src/gwmodelpp/VariableForwardSelector.cpp:9-77) as GWRBasic::fit runs it
 * (GWRBasic.cpp:50-72) over the columns after the intercept.  selected (cap k)
 * receives the chosen columns, *nselected their count.  The criterion list
 * (GWRBasic::indepVarsSelectionCriterionList) has m(m+1)/2 models for m
 * candidate columns: model q has model_nvars[q] columns in
 * model_vars[q*k .. q*k+k) (-1 padded) and criterion model_crit[q]; at most
 * cap_models are written, *nmodels receives the full count. */
int gwm_hip_indep_vars_select(gwm_hip_ctx* ctx, int has_intercept, double threshold, int* selected, int* nselected,
                              int* model_vars, int* model_nvars, double* model_crit, int cap_models, int* nmodels);

/* Allocation-free criterion.  No counterpart in the reference: a variant of
 * gwm_hip_criterion for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The inputs are the caller's device arrays in the caller's
 * order: coords n x 2, X n x k column-major, y n (no gwm_hip_set_data, no
 * spatial sort, so no pruning); FP64 arithmetic only.  Every other device
 * buffer lives in one static array of 32 MiB per device that the library
 * holds until the process ends; problems whose tables do not fit (about
 * n <= 12,000 at k = 8) are rejected with GWM_ERR_INVALID and a message
 * naming the size.  Calls on one device are serialised by the library
 * (they share its array); the call synchronises the context stream.
 * Results agree with gwm_hip_criterion on the same data within FP64
 * summation-order differences.  Status and *value as gwm_hip_criterion. */
int gwm_hip_criterion_noalloc(gwm_hip_ctx* ctx, int n, int k, const double* d_coords, const double* d_X,
                              const double* d_y, int distance, int kernel, int adaptive, double bw, int criterion,
                              double* value);

/* One distance per variable for gwm_hip_mgwr_run: replaces
 * SpatialMultiscaleAlgorithm::setSpatialWeights' distance part
 * (include/gwmodelpp/SpatialMultiscaleAlgorithm.h:66, 84-87; GWRMultiscale.cpp:1476-1483):
 * variable i's bandwidth searches and fits use distances[i]
 * (GWM_EUCLIDEAN or GWM_GREAT_CIRCLE on the same coordinates); the initial
 * all-variable search and fit use distances[0], as mInitSpatialWeight is
 * spatial weight 0.  k = 0 clears the setting: every variable then uses the
 * distance given to gwm_hip_set_data.  The run rejects a count other than
 * the data's k. */
int gwm_hip_mgwr_set_distances(gwm_hip_ctx* ctx, int k, const int* distances);

/* GWRMultiscale::setHasHatMatrix (GWRMultiscale.h:398) for gwm_hip_mgwr_run: keep
 * S0 and the per-variable operators (GWRMultiscale.cpp:163-168, 215-225,
 * 296-301) on the device, (k + 3) n^2 doubles (GWM_ERR_OOM when the device
 * lacks the memory).  The fitAll sweep then runs in FP64 whatever the
 * precision.  Off by default. */
int gwm_hip_mgwr_set_hat_matrix(gwm_hip_ctx* ctx, int enabled);
/* {tr S0, tr S0^T S0} of the last gwm_hip_mgwr_run (GWRMultiscale.cpp:176-179;
 * zeros without the hat matrix). */
int gwm_hip_mgwr_last_shat(gwm_hip_ctx* ctx, double* shat);

/* Golden-section bandwidth selection on device-resident data: replaces
 * BandwidthSelector::optimize (src/gwmodelpp/BandwidthSelector.cpp:6-81)
 * driving getCriterion.  Bounds follow GWRBasic::fit (GWRBasic.cpp:76-79):
 * lower = has_lower ? lower : (adaptive ? 20 : 0), upper = has_upper ? upper :
 * (adaptive ? n : maxDistance).  trace_* receive the criterion list sorted by
 * bandwidth (BandwidthSelector::bandwidthCriterion, :83-95). */
int gwm_hip_select_bandwidth(gwm_hip_ctx* ctx, int kernel, int adaptive, int criterion, int precision, int has_lower,
                             double lower, int has_upper, double upper, double* bw_out, double* trace_bw,
                             double* trace_value, int trace_cap, int* trace_len);

/* GWRBasic::fit() end to end (src/gwmodelpp/GWRBasic.cpp:36-103): optional
 * bandwidth selection, fit, CalcDiagnostic (:21-34) and the SE scaling.
 * diag[7] = RSS, AIC, AICc, ENP, EDF, R2, R2adj (RegressionDiagnostic order). */
int gwm_hip_gwr_run(gwm_hip_ctx* ctx, int kernel, int adaptive, double bw, int select_bw, int criterion,
                    int precision, int has_lower, double lower, int has_upper, double upper, double* betas,
                    double* betasSE, double* shat, double* qDiag, double* S, double* diag, double* bw_out,
                    double* trace_bw, double* trace_value, int trace_cap, int* trace_len);

/* CRSDistance::maxDistance (src/gwmodelpp/spatialweight/CRSDistance.cpp:133-143). */
int gwm_hip_max_distance(gwm_hip_ctx* ctx, double* out);
/* CRSDistance::distance(focus) (CRSDistance.cpp:122-131): n distances. */
int gwm_hip_distance_row(gwm_hip_ctx* ctx, int focus, double* out);
/* The squared-distance boundary of the FP64 bisquare and boxcar weights (test
 * hook, no device needed): the smallest double x whose correctly rounded
 * square root is >= bw, so that d^2 < x exactly when the reference's test
 * d < bw holds (BandwidthWeight.h:68-121); 0 for bw <= 0. */
int gwm_hip_sqrt_boundary(double bw, double* out);
/* Adaptive bandwidth resolved per focus point (BandwidthWeight.cpp:40-55): n values. */
int gwm_hip_adaptive_bandwidths(gwm_hip_ctx* ctx, double bw, double* out);

/* Multiscale single-column kernels on the context's coordinates:
 * bandwidthSizeCriterionVar{CV,AIC}Cuda (src/gwmodelpp/GWRMultiscale.cpp:1246-1392)
 * and fitVarCuda (:1062-1140).  x, y: n-vectors (host). */
int gwm_hip_var_criterion(gwm_hip_ctx* ctx, const double* x, const double* y, int kernel, int adaptive, double bw,
                          int criterion, int precision, double* value);
int gwm_hip_var_fit(gwm_hip_ctx* ctx, const double* x, const double* y, int kernel, int adaptive, double bw,
                    int precision, double* beta);

/* GWRMultiscale::fit() end to end (src/gwmodelpp/GWRMultiscale.cpp:68-197, backfitting
 * :208-317), without the hat matrix (setHasHatMatrix(false)) unless
 * gwm_hip_mgwr_set_hat_matrix enabled it.  Per-variable arrays have k entries;
 * init_type 0 Null / 1 Initial / 2 Specified; approach uses the GWM_CRIT_* ids;
 * crit_type 0 CVR / 1 dCVR.  Outputs: betas n x k, bws_out k, diag[7]
 * (CalcDiagnostic, :41-66: AIC is left 0 as the reference leaves it unset; the
 * hat-based terms use tr S = tr S^T S = 0 without the hat matrix), rss_out,
 * iterations_out. */
int gwm_hip_mgwr_run(gwm_hip_ctx* ctx, const int* kernels, const int* adaptive, const double* bws,
                     const int* init_type, const int* approach, const int* centered, const double* bw_threshold,
                     int has_intercept, int crit_type, double crit_threshold, int max_iter, int retry_times,
                     int adaptive_lower, int has_lower, double lower, int has_upper, double upper, int precision,
                     double* betas, double* bws_out, double* diag, double* rss_out, int* iterations_out);

/* Mixed precision: relative error bound used to certify the golden-section
 * decisions taken on mixed-precision criteria (default 1e-7, about 10x the
 * largest mixed-vs-FP64 criterion error measured on the test and bench workloads); a decision
 * closer than the bound is re-taken on FP64 re-evaluations.  No counterpart in
 * the reference, which has no mixed-precision path. */
int gwm_hip_set_certify_tolerance(gwm_hip_ctx* ctx, double rel_tol);
/* Mixed precision: the error of a mixed local solve grows with the
 * conditioning of its local design.  Every mixed evaluation computes, per
 * focus point, the indicator max_a A_aa / L_aa^2 (A = X^T W X, L its Cholesky
 * factor: the inverse smallest squared pivot of the Jacobi-scaled design,
 * >= 1, infinite when not SPD), and recomputes in FP64 -- the FP64 path's own
 * kernel and data segments, hence its bits -- every focus point whose
 * indicator exceeds the limit: criterion_limit in criterion evaluations
 * (default 10), fit_limit in fits (default 2).  gwm_hip_last_condition
 * returns the indicator maximum of the last mixed evaluation and the number
 * of focus points it recomputed.  No counterpart in the reference. */
int gwm_hip_set_refine_condition(gwm_hip_ctx* ctx, double criterion_limit, double fit_limit);

/* Spatial pruning (default on): the points are kept in a spatial order in 64-
 * point chunks with bounding boxes, and a sweep skips every chunk whose
 * points are all farther from a workgroup's focus points than the distance
 * at which the kernel weight is exactly zero in the sweep's arithmetic
 * (compact kernels: the bandwidth; Gaussian / exponential: 8.9 / 40
 * bandwidths in FP64, whose weights below e^-40 = 4.2e-18 are taken as zero,
 * and FP16 underflow of the split weight in mixed precision).  Skipped
 * chunks contribute exactly zero, so results are bit-identical with pruning
 * on or off.  No counterpart in the reference, which sums over all n. */
int gwm_hip_set_pruning(gwm_hip_ctx* ctx, int enabled);
int gwm_hip_last_condition(gwm_hip_ctx* ctx, double* cond, long long* refined);
/* Absolute error bound that the certified golden section assumes for the
 * last mixed-precision criterion value of this context (0 after an FP64
 * evaluation): the largest of tol |f| (3 tol |f| for CV), the rounding term
 * r = 8 * 2^-22 * kappa * sqrt(sum_i (res_i yhat_i)^2) (CV: r; AICc: n r /
 * RSS; kappa the indicator maximum capped at the recomputation limit) and,
 * for AICc, tol * S * tr(S) with S = n (2n - 2) / (n - 2 - tr(S))^2. */
int gwm_hip_last_error_bound(gwm_hip_ctx* ctx, double* bound);

/* Online autotuning of the single-problem sweeps (no counterpart in the
 * reference; on by default).  Per class (distance, sweep mode, precision,
 * column tiles, n to a power of two) the first sweeps rotate through the
 * launch shapes that leave results bit-identical -- one or two focus tiles
 * per wave, tiles visited in order or by a coprime stride -- each timed twice
 * with HIP events; the fastest is used from then on, process-wide.  Larger
 * pruned sweeps keep the default shape and run their (tile, segment)
 * workgroups longest first (a counting pass over the chunk boxes); disabling
 * autotuning also restores the plain order.  Batched server launches keep
 * the default shape and order. */
/* enabled: 0 off (default shape, plain order), 1 on; 2..5 force shape 0..3 on
 * every sweep -- 0 and 1: the default focus tiles per wave, by the coprime
 * stride / in order; 2 and 3: one focus tile per wave, likewise -- so that
 * tests can compare every shape bit for bit. */
int gwm_hip_set_autotune(gwm_hip_ctx* ctx, int enabled);
/* The tuned classes, one text line each (shape chosen, best time of every
 * shape in ms), NUL-terminated and truncated to cap bytes. */
int gwm_hip_autotune_report(char* buf, int cap);

/* Distance table of great-circle problems (no counterpart in the reference,
 * whose every criterion evaluation forms each pair's distance again,
 * GWRBasic.cpp:440-535).  Within one gwm_hip_select_bandwidth,
 * gwm_hip_gwr_run or gwm_hip_mgwr_run call, the first mixed-precision sweep
 * of great-circle data stores the FP32 squared distances of all pairs in
 * device memory (4 n^2 bytes: 40 GB at n = 100,000), and every later mixed
 * sweep of the call reads them instead of forming them: the same values, so
 * the same results.  The table is released when the call returns.  It is
 * built only when it needs at most max_bytes (default 96 GiB) and the device
 * keeps 8 GiB free beside it; max_bytes = 0 turns it off.  FP64 sweeps never
 * use it. */
int gwm_hip_set_distance_table(gwm_hip_ctx* ctx, double max_bytes);

/* FP64 distance table of certified searches (no counterpart in the
 * reference).  In a mixed-precision golden-section search over fixed
 * bandwidths of great-circle data, the first FP64 (certification) CV or AICc
 * sweep stores the FP64 squared distance of every pair that the bracket's
 * upper end reaches, exactly as the FP64 sweep forms it, and the search's
 * later FP64 sweeps read them instead of forming them: the same values, so
 * the same criteria and iterates.  The table is released when the search
 * returns.  It is built only when it needs at most max_bytes (default 48 GiB)
 * and the device keeps 8 GiB free beside it; max_bytes = 0 turns it off. */
int gwm_hip_set_certify_table(gwm_hip_ctx* ctx, double max_bytes);

/* Neighbour table of adaptive bandwidths (no counterpart in the reference,
 * whose BandwidthWeight::weight sorts every focus point's n distances again
 * for each bandwidth, BandwidthWeight.cpp:40-55).  At the second adaptive
 * select of a problem with floor(bw) < K, the context stores, per point, its
 * K smallest selection keys in ascending order (K: the largest power of two
 * up to 8192 below n whose n K 8 bytes fit max_bytes; 3.3 GB at n = 50,000),
 * and resolves every later adaptive bandwidth with floor(bw) < K from it:
 * the very order statistics the select returns, so the same thresholds.  The
 * table stays until the data change.  It is built only when the device keeps
 * 8 GiB free beside it; max_bytes = 0 turns it off (default 16 GiB), and a
 * cap below a built table's size releases it. */
int gwm_hip_set_neighbour_table(gwm_hip_ctx* ctx, double max_bytes);

/* Timing of the device kernels with HIP events on the context stream
 * (enabled by gwm_hip_set_timing).  name: "contract_<fp64|mixed>_<cv|plain|
 * hat|qdiag>" for the sweeps, "select" (adaptive bandwidths by the select),
 * "select_lookup" (by the neighbour table), "nbr_table" (its build), "solve",
 * "d2_table" (the distance table), "f64_table" (the certification table). */
int gwm_hip_set_timing(gwm_hip_ctx* ctx, int enabled);
int gwm_hip_kernel_time(gwm_hip_ctx* ctx, const char* name, long long* launches, double* total_ms);
/* Focus x data pairs the named sweeps actually processed while timing was on
 * (spatial pruning skips the rest); padding lanes and points count. */
int gwm_hip_kernel_pairs(gwm_hip_ctx* ctx, const char* name, double* pairs);
int gwm_hip_reset_timers(gwm_hip_ctx* ctx);
/* Local solves (focus x bandwidth pairs) evaluated since the last reset. */
int gwm_hip_solve_count(gwm_hip_ctx* ctx, long long* solves);

/* ------------------------------------------------------------------------
 * Batched / continuous-batching execution.  No counterpart in the reference,
 * which runs one problem per fit() and has no batching or serving layer.
 *
 * A server owns one device stream and runs many GWRBasic requests at once: a
 * fit at a given bandwidth, or a golden-section bandwidth search followed by
 * the fit (select_bw).  Requests are admitted first come, first served while
 * max_inflight slots and the device memory budget allow (a request's footprint
 * is O(n k^2), never n^2); the criterion and fit evaluations of all in-flight
 * requests are executed together in rounds of batched launches.  A round has a
 * work budget of about round_ms of device time (default 25), shared shortest-
 * remaining-work first, so a large evaluation advances in chunks of focus
 * points over several rounds while small ones complete in one.  A request's
 * result does not depend on which other requests share its rounds.  Output
 * pointers (nullable; betas and betasSE n x k, shat 2, qDiag n, diag 7, bw_out
 * 1) must stay valid until gwm_hip_server_wait returns for the request.
 * hbm_budget_bytes <= 0 means 80% of the free device memory at creation.
 */
typedef struct gwm_hip_server gwm_hip_server;
int gwm_hip_server_create(int device, int precision, double hbm_budget_bytes, int max_inflight,
                          gwm_hip_server** out);
int gwm_hip_server_set_certify_tolerance(gwm_hip_server* server, double rel_tol);
int gwm_hip_server_set_refine_condition(gwm_hip_server* server, double criterion_limit, double fit_limit);
/* Round budget in ms of device time; <= 0 runs every posted evaluation whole
 * in each round. */
int gwm_hip_server_set_round_ms(gwm_hip_server* server, double round_ms);
/* Bytes of one request's partial sums (default 256 MiB, at least 1): a round
 * computes at most as many of a request's focus points as their partial sums
 * fit (whole chunks of 128, at least one chunk), and a request holds only that
 * slice, so its device footprint stays O(n k^2) with a bounded partial-sum
 * term.  Applies to requests submitted afterwards; results do not depend on
 * it. */
int gwm_hip_server_set_partial_budget(gwm_hip_server* server, double bytes);
/* The server's tuning of batched sweep launches (no counterpart in the
 * reference): a class (distance, sweep mode, precision, column tiles) whose
 * default is two focus tiles per wave runs its first three launches of at
 * least 512 workgroups with two focus tiles per wave and three with one,
 * timed with HIP events per launch and normalised by the
 * pairs the launch swept (counted on the device), and keeps the faster
 * shape; results do not
 * depend on it.  One text line per tuned class (chosen shape, microseconds
 * per million pairs of each shape, rounds timed), NUL-terminated and
 * truncated to cap bytes. */
int gwm_hip_server_autotune_report(gwm_hip_server* server, char* buf, int cap);
int gwm_hip_server_submit(gwm_hip_server* server, int n, int k, const double* coords, const double* X,
                          const double* y, int distance, int kernel, int adaptive, double bw, int select_bw,
                          int criterion, int has_lower, double lower, int has_upper, double upper, double* betas,
                          double* betasSE, double* shat, double* qDiag, double* diag, double* bw_out,
                          long long* request_id);
/* Blocks until the request is done; status is the request's GWM_* code,
 * latency_ms runs from submission to completion, solves counts its local
 * solves.  The request id is released. */
int gwm_hip_server_wait(gwm_hip_server* server, long long request_id, int* status, double* latency_ms,
                        long long* solves);
int gwm_hip_server_stats(gwm_hip_server* server, long long* rounds, long long* solves, double* busy_ms,
                         double* peak_bytes, int* peak_inflight);
const char* gwm_hip_server_last_error(gwm_hip_server* server);
void gwm_hip_server_destroy(gwm_hip_server* server);

#ifdef __cplusplus
}
#endif

#endif /* GWMODEL_HIP_H */
