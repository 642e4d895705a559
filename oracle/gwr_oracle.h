/*
 * CPU FP64 oracle for the GWR hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A dependency-free restatement of libgwmodel's Serial/OpenMP GWR path
 * (GWRBasic, GWRMultiscale, BandwidthSelector, CRSDistance, BandwidthWeight).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker / CPU baseline. The product library never
 * links or calls it.
 *
 * Matrices are column-major FP64 (n rows), like Armadillo's arma::mat.
 * Every function cites the reference file:line it restates.
 */
#ifndef GWR_ORACLE_H
#define GWR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Kernel ids follow BandwidthWeight::KernelFunctionType
 * (include/gwmodelpp/spatialweight/BandwidthWeight.h:46-53). */
enum { ORC_GAUSSIAN = 0, ORC_EXPONENTIAL = 1, ORC_BISQUARE = 2, ORC_TRICUBE = 3, ORC_BOXCAR = 4 };
/* Criterion ids follow GWRBasic::BandwidthSelectionCriterionType { AIC, CV }. */
enum { ORC_CRIT_AIC = 0, ORC_CRIT_CV = 1 };

double orc_spgcdist(double lon1, double lon2, double lat1, double lat2);
/* d(focus, j) for all j (CRSDistance::distance, CRSDistance.cpp:122-131). */
int orc_distance_vec(int n, const double* coords, int geographic, int focus, double* out);
/* BandwidthWeight::weight(vec) (BandwidthWeight.cpp:35-61). */
int orc_weight_vec(int n, const double* dist, int kernel, int adaptive, double bw, double* out);
/* Threshold distance that an adaptive bandwidth resolves to, per focus. */
int orc_adaptive_threshold(int n, const double* coords, int geographic, double bw, int focus, double* out);
/* CRSDistance::maxDistance (CRSDistance.cpp:133-143). */
double orc_max_distance(int n, const double* coords, int geographic, int nthreads);

/* Criterion at one bandwidth (GWRBasic.cpp:440-535). Returns the value, or
 * DBL_MAX on a non-SPD local design / non-finite result. If focus_count > 0
 * only focus points [0, focus_count) are visited (timing subsets only). */
double orc_gwr_criterion(int n, int k, const double* coords, const double* X, const double* y,
                         int geographic, int kernel, int adaptive, double bw, int criterion,
                         int nthreads, int focus_count);

/* GWRBasic::fitOmp (GWRBasic.cpp:386-438). Outputs: betas n*k, betasSE n*k
 * (raw row sums of C_i.^2, before the sigma scaling done in fit()), shat[2],
 * qDiag n, S n*n when S != NULL. Returns 0 on success, 1 on non-SPD.
 * focus_count as above. */
int orc_gwr_fit_core(int n, int k, const double* coords, const double* X, const double* y,
                     int geographic, int kernel, int adaptive, double bw, int nthreads,
                     double* betas, double* betasSE, double* shat, double* qDiag, double* S,
                     int focus_count);

/* GWRBasic::predictSerial (GWRBasic.cpp:167-191): betas (m*k) at the m
 * locations (m*2, column-major) from the data (coords, X, y); adaptive
 * bandwidths count data points nearest each location.  Returns 0 ok, 1 when
 * a local design is not SPD (those rows NaN; the reference throws). */
int orc_gwr_predict(int n, int k, const double* coords, const double* X, const double* y, int geographic,
                    int m, const double* locations, int kernel, int adaptive, double bw, int nthreads,
                    double* betas);

/* GWRBasic::indepVarsSelectionCriterionSerial (GWRBasic.cpp:309-346): AICc of
 * the regressions on columns {0 if has_intercept} + vars with unit weights
 * (every location solves the same global OLS, :319); DBL_MAX when the design
 * is not SPD or the value is not finite (GWRBasic.cpp:893). */
double orc_indep_vars_criterion(int n, int k, const double* X, const double* y, int has_intercept,
                                const int* vars, int nvars);

/* VariableForwardSelector::optimize / sort / select
 * (VariableForwardSelector.cpp:9-77) over the columns of X after the
 * intercept.  selected (cap k) gets the chosen columns; the criterion list
 * (m(m+1)/2 models, m candidate columns) goes to model_vars (row q: k entries,
 * -1 padded), model_nvars and model_crit in the reference's sorted order. */
int orc_indep_vars_select(int n, int k, const double* X, const double* y, int has_intercept, double threshold,
                          int* selected, int* nselected, int* model_vars, int* model_nvars, double* model_crit,
                          int* nmodels);

/* GWRBasic::fit() orchestration (GWRBasic.cpp:36-128): optional golden-section
 * bandwidth selection, fit, diagnostics, SE scaling.
 * diag[7] = RSS, AIC, AICc, ENP, EDF, R2, R2adj (RegressionDiagnostic order).
 * trace_x / trace_f (capacity trace_cap) receive the golden-section criterion
 * list sorted by bandwidth (BandwidthSelector::bandwidthCriterion).
 * Returns 0 ok, 1 non-SPD, 2 invalid initial values. */
int orc_gwr_run(int n, int k, const double* coords, const double* X, const double* y,
                int geographic, int kernel, int adaptive, double bw,
                int select_bw, int criterion, int has_lower, double lower, int has_upper, double upper,
                int nthreads,
                double* betas, double* betasSE, double* shat, double* qDiag, double* S,
                double* diag, double* bw_out, double* trace_x, double* trace_f, int trace_cap,
                int* trace_len);

/* GWRMultiscale::fit() (GWRMultiscale.cpp:68-317) restated.
 * Per-variable arrays have k entries, geographic included (one SpatialWeight,
 * hence one distance, per variable; the initial all-variable search and fitAll
 * use variable 0's, GWRMultiscale.cpp:1476-1483). init_type: 0 Null, 1 Initial, 2 Specified.
 * approach: criterion ids as above. crit_type: 0 CVR, 1 dCVR.
 * Outputs: betas n*k, bws_out k, diag[7] (AIC left 0 as the reference leaves it unset),
 * rss0_out, iterations_out. Returns 0 ok, 1 non-SPD, 2 invalid initial values. */
int orc_mgwr_run(int n, int k, const double* coords, const double* X, const double* y,
                 const int* geographic, const int* kernels, const int* adaptive, const double* bws,
                 const int* init_type, const int* approach, const int* centered,
                 const double* bw_threshold, int has_intercept, int has_hat,
                 int crit_type, double crit_threshold, int max_iter, int retry_times,
                 int adaptive_lower, int has_lower, double lower, int has_upper, double upper,
                 int nthreads,
                 double* betas, double* bws_out, double* diag, double* rss0_out,
                 int* iterations_out);

#ifdef __cplusplus
}
#endif

#endif
