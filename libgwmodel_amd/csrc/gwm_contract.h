// Launch interface of the device kernels (host side of the HIP library).
#pragma once
#include <hip/hip_runtime.h>
#include "gwm_device.h"

namespace gwm {

// Contraction modes.
//   kModeCV    : columns [pairs | x*y] with w, self weight zeroed (leave-one-out CV)
//   kModeHat   : columns [pairs | x*y] with w and [pairs] with w^2 (X^T W^2 X)
//   kModeQDiag : columns of G = vech(g g^T) with w^2, bandwidth of the streamed index
//   kModePlain : columns [pairs | x*y] with w (AICc criterion, multiscale fits)
enum ContractMode { kModeCV = 0, kModeHat = 1, kModeQDiag = 2, kModePlain = 3 };

// One problem's sweep.  The FP64 kernel reads Z, the mixed kernel the FP16
// planes Zh/Zl with their per-column unscale factors.  nta/ntb are the
// problem's tile counts; a launch only groups jobs of equal counts.
struct SweepJob {
    Points P;
    const double* Z;         // row-major [rows >= ceil(n/64)*64][nta*16], zero padded
    const _Float16* Zh;      // [rows/32][nta*16][32] hi part
    const _Float16* Zl;      // same, lo part scaled by 2^11
    const double* colscale;  // per column unscale (power of two)
    const double* bw_pt;     // per-point bandwidth (adaptive), or nullptr for a fixed bw
    double bw;
    int kernel;
    int n;
    int n_pad;               // focus padding of the output, multiple of 128
    int seg_len;             // data points per segment, multiple of 64
    int nseg;
    int nta, ntb;
    double* out;             // [nseg][(nta+ntb)*16][n_pad]
    // optional focus list: output column t holds focus focus_idx[t],
    // t < nfocus (n_pad then pads nfocus); null = every point in order
    const int* focus_idx = nullptr;
    int nfocus = 0;
    // optional focus table (predict): when Q.n > 0 the focus points are the
    // points of Q (n_pad pads Q.n, focus_idx indexes Q, bw_pt is per point
    // of Q) and no focus is excluded from its own sums
    Points Q{};
    // pruning (see gwm_device.h): boxes of the 64-point data chunks (null: no
    // pruning) and, for adaptive qDiag sweeps, each chunk's largest bandwidth
    const double* boxes = nullptr;
    const double* chunk_bw = nullptr;
    // optional count of the focus x data pairs actually swept (pruning skips
    // the rest): one atomic add per workgroup and processed chunk
    unsigned long long* pairs = nullptr;
    // FP64 sweeps whose focus point's own weight is not the largest (CV:
    // zeroed; predict: the focus is not a data point): per focus point the
    // distance to its nearest other (data) point, indexed like bw_pt; the
    // exponent cut of the Gaussian and exponential kernels is taken relative
    // to that point's weight (make_kconst64_nn).  Null: relative to 1.
    const double* dnn = nullptr;
    // mixed great-circle sweeps of the data points against themselves: the
    // FP32 squared distances of every (focus, data) pair, precomputed by
    // launch_d2_table (layout there); the sweep is then launched with the
    // distance id kGcTable and reads them instead of forming them
    const float* d2tab = nullptr;
    int d2_ndb = 0;          // 32-point data blocks per focus group (rows / 32)
    // out holds the focus points (or list entries) from out_off on: column t
    // of the output is focus out_off + t (a server round's slice)
    int out_off = 0;
    // FP64 great-circle sweeps of one mixed-precision search at fixed
    // bandwidths no wider than tab_bw (its certification evaluations): the
    // FP64 squared distances of the pairs that bandwidth reaches, as the
    // sweep forms them, written once (launch_contract_f64_table) and read
    // instead of formed.  Item (tile, seg) holds tab_cnt[item] blocks of
    // kWaves x 16 x R x 64 doubles from block tab_off[item] on, one per chunk
    // within reach of tab_bw, in chunk order.
    double* f64tab = nullptr;
    const long long* tab_off = nullptr;
    int* tab_cnt = nullptr;
    double tab_bw = 0.0;
};

// Distance id of a mixed sweep that reads the precomputed squared great-circle
// distances (SweepJob::d2tab); pruning still uses the great-circle boxes.
constexpr int kGcTable = 2;

// Launch of one sweep: either a single job (items == nullptr; workgroup b
// covers tile b / nseg and segment b % nseg) or a batch (items: triples
// (job, tile, segment), one per workgroup, jobs[] in device memory).
struct SweepLaunch {
    SweepJob job0;
    const SweepJob* jobs;
    const int* items;
    int nitems;
    int tile_stride = 1;   // single job: workgroup order t visits tile (t * stride) mod tiles (set by the launcher)
    // launch shape of a single job, chosen by the autotuner (results do not
    // depend on it: every focus point sums its data in the same order):
    // shape_r = 1 asks for one focus tile per wave where the default is two;
    // linear_order = 1 visits the tiles in order instead of by the stride
    int shape_r = 0;
    int linear_order = 0;
    // optional work order of a single job (launch_sweep_plan): workgroup b
    // covers item order[b] = tile * nseg + segment; norder = tiles * nseg
    // must equal the launch's grid, or the launch fails
    const int* order = nullptr;
    int norder = 0;
};

// A stride coprime to `tiles`, near tiles / golden ratio: visiting tiles in
// the order (t * stride) mod tiles spreads consecutive workgroups over the
// domain, so that with spatially ordered points the workgroups that pruning
// leaves busy are spread over all XCDs instead of the few that hold the data
// segments next to the current tiles.
int tile_stride(int tiles);

// Single-column product table of a multiscale step (GWRMultiscale.cpp:
// 420-482, 560-636): Z[j][0] = x_j^2, Z[j][1] = x_j y_j, the other 14 of its
// 16 columns and the rows n .. rows_pad-1 zero.
hipError_t launch_var_table(const double* x, const double* y, int n, int rows_pad, double* Z, hipStream_t s);

// FP16 hi/lo planes ([rows/32][ldz][32]) and per-column unscale factors of a
// row-major FP64 table; colmax is ldz words of scratch.
hipError_t launch_split_fp16(const double* Z, int rows, int ldz, unsigned long long* colmax, _Float16* Zh,
                             _Float16* Zl, double* unscale, hipStream_t s);
// nta/ntb select the compile-time tile counts (those of every job); grid size is
// nitems (batch) or tiles * nseg (single job).
hipError_t launch_contract_f64(int dist, int mode, int nta, int ntb, const SweepLaunch& a, hipStream_t s);
// FP64 single-column sweep (nta = 1, ntb = 0, CV or plain mode, every point a
// focus point): out = [nseg][2][n_pad], the columns x^2 and x y of the
// single-column table, summed on the VALU (gwm_contract.hip).
hipError_t launch_contract_var_f64(int dist, int mode, const SweepJob& j, hipStream_t s);
// The FP64 distance table of a single-job great-circle CV / AICc sweep at a
// fixed bandwidth (SweepJob::f64tab), in the default launch shape: stage 3
// counts each item's chunks within reach of job0.bw into tab_cnt, stage 1
// writes their blocks (tab_off from the counts' prefix sums), stage 2 is the
// sweep at job0.bw <= tab_bw reading them.  Returns the doubles per block
// in *block (stage 3).
hipError_t launch_contract_f64_table(int stage, int mode, int nta, const SweepLaunch& a, hipStream_t s,
                                     long long* block = nullptr);
hipError_t launch_contract_mixed(int dist, int mode, int nta, int ntb, const SweepLaunch& a, hipStream_t s);
// the R (focus tiles per wave) each kernel uses for an (nta, ntb, mode,
// distance) instance; launchers and item tables must agree on it
int contract_r(int nta, int ntb, int precision_mixed, int mode, int dist);

// Squared great-circle distances (km^2, FP32) of every pair of the n points of
// P, exactly as the mixed sweep forms them: focus i < n_pad (rows past n repeat
// point n - 1), data j < rows (a multiple of 64; columns past n repeat point
// n - 1).  Layout: blocks of 16 focus x 32 data points, block (i / 16, j / 32)
// at ((i / 16) * (rows / 32) + j / 32) * 512 floats; inside a block the entry
// of focus 16 g + c and data 32 b + 8 q + e sits at (e / 4) * 256 + (16 q + c)
// * 4 + e % 4, so a wave reads its 16 x 32 operand tile as two contiguous
// kilobytes.  n_pad * rows * 4 bytes.
hipError_t launch_d2_table(const Points& P, int n, int n_pad, int rows, float* tab, hipStream_t s);

// Work order of a pruned single-job sweep whose focus points are the points
// of F in order (the data table itself, or the prediction locations) against
// n data points (gwm_plan.hip): counts[tile * nseg + seg] = data chunks of the
// segment within the zero-weight distance of the tile's box, then order = the
// items by decreasing count within each XCD group.  tile_pts = focus points
// per tile (64 R), tiles = ceil(F.n / tile_pts); counts and order hold
// tiles * nseg ints.  nseg <= kPlanMaxSeg; counts above kPlanMaxCount are
// ordered by buckets of seg_len / 64 / kPlanMaxCount chunks.
constexpr int kPlanMaxSeg = 2048;
constexpr int kPlanMaxCount = 128;
constexpr int kPlanGroups = 8;
hipError_t launch_sweep_plan(int dist, const Points& F, int n, const double* boxes, const double* bw_pt, double bw,
                             const double* chunk_bw, int kernel, int mixed, int tile_pts, int seg_len, int nseg,
                             int tiles, int* counts, int* order, hipStream_t s);

// Largest value of each 64-entry chunk of v[0..n) (chunk bandwidths for the
// pruned qDiag sweeps).
hipError_t launch_chunk_max(const double* v, int n, double* out, hipStream_t s);

// Adaptive bandwidth resolution (BandwidthWeight.cpp:40-55): per focus the
// interpolated distance between the floor(bw)-th and floor(bw)+1-th nearest
// neighbours, or (bw/n)*max distance when bw >= n.  With the chunk boxes of
// spatially ordered points (boxes, nullable) only candidate chunks are read.
// left (n ints) and nleft (one counter) are scratch for the points the
// one-wave fast path leaves to the workgroup select.
// focus (nullable): thresholds at these points instead of at P's own (predict).
hipError_t launch_adaptive_bw(int dist, const Points& P, double bw, double* out, const double* boxes, int* left,
                              unsigned int* nleft, hipStream_t s, const Points* focus = nullptr);
// Neighbour table: row f (K entries, K a power of two, 2 <= K <= 8192,
// K < n) holds the K smallest selection keys of point f among all points
// (squared distances, Euclidean; distances, great circle; as IEEE-754 bit
// patterns) in ascending order.  The lookup resolves an adaptive bandwidth
// with 1 <= bw and floor(bw) <= K - 1 from the table, bit-identical to
// launch_adaptive_bw.
hipError_t launch_neighbour_table(int dist, const Points& P, const double* boxes, int K, unsigned long long* table,
                                  hipStream_t s);
hipError_t launch_neighbour_lookup(int dist, const unsigned long long* table, int K, int n, double bw, double* out,
                                   hipStream_t s);
// Maximum pairwise distance (CRSDistance::maxDistance, CRSDistance.cpp:133-143).
hipError_t launch_max_distance(int dist, const Points& P, double* block_max, int nblocks, hipStream_t s);
// Distances from one focus to all points (test hook).
hipError_t launch_distance_row(int dist, const Points& P, int focus, double* out, hipStream_t s);

// kSolvePredict: betas only, at focus points that are not data points (X and
// y are not read; n counts the focus points).
enum SolveMode { kSolveCV = 0, kSolveAIC = 1, kSolveFit = 2, kSolveBeta = 3, kSolvePredict = 4 };

struct SolveArgs {
    const double* part;    // contraction partial sums [nseg][PT][n_pad]
    int nseg;
    int PT;                // total padded columns
    int colB;              // first column of the w^2 block
    int n;
    int n_pad;
    const double* X;       // n x k column-major
    const double* y;
    const double* bw_pt;   // per-focus bandwidth or nullptr
    double bw;
    int kernel;
    // outputs (per focus)
    double* res2;          // (y_i - x_i^T beta_i)^2
    double* sii;           // S_ii
    double* ssi;           // sum_j S_ij^2
    double* betas;         // n x k column-major (fit)
    double* se_raw;        // n x k column-major (fit)
    double* G;             // row-major [n][ldg] vech(g g^T) (fit)
    int ldg;
    double* gvec;          // row-major [n][k] g = A^-1 x_i (fit)
    int* fail;             // set to 1 on a non-SPD local design
    int f0 = 0;            // first focus of a batched slice
    double* cond = nullptr;  // per focus: max_a A_aa / L_aa^2 (Jacobi-scaled inverse pivot), or null
    // optional focus list (as in SweepJob): entry t solves focus focus_idx[t]
    // from partial-sum column t
    const int* focus_idx = nullptr;
    int nfocus = 0;
    double* ainv = nullptr;  // per focus: A^-1 row-major k x k (kSolveBeta), or null
    int part_off = 0;        // part holds focus points (or list entries) from part_off on
    // mixed-precision solves of CV criteria and predictions (with cond): per
    // focus the distance to the nearest other (data) point; a focus whose
    // largest weight, that point's, is below kMixedMinWeight gets cond =
    // infinity, which sends it to the FP64 recomputation
    const double* dnn = nullptr;
};

hipError_t launch_solve(int k, int mode, const SolveArgs& a, hipStream_t s);
// Batched solves of jobs with the same k: start[] (device, njobs+1 entries) is
// the prefix sum of the jobs' focus counts; entry t of job j solves focus
// jobs[j].f0 + t - start[j]; total = start[njobs].
hipError_t launch_solve_batch(int k, int mode, const SolveArgs* jobs, const int* start, int njobs, int total,
                              hipStream_t s);

struct QDiagJob {
    const double* part;
    int nseg, PT, n, n_pad;
    const double* X;
    const double* sii;
    double* qdiag;
    int f0 = 0;            // first point of a batched slice
    int part_off = 0;      // part holds points from part_off on
};
hipError_t launch_qdiag_batch(int k, const QDiagJob* jobs, const int* start, int njobs, int total, hipStream_t s);

struct ReduceJob {
    const double* a;
    const double* b;
    const double* c;
    int n;
    double* out;   // 3 sums, the failure flag, max of m
    int* fail;     // copied to out[3] and cleared; may be null
    const double* m = nullptr;   // out[4] = max over n (0 when null)
};
hipError_t launch_reduce_batch(const ReduceJob* jobs, int njobs, hipStream_t s);

struct SelectJob {
    Points P;
    double bw;
    double* out;
    int f0;                 // first focus of the slice
    const double* boxes;    // chunk boxes (pruning) or null
};
// Batched adaptive bandwidths: start[] is the prefix sum of the jobs' focus
// counts (as for the batched solves); left (total ints) and nleft are scratch.
hipError_t launch_adaptive_bw_batch(int dist, const SelectJob* jobs, const int* start, int njobs, int total,
                                    int* left, unsigned int* nleft, hipStream_t s);

// qDiag_j = 1 - 2 S_jj + x_j^T M_j x_j, M_j from the kModeQDiag contraction.
hipError_t launch_qdiag(int k, const double* part, int nseg, int PT, int n, int n_pad, const double* X,
                        const double* sii, double* qdiag, hipStream_t s);

// S matrix rows (only for small n): S_ij = w_ij g_i^T x_j.
hipError_t launch_smatrix(int dist, int k, const Points& P, const double* X, const double* gvec,
                          const double* bw_pt, double bw, int kernel, double* S, hipStream_t s);

// Focus points whose conditioning indicator exceeds limit (or is not finite):
// their indices go to list[] (unordered), their number to *count, which must
// be zero beforehand.
// Independent-variable selection: packed lower triangle of [X y]^T [X y]
// ((k+1)(k+2)/2 entries), then per candidate column set (sets: nsets x k,
// nvars[s] used) out[2s] = RSS and out[2s+1] = tr(S) of its OLS (NaN: not SPD).
hipError_t launch_gram(const double* X, const double* y, int n, int k, double* G, hipStream_t s);
// part: 2 * nsets * varsel_blocks(n) doubles of scratch.
int varsel_blocks(int n);
hipError_t launch_varsel(const double* X, const double* y, int n, int k, const double* G, const int* sets,
                         const int* nvars, int nsets, double* part, double* out, hipStream_t s);

// Multiscale hat matrix (n x n column-major, context point order):
// R (k slices) and S0 after fitAll from the per-point A^-1; the single-column
// hat matrix S of x at a bandwidth (dinv: n scratch, fail set when a
// denominator is not positive); out = a - b; out[0..1] = tr(S), sum S^2
// (part: 2048 scratch doubles).
hipError_t launch_hat_init(int dist, const Points& P, int k, const double* X, const double* ainv, const double* bw_pt,
                           double bw, int kernel, double* R, double* S0, hipStream_t s);
hipError_t launch_var_smatrix(int dist, const Points& P, const double* x, const double* bw_pt, double bw, int kernel,
                              double* dinv, int* fail, double* S, hipStream_t s);
hipError_t launch_sub(const double* a, const double* b, double* out, size_t count, hipStream_t s);
// C += A B for n x n column-major FP64 matrices on the matrix cores (gwm_gemm.hip).
hipError_t launch_dgemm_acc(int n, const double* A, const double* B, double* C, hipStream_t s);
hipError_t launch_trace(const double* S, int n, double* part, double* out, hipStream_t s);

hipError_t launch_compact(const double* cond, int n, double limit, int* list, int* count, hipStream_t s);
// The same for many problems in one launch (*count must be zero beforehand),
// and one launch that zeroes the ints ptrs[0..count).
struct CompactJob {
    const double* cond;
    int n;
    double limit;
    int* list;
    int* count;
};
hipError_t launch_compact_batch(const CompactJob* jobs, int njobs, int max_n, hipStream_t s);
hipError_t launch_zero_ints(int* const* ptrs, int count, hipStream_t s);

// Deterministic sums of up to 3 vectors of length n into out[0..2], and the
// maximum of m into out[3] (0 when m is null).
hipError_t launch_reduce(const double* a, const double* b, const double* c, const double* m, int n, double* out,
                         hipStream_t s);

// Single-column local solves for the multiscale path (GWRMultiscale.cpp:420-482,
// 560-636): part columns 0 = sum w x^2, 1 = sum w x y, colB = sum w^2 x^2.
// *fail |= 1 on a denominator that is not positive; with dnn (mixed CV
// evaluations, per point the distance to its nearest other point) *fail |= 2
// when a point's largest weight is below kMixedMinWeight.
hipError_t launch_solve_var(int mode, const double* part, int nseg, int PT, int colB, int n, int n_pad,
                            const double* x, const double* y, const double* bw_pt, double bw, int kernel,
                            double* res2, double* sii, double* beta, int* fail, hipStream_t s,
                            const double* dnn = nullptr);

}  // namespace gwm
