// Batched local solves: one thread per focus location, FP64 throughout.
//
// The reference inverts each k x k normal matrix with arma::inv_sympd and
// multiplies out beta_i = A_i^-1 b_i, C_i = A_i^-1 X^T W_i, s_i = x_i^T C_i
// (src/gwmodelpp/GWRBasic.cpp:209-220, 249-252, 285-290).  Here every local
// quantity comes from the Cholesky factor of A_i without forming C_i (k x n):
//   beta_i   = A^-1 b          (+ one FP64 iterative-refinement step)
//   S_ii     = w_ii x_i^T A^-1 x_i
//   sum_j S_ij^2 = g^T B g,    g = A^-1 x_i, B = X^T W_i^2 X
//   SE_raw_a = (A^-1 B A^-1)_aa = rowsum(C_i .^ 2)_a
// (both quadratic forms in B as the squared norms |L_B^T v|^2, B = L_B L_B^T)
// A failed Cholesky pivot (not > 0, or NaN) is the inv_sympd failure: the
// criterion becomes DBL_MAX and the fit throws (GWRBasic.cpp:222-226, 254-258).
#include "gwm_contract.h"

namespace gwm {

namespace {

template <int K>
struct Sym {
    static constexpr int M = K * (K + 1) / 2;
    // Z column of the pair (a, b), a <= b: pairs enumerated row by row
    __device__ static constexpr int zcol(int a, int b) { return a * K - (a * (a - 1)) / 2 + (b - a); }
    // packed lower storage of a symmetric matrix, element (i, j), j <= i
    __device__ static constexpr int low(int i, int j) { return i * (i + 1) / 2 + j; }
};

// sum_{g < nseg} p[g stride] in segment order -- the bits of the plain loop
// -- with the loads of eight segments issued together (the plain loop waits
// for each load before the next: 704 dependent loads per focus point at k = 8)
__device__ __forceinline__ double seg_sum(const double* p, size_t stride, int nseg) {
    double s = 0.0;
    for (int g0 = 0; g0 < nseg; g0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[(size_t)min(g0 + u, nseg - 1) * stride];   // in bounds
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (g0 + u < nseg) s += v[u];
    }
    return s;
}

template <int K>
__device__ __forceinline__ double sum_parts(const SolveArgs& a, int c, int i) {
    const size_t np = a.n_pad;
    const size_t col = (size_t)(i - a.part_off);
    return seg_sum(a.part + (size_t)c * np + col, (size_t)a.PT * np, a.nseg);
}

// Cholesky A = L L^T in packed lower storage.  Returns false on a pivot that
// is not positive (dpotrf's failure condition).
template <int K>
__device__ __forceinline__ bool cholesky(const double* A, double* L) {
#pragma unroll
    for (int j = 0; j < K; ++j) {
        double s = A[Sym<K>::low(j, j)];
#pragma unroll
        for (int p = 0; p < j; ++p) s -= L[Sym<K>::low(j, p)] * L[Sym<K>::low(j, p)];
        if (!(s > 0.0)) return false;
        double ljj = sqrt(s);
        L[Sym<K>::low(j, j)] = ljj;
#pragma unroll
        for (int i = j + 1; i < K; ++i) {
            double t = A[Sym<K>::low(i, j)];
#pragma unroll
            for (int p = 0; p < j; ++p) t -= L[Sym<K>::low(i, p)] * L[Sym<K>::low(j, p)];
            L[Sym<K>::low(i, j)] = t / ljj;
        }
    }
    return true;
}

// x = A^-1 v from the factor: L y = v, L^T x = y.
template <int K>
__device__ __forceinline__ void chol_solve(const double* L, const double* v, double* x) {
    double y[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        double s = v[i];
#pragma unroll
        for (int p = 0; p < i; ++p) s -= L[Sym<K>::low(i, p)] * y[p];
        y[i] = s / L[Sym<K>::low(i, i)];
    }
#pragma unroll
    for (int i = K - 1; i >= 0; --i) {
        double s = y[i];
#pragma unroll
        for (int p = i + 1; p < K; ++p) s -= L[Sym<K>::low(p, i)] * x[p];
        x[i] = s / L[Sym<K>::low(i, i)];
    }
}

template <int K>
__device__ __forceinline__ double quad_form(const double* Bp, const double* g) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < K; ++a) {
        s += Bp[Sym<K>::low(a, a)] * g[a] * g[a];
#pragma unroll
        for (int b = 0; b < a; ++b) s += 2.0 * Bp[Sym<K>::low(a, b)] * g[a] * g[b];
    }
    return s;
}

// v^T B v from the Cholesky factor of B (packed lower, B = L L^T): the sum of
// the squares of L^T v, never negative, as the reference's rowsum(C .^ 2) is
// not; the direct form (quad_form) can round below zero when the local
// design is nearly singular.
template <int K>
__device__ __forceinline__ double gram_form(const double* L, const double* v) {
    double s = 0.0;
#pragma unroll
    for (int a = 0; a < K; ++a) {
        double u = 0.0;
#pragma unroll
        for (int b = a; b < K; ++b) u += L[Sym<K>::low(b, a)] * v[b];
        s += u * u;
    }
    return s;
}

template <int K, int MODE>
__device__ __forceinline__ void solve_one(const SolveArgs& a, int t) {
    // partial sums in column t; the focus is t itself or focus_idx[t]
    const int i = a.focus_idx ? a.focus_idx[t] : t;
    const int n = a.n;
    constexpr int M = Sym<K>::M;

    double A[M], L[M], b[K], beta[K], xi[K];
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) A[Sym<K>::low(r, c)] = sum_parts<K>(a, Sym<K>::zcol(c, r), t);
#pragma unroll
    for (int r = 0; r < K; ++r) b[r] = sum_parts<K>(a, M + r, t);
    if (MODE != kSolvePredict)
#pragma unroll
        for (int r = 0; r < K; ++r) xi[r] = a.X[i + (size_t)r * n];

    if (!cholesky<K>(A, L)) {
        *a.fail = 1;
        if (MODE == kSolvePredict) {
#pragma unroll
            for (int p = 0; p < K; ++p) a.betas[i + (size_t)p * n] = NAN;
        } else {
            a.res2[i] = NAN;
        }
        if (MODE == kSolveAIC || MODE == kSolveFit) a.sii[i] = NAN;
        if (a.cond) a.cond[i] = INFINITY;
        return;
    }
    if (a.cond) {
        // conditioning indicator of the Jacobi-scaled design: the inverse of
        // its smallest squared Cholesky pivot (>= 1; large when a column is
        // nearly a combination of the earlier ones)
        double c = 1.0;
#pragma unroll
        for (int p = 0; p < K; ++p) {
            const double l = L[Sym<K>::low(p, p)];
            c = fmax(c, A[Sym<K>::low(p, p)] / (l * l));
        }
        // mixed precision: a focus point far from every other point has only
        // tiny weights, which the FP16 split does not resolve
        if (a.dnn && !(nearest_weight(a.kernel, a.dnn[i], a.bw_pt ? a.bw_pt[i] : a.bw) >= kMixedMinWeight))
            c = INFINITY;
        a.cond[i] = c;
    }
    chol_solve<K>(L, b, beta);
    {
        // one step of iterative refinement on the local normal equation
        double r[K], dlt[K];
#pragma unroll
        for (int p = 0; p < K; ++p) {
            double s = b[p];
#pragma unroll
            for (int c = 0; c < K; ++c) s -= A[p >= c ? Sym<K>::low(p, c) : Sym<K>::low(c, p)] * beta[c];
            r[p] = s;
        }
        chol_solve<K>(L, r, dlt);
#pragma unroll
        for (int p = 0; p < K; ++p) beta[p] += dlt[p];
    }
    if (MODE == kSolvePredict) {
#pragma unroll
        for (int p = 0; p < K; ++p) a.betas[i + (size_t)p * n] = beta[p];
        return;
    }
    // residual: y_i - sum_a beta_a x_a (GWRBase::Fitted order)
    double yhat = 0.0;
#pragma unroll
    for (int p = 0; p < K; ++p) yhat += beta[p] * xi[p];
    double res = a.y[i] - yhat;
    a.res2[i] = res * res;
    if (MODE == kSolveCV || MODE == kSolveAIC) {
        // (res * yhat)^2: the scale of the criterion's rounding error (an
        // error of yhat relative to itself moves res^2 by 2 res yhat times it)
        const double q = res * yhat;
        a.ssi[i] = q * q;
    }
    if (MODE == kSolveCV) return;
    if (MODE == kSolveBeta) {
#pragma unroll
        for (int p = 0; p < K; ++p) a.betas[i + (size_t)p * n] = beta[p];
        if (a.ainv) {
            // columns of A^-1 (symmetric: stored row-major as they come)
            double e[K], col[K];
            for (int p = 0; p < K; ++p) {
#pragma unroll
                for (int q = 0; q < K; ++q) e[q] = q == p ? 1.0 : 0.0;
                chol_solve<K>(L, e, col);
#pragma unroll
                for (int q = 0; q < K; ++q) a.ainv[(size_t)i * K * K + q * K + p] = col[q];
            }
        }
        return;
    }

    const double bi = a.bw_pt ? a.bw_pt[i] : a.bw;
    const double wii = kernel_f64(a.kernel, 0.0, bi);
    double g[K];
    chol_solve<K>(L, xi, g);
    double xg = 0.0;
#pragma unroll
    for (int p = 0; p < K; ++p) xg += xi[p] * g[p];
    a.sii[i] = wii * xg;
    if (MODE == kSolveAIC) return;

    // B = X^T W^2 X from the w^2 block, factored in A's storage (B is positive
    // semidefinite: the forms below are sums of squares).  A factor that fails
    // (B singular to rounding) leaves the direct form, clamped at zero.
    double* B = A;
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) B[Sym<K>::low(r, c)] = sum_parts<K>(a, a.colB + Sym<K>::zcol(c, r), t);
    const bool fb = cholesky<K>(B, B);
    if (!fb)
#pragma unroll
        for (int r = 0; r < K; ++r)
#pragma unroll
            for (int c = 0; c <= r; ++c) B[Sym<K>::low(r, c)] = sum_parts<K>(a, a.colB + Sym<K>::zcol(c, r), t);
    a.ssi[i] = fb ? gram_form<K>(B, g) : fmax(quad_form<K>(B, g), 0.0);
#pragma unroll
    for (int p = 0; p < K; ++p) a.betas[i + (size_t)p * n] = beta[p];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        double e[K], col[K];
#pragma unroll
        for (int p = 0; p < K; ++p) e[p] = p == c ? 1.0 : 0.0;
        chol_solve<K>(L, e, col);
        a.se_raw[i + (size_t)c * n] = fb ? gram_form<K>(B, col) : fmax(quad_form<K>(B, col), 0.0);
    }
#pragma unroll
    for (int p = 0; p < K; ++p) a.gvec[(size_t)i * K + p] = g[p];
    double* Gi = a.G + (size_t)i * a.ldg;
#pragma unroll
    for (int p = 0; p < K; ++p)
#pragma unroll
        for (int q = p; q < K; ++q) Gi[Sym<K>::zcol(p, q)] = g[p] * g[q];
}

template <int K, int MODE>
__global__ __launch_bounds__(128) void solve_kernel(SolveArgs a) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (a.focus_idx ? a.nfocus : a.n)) return;
    solve_one<K, MODE>(a, t);
}

// Batched solves: thread t of the launch belongs to the job whose focus range
// [start[j], start[j+1]) holds t (start is a prefix sum of the job sizes).
template <int K, int MODE>
__global__ __launch_bounds__(128) void solve_batch_kernel(const SolveArgs* jobs, const int* start, int njobs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= start[njobs]) return;
    const int j = find_job(start, njobs, t);
    solve_one<K, MODE>(jobs[j], jobs[j].f0 + t - start[j]);
}

template <int K>
__device__ __forceinline__ void qdiag_one(const QDiagJob& J, int j) {
    const int n = J.n;
    double x[K];
#pragma unroll
    for (int p = 0; p < K; ++p) x[p] = J.X[j + (size_t)p * n];
    double q = 0.0;
#pragma unroll
    for (int a = 0; a < K; ++a)
#pragma unroll
        for (int b = a; b < K; ++b) {
            const int c = Sym<K>::zcol(a, b);
            const double m = seg_sum(J.part + (size_t)c * J.n_pad + (j - J.part_off), (size_t)J.PT * J.n_pad, J.nseg);
            q += (a == b ? 1.0 : 2.0) * m * x[a] * x[b];
        }
    J.qdiag[j] = 1.0 - 2.0 * J.sii[j] + q;
}

template <int K>
__global__ __launch_bounds__(128) void qdiag_kernel(QDiagJob J) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J.n) return;
    qdiag_one<K>(J, j);
}

template <int K>
__global__ __launch_bounds__(128) void qdiag_batch_kernel(const QDiagJob* jobs, const int* start, int njobs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= start[njobs]) return;
    const int j = find_job(start, njobs, t);
    qdiag_one<K>(jobs[j], jobs[j].f0 + t - start[j]);
}

template <int DIST>
__global__ void smatrix_kernel(Points P, int k, const double* X, const double* gvec, const double* bw_pt, double bw,
                               int kernel, double* S) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    const int n = P.n;
    if (j >= n) return;
    double d = distance_f64<DIST>(P, i, j);
    double w = kernel_f64(kernel, d, bw_pt ? bw_pt[i] : bw);
    double s = 0.0;
    for (int p = 0; p < k; ++p) s += gvec[(size_t)i * k + p] * X[j + (size_t)p * n];
    S[i + (size_t)j * n] = w * s;
}

// Largest value of each 64-entry chunk.
__global__ __launch_bounds__(256) void chunk_max_kernel(const double* v, int n, double* out) {
    const int ch = blockIdx.x * blockDim.x + threadIdx.x;
    if (ch * 64 >= n) return;
    const int j1 = min(n, ch * 64 + 64);
    double m = v[ch * 64];
    for (int j = ch * 64 + 1; j < j1; ++j) m = fmax(m, v[j]);
    out[ch] = m;
}

// Indices of the focus points whose indicator exceeds the limit.
__global__ __launch_bounds__(256) void compact_kernel(const double* cond, int n, double limit, int* list, int* count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double c = cond[i];
    if (!(c <= limit)) list[atomicAdd(count, 1)] = i;
}

// compact_kernel over many problems: problem blockIdx.y, grid-stride in x.
__global__ __launch_bounds__(256) void compact_batch_kernel(const CompactJob* jobs) {
    const CompactJob J = jobs[blockIdx.y];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < J.n; i += gridDim.x * 256) {
        const double c = J.cond[i];
        if (!(c <= J.limit)) J.list[atomicAdd(J.count, 1)] = i;
    }
}

__global__ void zero_ints_kernel(int* const* ptrs, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) *ptrs[i] = 0;
}

// Deterministic fixed-order sums of a, b, c and the maximum of m (one
// workgroup): out[0..2] = sums, *mx = max (0 when m is null).
__device__ __forceinline__ void reduce_body(const double* a, const double* b, const double* c, const double* m, int n,
                                           double* out, double* mx) {
    __shared__ double sa[1024], sb[1024], sc[1024], sm[1024];
    double xa = 0.0, xb = 0.0, xc = 0.0, xm = 0.0;
    // each thread's elements in order, loaded eight at a time (the plain
    // loop waits for every load before the next)
    int i = threadIdx.x;
    for (; i + 7 * 1024 < n; i += 8 * 1024) {
        double va[8], vb[8], vc[8], vm[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int j = i + u * 1024;
            va[u] = a ? a[j] : 0.0;
            vb[u] = b ? b[j] : 0.0;
            vc[u] = c ? c[j] : 0.0;
            vm[u] = m ? m[j] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            if (a) xa += va[u];
            if (b) xb += vb[u];
            if (c) xc += vc[u];
            if (m) xm = fmax(xm, vm[u]);
        }
    }
    for (; i < n; i += 1024) {
        if (a) xa += a[i];
        if (b) xb += b[i];
        if (c) xc += c[i];
        if (m) xm = fmax(xm, m[i]);
    }
    sa[threadIdx.x] = xa;
    sb[threadIdx.x] = xb;
    sc[threadIdx.x] = xc;
    sm[threadIdx.x] = xm;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
            sc[threadIdx.x] += sc[threadIdx.x + w];
            sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sa[0];
        out[1] = sb[0];
        out[2] = sc[0];
        *mx = sm[0];
    }
}

__global__ __launch_bounds__(1024) void reduce3_kernel(const double* a, const double* b, const double* c,
                                                       const double* m, int n, double* out) {
    reduce_body(a, b, c, m, n, out, out + 3);
}

// Batched fixed-order sums: block b reduces job b's vectors (same order as a
// one-problem reduce3 launch); out[3] receives the job's failure flag, which
// is cleared for its next evaluation.
__global__ __launch_bounds__(1024) void reduce_batch_kernel(const ReduceJob* jobs) {
    const ReduceJob& J = jobs[blockIdx.x];
    reduce_body(J.a, J.b, J.c, J.m, J.n, J.out, J.out + 4);
    if (threadIdx.x == 0 && J.fail) {
        J.out[3] = (double)*J.fail;
        *J.fail = 0;
    }
}

// Single-column solves (GWRMultiscale.cpp:420-482, 560-636).
__global__ __launch_bounds__(256) void solve_var_kernel(int mode, const double* part, int nseg, int PT, int colB,
                                                        int n, int n_pad, const double* x, const double* y,
                                                        const double* bw_pt, double bw, int kernel, double* res2,
                                                        double* sii, double* beta, int* fail, const double* dnn) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dnn && !(nearest_weight(kernel, dnn[i], bw_pt ? bw_pt[i] : bw) >= kMixedMinWeight)) atomicOr(fail, 2);
    const double a = seg_sum(part + i, (size_t)PT * n_pad, nseg);
    const double b = seg_sum(part + (size_t)n_pad + i, (size_t)PT * n_pad, nseg);
    (void)colB;
    if (!(a > 0.0)) {
        atomicOr(fail, 1);
        if (res2) res2[i] = NAN;
        return;
    }
    const double inv = 1.0 / a;
    const double bt = inv * b;
    if (beta) beta[i] = bt;
    if (res2) {
        double r = y[i] - x[i] * bt;
        res2[i] = r * r;
    }
    if (mode != kSolveCV && sii) {
        const double bi = bw_pt ? bw_pt[i] : bw;
        sii[i] = x[i] * (inv * (x[i] * kernel_f64(kernel, 0.0, bi)));
    }
}


template <int K>
hipError_t launch_solve_k(int mode, const SolveArgs& a, hipStream_t s) {
    const int count = a.focus_idx ? a.nfocus : a.n;
    if (count <= 0) return hipSuccess;
    dim3 grid((count + 127) / 128);
    if (mode == kSolveCV) hipLaunchKernelGGL((solve_kernel<K, kSolveCV>), grid, dim3(128), 0, s, a);
    else if (mode == kSolveAIC) hipLaunchKernelGGL((solve_kernel<K, kSolveAIC>), grid, dim3(128), 0, s, a);
    else if (mode == kSolveBeta) hipLaunchKernelGGL((solve_kernel<K, kSolveBeta>), grid, dim3(128), 0, s, a);
    else if (mode == kSolvePredict) hipLaunchKernelGGL((solve_kernel<K, kSolvePredict>), grid, dim3(128), 0, s, a);
    else hipLaunchKernelGGL((solve_kernel<K, kSolveFit>), grid, dim3(128), 0, s, a);
    return hipGetLastError();
}

template <int K>
hipError_t launch_qdiag_k(const double* part, int nseg, int PT, int n, int n_pad, const double* X,
                          const double* sii, double* q, hipStream_t s) {
    QDiagJob J{part, nseg, PT, n, n_pad, X, sii, q};
    hipLaunchKernelGGL(qdiag_kernel<K>, dim3((n + 127) / 128), dim3(128), 0, s, J);
    return hipGetLastError();
}

template <int K>
hipError_t launch_solve_batch_k(int mode, const SolveArgs* jobs, const int* start, int njobs, int total,
                                hipStream_t s) {
    dim3 grid((total + 127) / 128);
    if (total <= 0) return hipSuccess;
    if (mode == kSolveCV) hipLaunchKernelGGL((solve_batch_kernel<K, kSolveCV>), grid, dim3(128), 0, s, jobs, start, njobs);
    else if (mode == kSolveAIC) hipLaunchKernelGGL((solve_batch_kernel<K, kSolveAIC>), grid, dim3(128), 0, s, jobs, start, njobs);
    else if (mode == kSolveBeta) hipLaunchKernelGGL((solve_batch_kernel<K, kSolveBeta>), grid, dim3(128), 0, s, jobs, start, njobs);
    else hipLaunchKernelGGL((solve_batch_kernel<K, kSolveFit>), grid, dim3(128), 0, s, jobs, start, njobs);
    return hipGetLastError();
}

template <int K>
hipError_t launch_qdiag_batch_k(const QDiagJob* jobs, const int* start, int njobs, int total, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(qdiag_batch_kernel<K>, dim3((total + 127) / 128), dim3(128), 0, s, jobs, start, njobs);
    return hipGetLastError();
}

}  // namespace

#define GWM_K_SWITCH(k, CALL)                                                                          \
    switch (k) {                                                                                       \
    case 1: return CALL(1); case 2: return CALL(2); case 3: return CALL(3); case 4: return CALL(4);    \
    case 5: return CALL(5); case 6: return CALL(6); case 7: return CALL(7); case 8: return CALL(8);    \
    case 9: return CALL(9); case 10: return CALL(10); case 11: return CALL(11); case 12: return CALL(12); \
    case 13: return CALL(13); case 14: return CALL(14); case 15: return CALL(15); case 16: return CALL(16); \
    default: return hipErrorInvalidValue;                                                              \
    }

hipError_t launch_solve(int k, int mode, const SolveArgs& a, hipStream_t s) {
#define GWM_SOLVE(K) launch_solve_k<K>(mode, a, s)
    GWM_K_SWITCH(k, GWM_SOLVE)
#undef GWM_SOLVE
}

hipError_t launch_qdiag(int k, const double* part, int nseg, int PT, int n, int n_pad, const double* X,
                        const double* sii, double* qdiag, hipStream_t s) {
#define GWM_QD(K) launch_qdiag_k<K>(part, nseg, PT, n, n_pad, X, sii, qdiag, s)
    GWM_K_SWITCH(k, GWM_QD)
#undef GWM_QD
}

hipError_t launch_solve_batch(int k, int mode, const SolveArgs* jobs, const int* start, int njobs, int total,
                              hipStream_t s) {
#define GWM_SB(K) launch_solve_batch_k<K>(mode, jobs, start, njobs, total, s)
    GWM_K_SWITCH(k, GWM_SB)
#undef GWM_SB
}

hipError_t launch_qdiag_batch(int k, const QDiagJob* jobs, const int* start, int njobs, int total, hipStream_t s) {
#define GWM_QB(K) launch_qdiag_batch_k<K>(jobs, start, njobs, total, s)
    GWM_K_SWITCH(k, GWM_QB)
#undef GWM_QB
}

hipError_t launch_reduce_batch(const ReduceJob* jobs, int njobs, hipStream_t s) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(reduce_batch_kernel, dim3(njobs), dim3(1024), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_smatrix(int dist, int k, const Points& P, const double* X, const double* gvec,
                          const double* bw_pt, double bw, int kernel, double* S, hipStream_t s) {
    dim3 grid((P.n + 255) / 256, P.n);
    if (dist == kEuclidean)
        hipLaunchKernelGGL(smatrix_kernel<kEuclidean>, grid, dim3(256), 0, s, P, k, X, gvec, bw_pt, bw, kernel, S);
    else
        hipLaunchKernelGGL(smatrix_kernel<kGreatCircle>, grid, dim3(256), 0, s, P, k, X, gvec, bw_pt, bw, kernel, S);
    return hipGetLastError();
}

hipError_t launch_chunk_max(const double* v, int n, double* out, hipStream_t s) {
    const int nc = (n + 63) / 64;
    if (nc <= 0) return hipSuccess;
    hipLaunchKernelGGL(chunk_max_kernel, dim3((nc + 255) / 256), dim3(256), 0, s, v, n, out);
    return hipGetLastError();
}

hipError_t launch_compact(const double* cond, int n, double limit, int* list, int* count, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(compact_kernel, dim3((n + 255) / 256), dim3(256), 0, s, cond, n, limit, list, count);
    return hipGetLastError();
}

hipError_t launch_compact_batch(const CompactJob* jobs, int njobs, int max_n, hipStream_t s) {
    if (njobs <= 0 || max_n <= 0) return hipSuccess;
    if (njobs > 65535) return hipErrorInvalidValue;
    const int bx = std::min((max_n + 255) / 256, 1024);
    hipLaunchKernelGGL(compact_batch_kernel, dim3(bx, njobs), dim3(256), 0, s, jobs);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int* const* ptrs, int count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((count + 255) / 256), dim3(256), 0, s, ptrs, count);
    return hipGetLastError();
}

hipError_t launch_reduce(const double* a, const double* b, const double* c, const double* m, int n, double* out,
                         hipStream_t s) {
    hipLaunchKernelGGL(reduce3_kernel, dim3(1), dim3(1024), 0, s, a, b, c, m, n, out);
    return hipGetLastError();
}

hipError_t launch_solve_var(int mode, const double* part, int nseg, int PT, int colB, int n, int n_pad,
                            const double* x, const double* y, const double* bw_pt, double bw, int kernel,
                            double* res2, double* sii, double* beta, int* fail, hipStream_t s, const double* dnn) {
    hipLaunchKernelGGL(solve_var_kernel, dim3((n + 255) / 256), dim3(256), 0, s, mode, part, nseg, PT, colB, n, n_pad,
                       x, y, bw_pt, bw, kernel, res2, sii, beta, fail, dnn);
    return hipGetLastError();
}

}  // namespace gwm
