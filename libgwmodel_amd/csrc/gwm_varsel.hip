// Independent-variable selection criterion on the device.
//
// GWRBasic::indepVarsSelectionCriterion{Serial,Omp,Cuda}
// (src/gwmodelpp/GWRBasic.cpp:309-346, 537-580, 838-900) solves a local
// regression at every location with unit weights (:319), i.e. the same global
// OLS n times, and scores the candidate column set by AICc with
// tr(S) = sum_i x_i^T (X^T X)^-1 x_i.  Here the Gram matrix of [X y] is formed
// once per selection (one workgroup per entry, a fixed-order reduction), and
// each candidate set is a group of workgroups over slices of the points: each
// factors the set's sub-Gram (Cholesky, beta with one refinement step) and
// sums the squared residuals (GWRBase::RSS order within a point) and hat
// diagonals of its slice; a last kernel adds the slices in order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gwm_contract.h"

namespace gwm {
namespace {

constexpr int kVarMax = 17;   // columns of a candidate set (16) + y

// Workgroup sum of v (256 lanes), fixed order; result valid in every lane.
__device__ __forceinline__ double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}

// Entry e of the packed lower triangle of [X y]^T [X y] ((k+1)(k+2)/2 entries,
// row-major: e = r(r+1)/2 + c, c <= r; column k is y).
__global__ __launch_bounds__(256) void gram_kernel(const double* X, const double* y, int n, int k, double* G) {
    __shared__ double red[256];
    const int e = blockIdx.x;
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= e) ++r;
    const int c = e - r * (r + 1) / 2;
    const double* a = r == k ? y : X + (size_t)r * n;
    const double* b = c == k ? y : X + (size_t)c * n;
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) s += a[j] * b[j];
    s = block_sum(s, red);
    if (threadIdx.x == 0) G[e] = s;
}

__device__ __forceinline__ double gram_at(const double* G, int r, int c) {
    return r >= c ? G[r * (r + 1) / 2 + c] : G[c * (c + 1) / 2 + r];
}

// nblk workgroups per candidate set (set = blockIdx / nblk), each over one
// contiguous slice of the points: sets[s * k + a], a < nvars[s], are column
// indices of X.  part[2 (s nblk + b)] = the slice's RSS, [.. + 1] its tr(S);
// NaN for both when the sub-Gram is not positive definite (inv_sympd's failure).
__global__ __launch_bounds__(256) void varsel_kernel(const double* X, const double* y, int n, int k,
                                                     const double* G, const int* sets, const int* nvars, int nblk,
                                                     double* part) {
    __shared__ double red[256];
    __shared__ double L[kVarMax * kVarMax], A[kVarMax * kVarMax];
    __shared__ double beta[kVarMax], b[kVarMax], z[kVarMax], r[kVarMax], bt[kVarMax], d[kVarMax];
    __shared__ int cols[kVarMax];
    __shared__ int ok;
    const int s = blockIdx.x / nblk, blk = blockIdx.x % nblk;
    const int p = nvars[s];
    const int chunk = (n + nblk - 1) / nblk;
    const int j0 = blk * chunk, j1 = min(n, j0 + chunk);
    if (threadIdx.x < p) cols[threadIdx.x] = sets[(size_t)s * k + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        // Cholesky of A = X_s^T X_s (lower, row-major p x p) and beta = A^-1 b
        int good = 1;
        for (int i = 0; i < p; ++i) {
            b[i] = gram_at(G, k, cols[i]);
            for (int j = 0; j <= i; ++j) A[i * kVarMax + j] = gram_at(G, cols[i], cols[j]);
        }
        for (int j = 0; j < p && good; ++j) {
            double dj = A[j * kVarMax + j];
            for (int q = 0; q < j; ++q) dj -= L[j * kVarMax + q] * L[j * kVarMax + q];
            if (!(dj > 0.0)) {
                good = 0;
                break;
            }
            const double ljj = sqrt(dj);
            L[j * kVarMax + j] = ljj;
            for (int i = j + 1; i < p; ++i) {
                double t = A[i * kVarMax + j];
                for (int q = 0; q < j; ++q) t -= L[i * kVarMax + q] * L[j * kVarMax + q];
                L[i * kVarMax + j] = t / ljj;
            }
        }
        if (good) {
            auto solve = [&](const double* v, double* x) {
                for (int i = 0; i < p; ++i) {
                    double t = v[i];
                    for (int q = 0; q < i; ++q) t -= L[i * kVarMax + q] * z[q];
                    z[i] = t / L[i * kVarMax + i];
                }
                for (int i = p - 1; i >= 0; --i) {
                    double t = z[i];
                    for (int q = i + 1; q < p; ++q) t -= L[q * kVarMax + i] * x[q];
                    x[i] = t / L[i * kVarMax + i];
                }
            };
            solve(b, bt);
            // one refinement step on A beta = b
            for (int i = 0; i < p; ++i) {
                double t = b[i];
                for (int j = 0; j < p; ++j) t -= (i >= j ? A[i * kVarMax + j] : A[j * kVarMax + i]) * bt[j];
                r[i] = t;
            }
            solve(r, d);
            for (int i = 0; i < p; ++i) beta[i] = bt[i] + d[i];
        }
        ok = good;
    }
    __syncthreads();
    double* out = part + 2 * (size_t)blockIdx.x;
    if (!ok) {
        if (threadIdx.x == 0) out[0] = out[1] = NAN;
        return;
    }
    double rss = 0.0, tr = 0.0;
    for (int j = j0 + threadIdx.x; j < j1; j += 256) {
        double x[kVarMax], u[kVarMax];
        double f = 0.0;
        for (int a = 0; a < p; ++a) {
            x[a] = X[j + (size_t)cols[a] * n];
            f += beta[a] * x[a];
        }
        const double res = y[j] - f;
        rss += res * res;
        // x^T A^-1 x = |L^-1 x|^2
        double h = 0.0;
        for (int i = 0; i < p; ++i) {
            double t = x[i];
            for (int q = 0; q < i; ++q) t -= L[i * kVarMax + q] * u[q];
            u[i] = t / L[i * kVarMax + i];
            h += u[i] * u[i];
        }
        tr += h;
    }
    rss = block_sum(rss, red);
    tr = block_sum(tr, red);
    if (threadIdx.x == 0) {
        out[0] = rss;
        out[1] = tr;
    }
}

// out[2s], out[2s + 1]: the slices' sums of set s, in slice order.
__global__ void varsel_final_kernel(const double* part, int nblk, int nsets, double* out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsets) return;
    double rss = 0.0, tr = 0.0;
    for (int b = 0; b < nblk; ++b) {
        rss += part[2 * ((size_t)s * nblk + b)];
        tr += part[2 * ((size_t)s * nblk + b) + 1];
    }
    out[2 * s] = rss;
    out[2 * s + 1] = tr;
}

}  // namespace

hipError_t launch_gram(const double* X, const double* y, int n, int k, double* G, hipStream_t s) {
    if (k < 1 || k + 1 > kVarMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gram_kernel, dim3((k + 1) * (k + 2) / 2), dim3(256), 0, s, X, y, n, k, G);
    return hipGetLastError();
}

int varsel_blocks(int n) { return std::max(1, std::min(256, (n + 4095) / 4096)); }

hipError_t launch_varsel(const double* X, const double* y, int n, int k, const double* G, const int* sets,
                         const int* nvars, int nsets, double* part, double* out, hipStream_t s) {
    if (nsets <= 0) return hipSuccess;
    if (k < 1 || k + 1 > kVarMax) return hipErrorInvalidValue;
    const int nblk = varsel_blocks(n);
    hipLaunchKernelGGL(varsel_kernel, dim3(nsets * nblk), dim3(256), 0, s, X, y, n, k, G, sets, nvars, nblk, part);
    hipLaunchKernelGGL(varsel_final_kernel, dim3((nsets + 63) / 64), dim3(64), 0, s, part, nblk, nsets, out);
    return hipGetLastError();
}

}  // namespace gwm
