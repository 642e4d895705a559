// Hat matrix of the multiscale backfitting, resident in HBM.
//
// GWRMultiscale keeps S0 (n x n) and one operator R_i (n x n) per variable
// (mSArray, GWRMultiscale.cpp:163-168): after fitAll, R_i row j =
// x_ji (A_j^-1 X^T W_j)[i, :] (:215-225) and S0 = sum_i R_i; each backfitting
// step of variable i with its single-column hat matrix S_i sets
// R_i <- S_i (R_i - S0) + S_i and S0 <- R_i(new) - (R_i(old) - S0) (:296-301);
// the diagnostics use tr(S0) and tr(S0^T S0) (:176-179).  Here every matrix
// is column-major n x n in the context's point order (traces do not depend
// on the order), the weights are recomputed on the fly from the FP64
// distance and kernel, and the product S_i M is a plain FP64 GEMM (rocBLAS).
#include <hip/hip_runtime.h>

#include "gwm_contract.h"
#include "gwm_device.h"

namespace gwm {
namespace {

// R_i[j, l] = x_ji * sum_b Ainv_j[i, b] (x_lb w_jl) for every variable i and
// S0[j, l] = sum_i R_i[j, l]; one thread per (j, l), j fastest.
template <int DIST>
__global__ __launch_bounds__(256) void hat_init_kernel(Points P, int k, const double* X, const double* ainv,
                                                       const double* bw_pt, double bw, int kernel, double* R,
                                                       double* S0) {
    const int n = P.n;
    const int jb = (n + 255) / 256;
    const int l = blockIdx.x / jb;
    const int j = (blockIdx.x % jb) * 256 + threadIdx.x;
    if (j >= n || l >= n) return;
    const GPoints G = global_points(P);
    const double w = kernel_f64(kernel, distance_f64<DIST>(G, j, l), bw_pt ? bw_pt[j] : bw);
    const size_t nn = (size_t)n * n, at = (size_t)j + (size_t)l * n;
    const double* A = ainv + (size_t)j * k * k;
    double s0 = 0.0;
    for (int i = 0; i < k; ++i) {
        double c = 0.0;
        for (int b = 0; b < k; ++b) c += A[i * k + b] * (X[l + (size_t)b * n] * w);
        const double r = X[j + (size_t)i * n] * c;
        R[(size_t)i * nn + at] = r;
        s0 += r;
    }
    S0[at] = s0;
}

// d_j = sum_l (x_l w_jl) x_l for the single column x (fitVar's X^T W X), one
// workgroup per row j, fixed-order reduction; fail when d_j is not positive.
template <int DIST>
__global__ __launch_bounds__(256) void var_denominator_kernel(Points P, const double* x, const double* bw_pt, double bw,
                                                              int kernel, double* dinv, int* fail) {
    __shared__ double red[256];
    const int n = P.n, j = blockIdx.x, t = threadIdx.x;
    const GPoints G = global_points(P);
    const double b = bw_pt ? bw_pt[j] : bw;
    double s = 0.0;
    for (int l = t; l < n; l += 256) s += (x[l] * kernel_f64(kernel, distance_f64<DIST>(G, j, l), b)) * x[l];
    red[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) {
        const double d = red[0];
        if (!(d > 0.0)) *fail = 1;
        dinv[j] = 1.0 / d;
    }
}

// S_i[j, l] = x_j (d_j^-1 (x_l w_jl)) (fitVar's si = x(i) * ci, GWRMultiscale.cpp:445-447).
template <int DIST>
__global__ __launch_bounds__(256) void var_smatrix_kernel(Points P, const double* x, const double* bw_pt, double bw,
                                                          int kernel, const double* dinv, double* S) {
    const int n = P.n;
    const int jb = (n + 255) / 256;
    const int l = blockIdx.x / jb;
    const int j = (blockIdx.x % jb) * 256 + threadIdx.x;
    if (j >= n || l >= n) return;
    const GPoints G = global_points(P);
    const double w = kernel_f64(kernel, distance_f64<DIST>(G, j, l), bw_pt ? bw_pt[j] : bw);
    S[(size_t)j + (size_t)l * n] = x[j] * (dinv[j] * (x[l] * w));
}

__global__ __launch_bounds__(256) void sub_kernel(const double* a, const double* b, double* out, size_t count) {
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < count; e += (size_t)gridDim.x * 256)
        out[e] = a[e] - b[e];
}

// Partial sums of tr(S) and sum S^2 per workgroup (fixed grid, fixed order).
constexpr int kTraceGrid = 1024;
__global__ __launch_bounds__(256) void trace_partial_kernel(const double* S, int n, double* part) {
    __shared__ double r1[256], r2[256];
    const size_t nn = (size_t)n * n;
    double tr = 0.0, sq = 0.0;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < nn; e += (size_t)kTraceGrid * 256) {
        const double v = S[e];
        sq += v * v;
        if (e % ((size_t)n + 1) == 0) tr += v;
    }
    r1[threadIdx.x] = tr;
    r2[threadIdx.x] = sq;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r1[0];
        part[2 * blockIdx.x + 1] = r2[0];
    }
}

__global__ __launch_bounds__(256) void trace_final_kernel(const double* part, double* out) {
    __shared__ double r1[256], r2[256];
    double tr = 0.0, sq = 0.0;
    for (int b = threadIdx.x; b < kTraceGrid; b += 256) {
        tr += part[2 * b];
        sq += part[2 * b + 1];
    }
    r1[threadIdx.x] = tr;
    r2[threadIdx.x] = sq;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = r1[0];
        out[1] = r2[0];
    }
}

int pair_blocks(int n) { return ((n + 255) / 256) * n; }

}  // namespace

hipError_t launch_hat_init(int dist, const Points& P, int k, const double* X, const double* ainv, const double* bw_pt,
                           double bw, int kernel, double* R, double* S0, hipStream_t s) {
    const dim3 grid(pair_blocks(P.n));
    if (dist == kEuclidean)
        hipLaunchKernelGGL(hat_init_kernel<kEuclidean>, grid, dim3(256), 0, s, P, k, X, ainv, bw_pt, bw, kernel, R, S0);
    else
        hipLaunchKernelGGL(hat_init_kernel<kGreatCircle>, grid, dim3(256), 0, s, P, k, X, ainv, bw_pt, bw, kernel, R,
                           S0);
    return hipGetLastError();
}

hipError_t launch_var_smatrix(int dist, const Points& P, const double* x, const double* bw_pt, double bw, int kernel,
                              double* dinv, int* fail, double* S, hipStream_t s) {
    const dim3 grid(pair_blocks(P.n));
    if (dist == kEuclidean) {
        hipLaunchKernelGGL(var_denominator_kernel<kEuclidean>, dim3(P.n), dim3(256), 0, s, P, x, bw_pt, bw, kernel,
                           dinv, fail);
        hipLaunchKernelGGL(var_smatrix_kernel<kEuclidean>, grid, dim3(256), 0, s, P, x, bw_pt, bw, kernel, dinv, S);
    } else {
        hipLaunchKernelGGL(var_denominator_kernel<kGreatCircle>, dim3(P.n), dim3(256), 0, s, P, x, bw_pt, bw, kernel,
                           dinv, fail);
        hipLaunchKernelGGL(var_smatrix_kernel<kGreatCircle>, grid, dim3(256), 0, s, P, x, bw_pt, bw, kernel, dinv, S);
    }
    return hipGetLastError();
}

hipError_t launch_sub(const double* a, const double* b, double* out, size_t count, hipStream_t s) {
    const size_t blocks = std::min<size_t>((count + 255) / 256, 65536);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(sub_kernel, dim3((unsigned)blocks), dim3(256), 0, s, a, b, out, count);
    return hipGetLastError();
}

hipError_t launch_trace(const double* S, int n, double* part, double* out, hipStream_t s) {
    hipLaunchKernelGGL(trace_partial_kernel, dim3(kTraceGrid), dim3(256), 0, s, S, n, part);
    hipLaunchKernelGGL(trace_final_kernel, dim3(1), dim3(256), 0, s, part, out);
    return hipGetLastError();
}

}  // namespace gwm
