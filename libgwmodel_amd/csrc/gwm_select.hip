// Adaptive bandwidth resolution and distance utilities.
//
// BandwidthWeight::weight (src/gwmodelpp/spatialweight/BandwidthWeight.cpp:40-55)
// sorts the n distances of a focus point and interpolates between the
// floor(bw)-th and floor(bw)+1-th smallest (the focus itself has rank 1):
//     fixbw = d0 + (d1 - d0) * (bw - floor(bw)),
// or uses (bw/n) * max(dist) when bw >= n.  Here the two order statistics are
// found without sorting, by an MSD radix select (8-bit digits, 8 passes) over
// the IEEE-754 bit patterns of the non-negative selection values, which order
// like the values themselves.  One workgroup per focus point; the histogram
// lives in LDS.  The result is bit-identical to the sort because the selected
// values are the very doubles the sort would return.
#include "gwm_contract.h"

namespace gwm {

namespace {

__device__ __forceinline__ unsigned long long key_of(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ double val_of(unsigned long long k) { return __longlong_as_double((long long)k); }

template <int DIST, class PTS, class OUT>
__device__ __forceinline__ void adaptive_bw_one(const PTS& P, double bw, OUT out, int f) {
    const int n = P.n;
    const int tid = threadIdx.x;
    __shared__ unsigned int hist[256];
    __shared__ unsigned long long s_prefix;
    __shared__ long long s_rank;
    __shared__ unsigned int s_equal;
    __shared__ unsigned long long s_next;
    __shared__ double s_red[256];

    const double dn = bw / n;
    if (!(dn < 1)) {
        // (bw/n) * max(dist); distances are >= 0 and include the focus (0)
        double m = 0.0;
        for (int j = tid; j < n; j += 256) m = fmax(m, distance_f64<DIST>(P, f, j));
        s_red[tid] = m;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) s_red[tid] = fmax(s_red[tid], s_red[tid + w]);
            __syncthreads();
        }
        if (tid == 0) out[f] = dn * s_red[0];
        return;
    }
    const double b0 = floor(bw), bx = bw - b0;
    if (tid == 0) {
        s_prefix = 0ull;
        s_rank = (long long)b0 - 1;   // 0-based rank of d0
    }
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 56 - 8 * pass;
        hist[tid] = 0u;
        __syncthreads();
        const unsigned long long prefix = s_prefix;
        for (int j = tid; j < n; j += 256) {
            unsigned long long key = key_of(select_value<DIST>(P, f, j));
            bool match = pass == 0 || (key >> (shift + 8)) == (prefix >> (shift + 8));
            if (match) atomicAdd(&hist[(key >> shift) & 255ull], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            long long rank = s_rank;
            int bin = 0;
            for (; bin < 255; ++bin) {
                if (rank < (long long)hist[bin]) break;
                rank -= hist[bin];
            }
            s_prefix = prefix | ((unsigned long long)bin << shift);
            s_rank = rank;
            s_equal = hist[bin];
        }
        __syncthreads();
    }
    const unsigned long long key0 = s_prefix;
    // second order statistic: equal to d0 when ties cover rank b0, else the
    // smallest value above d0
    const bool tie = (s_rank + 1) < (long long)s_equal;
    double d1v;
    if (tie) {
        d1v = val_of(key0);
    } else {
        if (tid == 0) s_next = ~0ull;
        __syncthreads();
        unsigned long long best = ~0ull;
        for (int j = tid; j < n; j += 256) {
            unsigned long long key = key_of(select_value<DIST>(P, f, j));
            if (key > key0 && key < best) best = key;
        }
        atomicMin(&s_next, best);
        __syncthreads();
        d1v = val_of(s_next);
    }
    if (tid == 0) {
        double d0 = select_to_distance<DIST>(val_of(key0));
        double d1 = select_to_distance<DIST>(d1v);
        out[f] = d0 + (d1 - d0) * bx;
    }
}

template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_kernel(Points P, double bw, double* out) {
    adaptive_bw_one<DIST>(P, bw, out, blockIdx.x);
}

template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_batch_kernel(const SelectJob* jobs, const int* start, int njobs) {
    const int t = blockIdx.x;
    const int j = find_job(start, njobs, t);
    const SelectJob J = jobs[j];
    adaptive_bw_one<DIST>(global_points(J.P), J.bw, gptr(J.out), J.f0 + t - start[j]);
}

template <int DIST>
__global__ __launch_bounds__(256) void max_distance_kernel(Points P, double* block_max) {
    __shared__ double s_red[256];
    const int n = P.n;
    const int f = blockIdx.x * 256 + threadIdx.x;
    double m = 0.0;
    if (f < n)
        for (int j = 0; j < n; ++j) m = fmax(m, distance_f64<DIST>(P, f, j));
    s_red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_red[threadIdx.x] = fmax(s_red[threadIdx.x], s_red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_max[blockIdx.x] = s_red[0];
}

template <int DIST>
__global__ void distance_row_kernel(Points P, int focus, double* out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P.n) out[j] = distance_f64<DIST>(P, focus, j);
}

}  // namespace

hipError_t launch_adaptive_bw(int dist, const Points& P, double bw, double* out, hipStream_t s) {
    if (dist == kEuclidean)
        hipLaunchKernelGGL(adaptive_bw_kernel<kEuclidean>, dim3(P.n), dim3(256), 0, s, P, bw, out);
    else
        hipLaunchKernelGGL(adaptive_bw_kernel<kGreatCircle>, dim3(P.n), dim3(256), 0, s, P, bw, out);
    return hipGetLastError();
}

hipError_t launch_adaptive_bw_batch(int dist, const SelectJob* jobs, const int* start, int njobs, int total,
                                    hipStream_t s) {
    if (total <= 0) return hipSuccess;
    if (dist == kEuclidean)
        hipLaunchKernelGGL(adaptive_bw_batch_kernel<kEuclidean>, dim3(total), dim3(256), 0, s, jobs, start, njobs);
    else
        hipLaunchKernelGGL(adaptive_bw_batch_kernel<kGreatCircle>, dim3(total), dim3(256), 0, s, jobs, start, njobs);
    return hipGetLastError();
}

hipError_t launch_max_distance(int dist, const Points& P, double* block_max, int nblocks, hipStream_t s) {
    if (dist == kEuclidean)
        hipLaunchKernelGGL(max_distance_kernel<kEuclidean>, dim3(nblocks), dim3(256), 0, s, P, block_max);
    else
        hipLaunchKernelGGL(max_distance_kernel<kGreatCircle>, dim3(nblocks), dim3(256), 0, s, P, block_max);
    return hipGetLastError();
}

hipError_t launch_distance_row(int dist, const Points& P, int focus, double* out, hipStream_t s) {
    int blocks = (P.n + 255) / 256;
    if (dist == kEuclidean)
        hipLaunchKernelGGL(distance_row_kernel<kEuclidean>, dim3(blocks), dim3(256), 0, s, P, focus, out);
    else
        hipLaunchKernelGGL(distance_row_kernel<kGreatCircle>, dim3(blocks), dim3(256), 0, s, P, focus, out);
    return hipGetLastError();
}

}  // namespace gwm
