// Adaptive bandwidth resolution and distance utilities.
//
// BandwidthWeight::weight (src/gwmodelpp/spatialweight/BandwidthWeight.cpp:40-55)
// sorts the n distances of a focus point and interpolates between the
// floor(bw)-th and floor(bw)+1-th smallest (the focus itself has rank 1):
//     fixbw = d0 + (d1 - d0) * (bw - floor(bw)),
// or uses (bw/n) * max(dist) when bw >= n.  Here the two order statistics are
// found without sorting, by an MSD radix select over the IEEE-754 bit patterns
// of the non-negative selection values, which order like the values
// themselves.  One workgroup per focus point; the histogram lives in LDS.
// Digits are 12 bits (sign + exponent, then the top mantissa bits), then 8.
// Passes stream the n distances (recomputed each time) until the bin holding
// the wanted rank has at most kCap keys; one more pass gathers that bin into
// LDS (and the smallest key above it), and the remaining digits run on the
// LDS copy.  Typically 2-3 streaming passes instead of one per byte.  The
// result is bit-identical to the sort because the selected values are the
// very doubles the sort would return.
#include "gwm_contract.h"

namespace gwm {

namespace {

__device__ __forceinline__ unsigned long long key_of(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ double val_of(unsigned long long k) { return __longlong_as_double((long long)k); }

constexpr int kSelBins = 4096;   // 12-bit digits
constexpr int kSelCap = 2048;    // keys gathered into LDS

template <int DIST, class PTS, class OUT>
__device__ __forceinline__ void adaptive_bw_one(const PTS& P, double bw, OUT out, int f) {
    const int n = P.n;
    const int tid = threadIdx.x;
    __shared__ unsigned int hist[kSelBins];
    __shared__ unsigned int chunk_sum[256];
    __shared__ unsigned long long cand[kSelCap];
    __shared__ unsigned long long s_prefix, s_above, s_next;
    __shared__ long long s_rank;
    __shared__ unsigned int s_count, s_ncand;
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
    int shift = 64;
    bool in_lds = false;
    int ncand = 0;
    while (shift > 0) {
        const int width = shift > 40 ? 12 : 8;   // 12, 12, then 8 x 5
        shift -= width;
        const int nb = 1 << width;
        const int hs = shift + width;           // bits above this digit (64: none)
        for (int i = tid; i < nb; i += 256) hist[i] = 0u;
        __syncthreads();
        const unsigned long long prefix = s_prefix;
        const unsigned long long mask = (unsigned long long)(nb - 1);
        if (in_lds) {
            for (int i = tid; i < ncand; i += 256) {
                const unsigned long long key = cand[i];
                if ((key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & mask], 1u);
            }
        } else {
            for (int j = tid; j < n; j += 256) {
                const unsigned long long key = key_of(select_value<DIST>(P, f, j));
                if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & mask], 1u);
            }
        }
        __syncthreads();
        // the bin holding rank s_rank: per-thread chunk sums, then a scan by thread 0
        const int per = nb / 256;
        {
            unsigned int sum = 0u;
            for (int i = 0; i < per; ++i) sum += hist[tid * per + i];
            chunk_sum[tid] = sum;
        }
        __syncthreads();
        if (tid == 0) {
            long long rank = s_rank;
            int c = 0;
            for (; c < 255; ++c) {
                if (rank < (long long)chunk_sum[c]) break;
                rank -= chunk_sum[c];
            }
            int bin = c * per;
            for (int i = 0; i < per - 1; ++i, ++bin) {
                if (rank < (long long)hist[bin]) break;
                rank -= hist[bin];
            }
            s_prefix = prefix | ((unsigned long long)bin << shift);
            s_rank = rank;
            s_count = hist[bin];
        }
        __syncthreads();
        if (!in_lds && shift > 0 && s_count <= (unsigned int)kSelCap) {
            // gather the bin (its keys share the top 64 - shift bits) and the
            // smallest key above it
            if (tid == 0) {
                s_ncand = 0u;
                s_above = ~0ull;
            }
            __syncthreads();
            const unsigned long long top = s_prefix >> shift;
            unsigned long long above = ~0ull;
            for (int j = tid; j < n; j += 256) {
                const unsigned long long key = key_of(select_value<DIST>(P, f, j));
                const unsigned long long kt = key >> shift;
                if (kt == top) cand[atomicAdd(&s_ncand, 1u)] = key;
                else if (kt > top && key < above) above = key;
            }
            atomicMin(&s_above, above);
            __syncthreads();
            ncand = (int)s_ncand;
            in_lds = true;
        }
    }
    // all 64 bits resolved: key0 = s_prefix, s_count keys equal it
    const unsigned long long key0 = s_prefix;
    const bool tie = (s_rank + 1) < (long long)s_count;
    double d1v;
    if (tie) {
        d1v = val_of(key0);
    } else {
        // second order statistic: the smallest key above key0
        if (tid == 0) s_next = in_lds ? s_above : ~0ull;
        __syncthreads();
        unsigned long long best = ~0ull;
        if (in_lds) {
            for (int i = tid; i < ncand; i += 256) {
                const unsigned long long key = cand[i];
                if (key > key0 && key < best) best = key;
            }
        } else {
            for (int j = tid; j < n; j += 256) {
                const unsigned long long key = key_of(select_value<DIST>(P, f, j));
                if (key > key0 && key < best) best = key;
            }
        }
        atomicMin(&s_next, best);
        __syncthreads();
        d1v = val_of(s_next);
    }
    if (tid == 0) {
        const double d0 = select_to_distance<DIST>(val_of(key0));
        const double d1 = select_to_distance<DIST>(d1v);
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
