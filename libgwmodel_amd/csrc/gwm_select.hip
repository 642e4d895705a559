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
//
// With the points in spatial order and the bounding boxes of their 64-point
// chunks (pruning), the passes stream only candidate chunks: an upper bound U
// on the floor(bw)+1-th smallest key is the m-th smallest chunk upper bound
// (m = ceil((floor(bw)+1)/64) full chunks hold at least floor(bw)+1 points no
// farther than U), and every point with a key <= U lies in a chunk whose
// lower bound is <= U.  The two order statistics are therefore the same among
// the candidates as among all n points.
#include "gwm_contract.h"

namespace gwm {

namespace {

__device__ __forceinline__ unsigned long long key_of(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ double val_of(unsigned long long k) { return __longlong_as_double((long long)k); }

constexpr int kSelBins = 4096;   // 12-bit digits
constexpr int kSelCap = 2048;    // keys gathered into LDS
constexpr int kSelChunks = 4096; // candidate chunk list capacity (pruning for n <= 262144)

// Lower and upper bounds of the selection key (squared distance, Euclidean;
// distance, great circle) between point f and the points of a chunk box.
template <int DIST, class PTS>
__device__ __forceinline__ void key_bounds(const PTS& P, int f, GWM_GLOBAL const double* box, double* lb,
                                           double* ub) {
    if constexpr (DIST == kEuclidean) {
        const double x[2] = {P.px[f], P.py[f]};
        double g2 = 0.0, h2 = 0.0;
#pragma unroll
        for (int d = 0; d < 2; ++d) {
            const double lo = box[2 * d], hi = box[2 * d + 1];
            const double g = fmax(0.0, fmax(lo - x[d], x[d] - hi));
            const double h = fmax(fabs(x[d] - lo), fabs(x[d] - hi));
            g2 += g * g;
            h2 += h * h;
        }
        *lb = g2 * (1.0 - 1e-12);
        *ub = h2 * (1.0 + 1e-12);
    } else {
        const double x[3] = {P.hx[f], P.hy[f], P.hz[f]};
        double g2 = 0.0, h2 = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double lo = box[2 * d], hi = box[2 * d + 1];
            const double g = fmax(0.0, fmax(lo - x[d], x[d] - hi));
            const double h = fmax(fabs(x[d] - lo), fabs(x[d] - hi));
            g2 += g * g;
            h2 += h * h;
        }
        // distance = a * central angle * (1 + flattening terms of at most 4f)
        const double a2 = 2.0 * 6378.137;
        *lb = 0.98 * a2 * asin(fmin(1.0, sqrt(g2)));
        *ub = 1.02 * a2 * asin(fmin(1.0, sqrt(h2)));
    }
}

template <int DIST, class PTS, class OUT>
__device__ __forceinline__ void adaptive_bw_one(const PTS& P, double bw, OUT out, int f,
                                                GWM_GLOBAL const double* boxes) {
    const int n = P.n;
    const int tid = threadIdx.x;
    constexpr int BD = DIST == kEuclidean ? 4 : 6;   // doubles per chunk box
    __shared__ unsigned int hist[kSelBins];
    __shared__ unsigned int chunk_sum[256];
    __shared__ unsigned long long cand[kSelCap];
    __shared__ unsigned long long s_prefix, s_above, s_next;
    __shared__ long long s_rank;
    __shared__ unsigned int s_count, s_ncand, s_nclist;
    __shared__ double s_red[256];
    __shared__ int clist[kSelChunks];

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
    // pruning: candidate chunks (see the header)
    const int nch = (n + 63) / 64, nfull = n / 64;
    const long long need = (long long)b0 + 1;   // keys up to rank floor(bw) are needed
    const bool pruned = boxes && nch <= kSelChunks && need <= 64LL * nfull;
    int nclist = 0;
    if (pruned) {
        // U: upper edge of the 24-bit radix bin holding the m-th smallest
        // chunk upper bound over the full chunks
        const long long m = (need + 63) / 64;
        unsigned long long prefix = 0ull;
        long long rank = m - 1;
        int shift = 64;
        for (int pass = 0; pass < 2; ++pass) {
            shift -= 12;
            const int hs = shift + 12;
            for (int i = tid; i < kSelBins; i += 256) hist[i] = 0u;
            __syncthreads();
            for (int c = tid; c < nfull; c += 256) {
                double lb, ub;
                key_bounds<DIST>(P, f, boxes + (size_t)c * BD, &lb, &ub);
                const unsigned long long key = key_of(ub);
                if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & 4095ull], 1u);
            }
            __syncthreads();
            {
                unsigned int sum = 0u;
                for (int i = 0; i < 16; ++i) sum += hist[tid * 16 + i];
                chunk_sum[tid] = sum;
            }
            __syncthreads();
            if (tid == 0) {
                long long r = rank;
                int c = 0;
                for (; c < 255; ++c) {
                    if (r < (long long)chunk_sum[c]) break;
                    r -= chunk_sum[c];
                }
                int bin = c * 16;
                for (int i = 0; i < 15; ++i, ++bin) {
                    if (r < (long long)hist[bin]) break;
                    r -= hist[bin];
                }
                s_prefix = prefix | ((unsigned long long)bin << shift);
                s_rank = r;
            }
            __syncthreads();
            prefix = s_prefix;
            rank = s_rank;
            __syncthreads();
        }
        const double U = val_of(prefix | ((1ull << shift) - 1ull));
        if (tid == 0) s_nclist = 0u;
        __syncthreads();
        for (int c = tid; c < nch; c += 256) {
            double lb, ub;
            key_bounds<DIST>(P, f, boxes + (size_t)c * BD, &lb, &ub);
            if (lb <= U) clist[atomicAdd(&s_nclist, 1u)] = c;
        }
        __syncthreads();
        nclist = (int)s_nclist;
    }
    // the selection keys of every candidate point j
    auto for_each_key = [&](auto&& fn) {
        if (pruned) {
            for (int e = tid; e < nclist * 64; e += 256) {
                const int j = clist[e >> 6] * 64 + (e & 63);
                if (j < n) fn(key_of(select_value<DIST>(P, f, j)));
            }
        } else {
            for (int j = tid; j < n; j += 256) fn(key_of(select_value<DIST>(P, f, j)));
        }
    };
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
            for_each_key([&](unsigned long long key) {
                if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & mask], 1u);
            });
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
            for_each_key([&](unsigned long long key) {
                const unsigned long long kt = key >> shift;
                if (kt == top) cand[atomicAdd(&s_ncand, 1u)] = key;
                else if (kt > top && key < above) above = key;
            });
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
            for_each_key([&](unsigned long long key) {
                if (key > key0 && key < best) best = key;
            });
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
__global__ __launch_bounds__(256) void adaptive_bw_kernel(Points P, double bw, double* out, const double* boxes) {
    adaptive_bw_one<DIST>(global_points(P), bw, gptr(out), blockIdx.x, gptr(boxes));
}

template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_batch_kernel(const SelectJob* jobs, const int* start, int njobs) {
    const int t = blockIdx.x;
    const int j = find_job(start, njobs, t);
    const SelectJob J = jobs[j];
    adaptive_bw_one<DIST>(global_points(J.P), J.bw, gptr(J.out), J.f0 + t - start[j], gptr(J.boxes));
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

hipError_t launch_adaptive_bw(int dist, const Points& P, double bw, double* out, const double* boxes, hipStream_t s) {
    if (dist == kEuclidean)
        hipLaunchKernelGGL(adaptive_bw_kernel<kEuclidean>, dim3(P.n), dim3(256), 0, s, P, bw, out, boxes);
    else
        hipLaunchKernelGGL(adaptive_bw_kernel<kGreatCircle>, dim3(P.n), dim3(256), 0, s, P, bw, out, boxes);
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
