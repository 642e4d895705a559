// Work-ordered launch of a pruned single-problem sweep.
//
// A sweep workgroup covers one focus tile and one data segment.  With pruning
// and a narrow bandwidth most (tile, segment) items skip every chunk, and the
// rest differ in work by up to the segment's chunk count; in plain launch
// order the last busy workgroups start late and the sweep ends on a tail.  The
// plan counts, for every item, the data chunks whose box is within the
// zero-weight distance of the tile's box (the test the sweep applies), and
// orders the items longest first within each XCD group (items of segments
// s = g mod G go to workgroups b = g mod G, as the plain order does).  The
// order only schedules: every item writes its own partial sums, so results
// are the same bits in any order.
#include <hip/hip_runtime.h>

#include "gwm_contract.h"
#include "gwm_device.h"

namespace gwm {
namespace {

// Counts of one focus tile against every data chunk; one workgroup per tile.
// The tile's box and largest bandwidth are those the sweep forms from the
// same focus coordinates (workgroup_box).
template <int DIST>
__global__ __launch_bounds__(256) void plan_count_kernel(Points F, int n, const double* boxes, const double* bw_pt,
                                                         double bw, const double* chunk_bw, int kernel, int mixed,
                                                         int tile_pts, int seg_len, int nseg, int* counts) {
    constexpr int BD = DIST == kEuclidean ? 2 : 3;
    __shared__ int cnt[kPlanMaxSeg];
    const int tile = blockIdx.x, t = threadIdx.x;
    const int nch = (n + 63) / 64;
    GWM_GLOBAL const double* B = gptr(boxes);
    const GPoints G = global_points(F);
    for (int s = t; s < nseg; s += 256) cnt[s] = 0;
    double lo[BD], hi[BD], b = 0.0;
#pragma unroll
    for (int d = 0; d < BD; ++d) {
        lo[d] = INFINITY;
        hi[d] = -INFINITY;
    }
    const int f = tile * tile_pts + t;
    if (t < tile_pts && f < F.n) {
        const double x[3] = {DIST == kEuclidean ? G.px[f] : G.hx[f], DIST == kEuclidean ? G.py[f] : G.hy[f],
                             DIST == kEuclidean ? 0.0 : G.hz[f]};
#pragma unroll
        for (int d = 0; d < BD; ++d) lo[d] = hi[d] = x[d];
        b = bw_pt ? gptr(bw_pt)[f] : bw;
    }
    double fb;
    workgroup_box<BD>(lo, hi, b, &fb);
    if (!bw_pt) fb = bw;
    const double thr = prune_threshold<DIST>(zero_weight_distance(kernel, fb, mixed != 0));
    for (int ch = t; ch < nch; ch += 256) {
        const double th =
            chunk_bw ? prune_threshold<DIST>(zero_weight_distance(kernel, gptr(chunk_bw)[ch], mixed != 0)) : thr;
        if (!(box_gap2<DIST>(lo, hi, B + (size_t)ch * 2 * BD) > th)) atomicAdd(&cnt[(ch * 64) / seg_len], 1);
    }
    __syncthreads();
    for (int s = t; s < nseg; s += 256) counts[(size_t)tile * nseg + s] = cnt[s];
}

// Counting sort of the items by decreasing count within each of the G groups
// (group of item q: (q mod nseg) mod G; every group holds nitems / G items
// because G divides nseg): order[G i + g] = the i-th item of group g.  One
// workgroup; the order within equal counts is that of the atomics (it does
// not matter).
__global__ __launch_bounds__(1024) void plan_order_kernel(const int* counts, int nitems, int nseg, int maxc, int G,
                                                          int* order) {
    __shared__ int hist[kPlanGroups][kPlanMaxCount + 1];
    const int t = threadIdx.x;
    for (int e = t; e < kPlanGroups * (kPlanMaxCount + 1); e += 1024) (&hist[0][0])[e] = 0;
    __syncthreads();
    // counts beyond kPlanMaxCount buckets are ranked by bucket
    const int nb = min(maxc, kPlanMaxCount);
    auto bucket = [&](int c) {
        c = min(max(c, 0), maxc);
        return maxc > kPlanMaxCount ? (int)((long long)c * kPlanMaxCount / maxc) : c;
    };
    for (int q = t; q < nitems; q += 1024) atomicAdd(&hist[(q % nseg) % G][bucket(counts[q])], 1);
    __syncthreads();
    if (t < G) {
        int run = 0;
        for (int c = nb; c >= 0; --c) {
            const int h = hist[t][c];
            hist[t][c] = run;
            run += h;
        }
    }
    __syncthreads();
    for (int q = t; q < nitems; q += 1024) {
        const int c = bucket(counts[q]);
        const int g = (q % nseg) % G;
        const int pos = atomicAdd(&hist[g][c], 1);
        order[(size_t)G * pos + g] = q;
    }
}

}  // namespace

hipError_t launch_sweep_plan(int dist, const Points& F, int n, const double* boxes, const double* bw_pt, double bw,
                             const double* chunk_bw, int kernel, int mixed, int tile_pts, int seg_len, int nseg,
                             int tiles, int* counts, int* order, hipStream_t s) {
    const int maxc = seg_len / 64;
    if (nseg < 1 || nseg > kPlanMaxSeg || maxc < 1 || tiles < 1 || tile_pts % 64 != 0 ||
        tile_pts > 256 || (long long)(nseg - 1) * seg_len >= n || (long long)nseg * seg_len < n)
        return hipErrorInvalidValue;
    if ((long long)tiles * tile_pts < F.n || (long long)(tiles - 1) * tile_pts >= F.n) return hipErrorInvalidValue;
    if (dist == kEuclidean)
        hipLaunchKernelGGL(plan_count_kernel<kEuclidean>, dim3(tiles), dim3(256), 0, s, F, n, boxes, bw_pt, bw,
                           chunk_bw, kernel, mixed, tile_pts, seg_len, nseg, counts);
    else
        hipLaunchKernelGGL(plan_count_kernel<kGreatCircle>, dim3(tiles), dim3(256), 0, s, F, n, boxes, bw_pt, bw,
                           chunk_bw, kernel, mixed, tile_pts, seg_len, nseg, counts);
    const int G = nseg % kPlanGroups == 0 ? kPlanGroups : 1;
    hipLaunchKernelGGL(plan_order_kernel, dim3(1), dim3(1024), 0, s, counts, tiles * nseg, nseg, maxc, G, order);
    return hipGetLastError();
}

}  // namespace gwm
