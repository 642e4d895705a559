// Problem set-up and per-context device state of the single-problem API:
// the spatial (Hilbert) order of the points, their chunk boxes, the device
// buffers, the adaptive-bandwidth cache and the great-circle distance table.
#include <cmath>

#include "gwm_internal.h"

namespace gwm {
namespace detail {

void free_problem(gwm_hip_ctx* c) {
    auto dfree = [c](auto*& p) { cfree(c, p); };
    dfree(c->px); dfree(c->py); dfree(c->hx); dfree(c->hy_); dfree(c->hz); dfree(c->sl);
    dfree(c->X); dfree(c->y); dfree(c->Z); dfree(c->G); dfree(c->part);
    dfree(c->boxes); dfree(c->chunk_bw); dfree(c->sel_left); dfree(c->sel_nleft); dfree(c->dnn);
    for (auto& t : c->parked) {
        dfree(t.boxes);
        dfree(t.dnn);
        t = gwm_hip_ctx::DistTables{};
    }
    for (auto& e : c->bw_cache) dfree(e.buf);
    c->bw_cache.clear();
    for (auto& t : c->nbr) {
        if (t.keys) hipFree(t.keys);
        t = gwm_hip_ctx::NbrTable{};
    }
    dfree(c->bwi); dfree(c->res2); dfree(c->cond); dfree(c->sii); dfree(c->rlist); dfree(c->rcount);
    dfree(c->rpart);
    c->rpart_bytes = 0; dfree(c->ssi); dfree(c->betas); dfree(c->se);
    dfree(c->gvec); dfree(c->qd); dfree(c->red); dfree(c->blockmax); dfree(c->vx); dfree(c->vy);
    dfree(c->vZ); dfree(c->S); dfree(c->fail);
    dfree(c->Zh); dfree(c->Zl); dfree(c->Gh); dfree(c->Gl); dfree(c->vZh); dfree(c->vZl);
    dfree(c->zscale); dfree(c->gscale); dfree(c->vscale); dfree(c->colmax);
    // the plan buffers and the distance tables come from hipMalloc on every context
    release_cert_table(c);
    if (c->d2tab) hipFree(c->d2tab);
    c->d2tab = nullptr;
    c->d2tab_bytes = 0;
    c->d2_valid = false;
    if (c->plan_counts) hipFree(c->plan_counts);
    if (c->plan_order) hipFree(c->plan_order);
    c->plan_counts = nullptr;
    c->plan_order = nullptr;
    c->plan_items = 0;
    c->max_dist = -1.0;
    c->part_bytes = 0;
    c->n = 0;
}

// Download a sorted-order device array (n x cols) into a caller-order buffer.
int download_caller(gwm_hip_ctx* c, const double* dev, double* host, int cols) {
    std::vector<double> tmp((size_t)c->n * cols);
    GWM_HIP_CHECK(c, hipMemcpyAsync(tmp.data(), dev, tmp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    to_caller(c, tmp.data(), host, cols);
    return GWM_OK;
}

// f(i) for i in [0, n) on the host's cores (element-wise work, so the
// result does not depend on the split).
// Great circle: the half unit vectors (hx, hy, hz) and sin(lat) of points
// given in degrees.
void sphere_tables(size_t N, const double* lon, const double* lat, double* hx, double* hy, double* hz, double* sl) {
    const double de2ra = kPi / 180;
    parallel_for(N, [&](size_t j) {
        const double lo = lon[j] * de2ra, la = lat[j] * de2ra;
        hx[j] = 0.5 * std::cos(la) * std::cos(lo);
        hy[j] = 0.5 * std::cos(la) * std::sin(lo);
        hz[j] = 0.5 * std::sin(la);
        sl[j] = std::sin(la);
    });
}

// Position of the cell (x, y) along the Hilbert curve of a 2^bits x 2^bits
// grid (the classic rotate-and-reflect walk, one quadrant per level).
static uint64_t hilbert_index(uint32_t x, uint32_t y, int bits) {
    uint64_t d = 0;
    for (uint32_t s = 1u << (bits - 1); s > 0; s >>= 1) {
        const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
        d += (uint64_t)s * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = s - 1 - (x & (s - 1)) + (x & ~(s - 1));
                y = s - 1 - (y & (s - 1)) + (y & ~(s - 1));
            }
            const uint32_t t = x;
            x = y;
            y = t;
        }
    }
    return d;
}

// Sorted order of the points along a Hilbert curve of their planar
// coordinates (x, y, or lon, lat in degrees): consecutive points are close,
// and unlike a Morton order the curve has no long jumps, so runs of 64
// points have small bounding boxes.
std::vector<int> spatial_order(int n, const double* xs, const double* ys) {
    const int bits = 16;
    const double* v[2] = {xs, ys};
    double lo[2], hi[2];
    for (int d = 0; d < 2; ++d) {
        lo[d] = *std::min_element(v[d], v[d] + n);
        hi[d] = *std::max_element(v[d], v[d] + n);
    }
    // 32-bit curve keys, then a stable two-pass radix sort (16-bit digits):
    // the order of (key, index) pairs, ties in index order
    std::vector<uint32_t> key(n);
    const double top = (double)((1u << bits) - 1);
    parallel_for(n, [&](size_t i) {
        uint32_t q[2];
        for (int d = 0; d < 2; ++d) {
            const double span = hi[d] - lo[d];
            const double t = span > 0 ? (v[d][i] - lo[d]) / span : 0.0;
            q[d] = (uint32_t)std::min(top, std::max(0.0, std::floor(t * top + 0.5)));
        }
        key[i] = (uint32_t)hilbert_index(q[0], q[1], bits);
    });
    std::vector<int> a(n), b(n);
    for (int i = 0; i < n; ++i) a[i] = i;
    std::vector<int> count(1 << 16);
    for (int shift = 0; shift < 32; shift += 16) {
        std::fill(count.begin(), count.end(), 0);
        for (int i = 0; i < n; ++i) ++count[(key[a[i]] >> shift) & 0xffff];
        int sum = 0;
        for (int& c : count) {
            const int t = c;
            c = sum;
            sum += t;
        }
        for (int i = 0; i < n; ++i) b[count[(key[a[i]] >> shift) & 0xffff]++] = a[i];
        a.swap(b);
    }
    return a;
}

// Bounding boxes [lo_0, hi_0, lo_1, hi_1, ...] of consecutive 64-point
// chunks, followed by those of the groups of 16 consecutive chunks (1,024
// points: group g covers chunks 16 g .. 16 g + 15), which let a sweep skip a
// whole group with one test.
std::vector<double> chunk_boxes(int n, int dims, const double* const* xs) {
    const int nc = (n + 63) / 64, ng = (nc + 15) / 16;
    std::vector<double> b((size_t)(nc + ng) * 2 * dims);
    for (int ch = 0; ch < nc; ++ch) {
        const int j0 = ch * 64, j1 = std::min(n, j0 + 64);
        for (int d = 0; d < dims; ++d) {
            double mn = xs[d][j0], mx = xs[d][j0];
            for (int j = j0 + 1; j < j1; ++j) {
                mn = std::min(mn, xs[d][j]);
                mx = std::max(mx, xs[d][j]);
            }
            b[(size_t)ch * 2 * dims + 2 * d] = mn;
            b[(size_t)ch * 2 * dims + 2 * d + 1] = mx;
        }
    }
    for (int g = 0; g < ng; ++g) {
        double* gb = b.data() + (size_t)(nc + g) * 2 * dims;
        const int c1 = std::min(nc, 16 * g + 16);
        for (int d = 0; d < dims; ++d) {
            double mn = b[(size_t)16 * g * 2 * dims + 2 * d], mx = b[(size_t)16 * g * 2 * dims + 2 * d + 1];
            for (int ch = 16 * g + 1; ch < c1; ++ch) {
                mn = std::min(mn, b[(size_t)ch * 2 * dims + 2 * d]);
                mx = std::max(mx, b[(size_t)ch * 2 * dims + 2 * d + 1]);
            }
            gb[2 * d] = mn;
            gb[2 * d + 1] = mx;
        }
    }
    return b;
}


int ensure_part(gwm_hip_ctx* c, size_t bytes) {
    if (bytes <= c->part_bytes) return GWM_OK;
    cfree(c, c->part);
    c->part_bytes = 0;
    int rc = dalloc(c, reinterpret_cast<char**>(&c->part), bytes);
    if (rc) return rc;
    c->part_bytes = bytes;
    return GWM_OK;
}

int ensure_plan(gwm_hip_ctx* c, size_t items) {
    if (items <= c->plan_items) return GWM_OK;
    hipFree(c->plan_counts);
    hipFree(c->plan_order);
    c->plan_counts = nullptr;
    c->plan_order = nullptr;
    c->plan_items = 0;
    GWM_HIP_CHECK(c, hipMalloc((void**)&c->plan_counts, items * sizeof(int)));
    GWM_HIP_CHECK(c, hipMalloc((void**)&c->plan_order, items * sizeof(int)));
    c->plan_items = items;
    return GWM_OK;
}

bool stopped(gwm_hip_ctx* c) { return c->stop && c->stop(c->stop_user); }

void timer_begin(gwm_hip_ctx* c) {
    if (c->timing) hipEventRecord(c->ev0, c->stream);
}
int timer_end(gwm_hip_ctx* c, const char* name) {
    if (!c->timing) return GWM_OK;
    GWM_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    GWM_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    GWM_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    auto& s = c->stats[name];
    s.launches += 1;
    s.ms += ms;
    return GWM_OK;
}

// The squared-distance table of a great-circle problem (launch_d2_table),
// built once per outermost search / fit / backfit call.  It is optional:
// when it exceeds d2_cap, or the device cannot hold it beside 8 GiB for
// everything else, the sweeps form their distances themselves.
int ensure_d2_table(gwm_hip_ctx* c) {
    if (c->d2_valid || c->d2_scope <= 0 || c->async_mem || c->dist != kGreatCircle) return GWM_OK;
    const size_t bytes = (size_t)c->n_pad * (size_t)c->rows_pad * sizeof(float);
    if (bytes > c->d2_cap) return GWM_OK;
    if (c->d2tab_bytes < bytes) {
        if (c->d2tab) hipFree(c->d2tab);
        c->d2tab = nullptr;
        c->d2tab_bytes = 0;
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < bytes + (size_t(8) << 30)) {
            (void)hipGetLastError();
            return GWM_OK;
        }
        if (hipMalloc((void**)&c->d2tab, bytes) != hipSuccess) {
            (void)hipGetLastError();
            c->d2tab = nullptr;
            return GWM_OK;
        }
        c->d2tab_bytes = bytes;
    }
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_d2_table(c->points(), c->n, c->n_pad, c->rows_pad, c->d2tab, c->stream));
    int rc = timer_end(c, "d2_table");
    if (rc) return rc;
    c->d2_valid = true;
    return GWM_OK;
}

// Rows of the neighbour table for n points under a byte cap: the largest
// power of two K <= 8192 with K < n and n K 8 bytes <= cap (0: none).
static int neighbour_k(int n, size_t cap) {
    int K = 8192;
    while (K >= 64 && (K >= n || (size_t)n * (size_t)K * sizeof(unsigned long long) > cap)) K >>= 1;
    return K >= 64 ? K : 0;
}

// Thresholds of an adaptive bandwidth into out: from the neighbour table of
// the active distance when it answers bw, else by the select.  The table is
// built at the second select it would have answered (one search is then
// known to be under way; a single fit never pays for it).
static int select_thresholds(gwm_hip_ctx* c, double bw, double* out) {
    auto& t = c->nbr[c->dist];
    const int K = t.keys ? t.K : neighbour_k(c->n, c->nbr_cap);
    const bool fits = K > 0 && floor(bw) <= K - 1;
    if (fits && !t.keys && ++t.misses >= 2) {
        const size_t bytes = (size_t)c->n * (size_t)K * sizeof(unsigned long long);
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b >= bytes + (size_t(8) << 30) &&
            hipMalloc((void**)&t.keys, bytes) == hipSuccess) {
            t.K = K;
            timer_begin(c);
            GWM_HIP_CHECK(c, launch_neighbour_table(c->dist, c->points(), c->prune ? c->boxes : nullptr, K, t.keys,
                                                    c->stream));
            int rc = timer_end(c, "nbr_table");
            if (rc) return rc;
        } else {
            (void)hipGetLastError();
            t.keys = nullptr;
            t.misses = -(1 << 30);   // no room: do not try again for this problem
        }
    }
    if (fits && t.keys) {
        timer_begin(c);
        GWM_HIP_CHECK(c, launch_neighbour_lookup(c->dist, t.keys, t.K, c->n, bw, out, c->stream));
        return timer_end(c, "select_lookup");
    }
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_adaptive_bw(c->dist, c->points(), bw, out, c->prune ? c->boxes : nullptr, c->sel_left,
                                        c->sel_nleft, c->stream));
    return timer_end(c, "select");
}

// Resolve the bandwidth: adaptive -> per-focus thresholds (a vector of the
// bandwidth cache, filled by the select on a miss).
int resolve_bandwidth(gwm_hip_ctx* c, int adaptive, double bw, const double** bw_pt) {
    *bw_pt = nullptr;
    if (!adaptive) return GWM_OK;
    if (!(bw >= 1.0)) {
        c->err = "adaptive bandwidth must be >= 1 neighbour";
        return GWM_ERR_INVALID;
    }
    for (auto& e : c->bw_cache)
        if (e.bw == bw && e.dist == c->dist) {
            e.used = ++c->bw_clock;
            *bw_pt = e.buf;
            return GWM_OK;
        }
    // Up to 1 GiB and 512 vectors: a multiscale backfit revisits about 15
    // bandwidths per variable in every iteration, in the same cyclic order,
    // which a cache smaller than that working set misses every time.
    const size_t cap = std::min<size_t>(512, std::max<size_t>(4, (size_t(1) << 30) / ((size_t)c->n * sizeof(double))));
    gwm_hip_ctx::BwEntry* slot = nullptr;
    if (c->bw_cache.size() < cap) {
        double* buf = nullptr;
        int rc = dalloc(c, &buf, c->n);
        if (rc) return rc;
        c->bw_cache.push_back(gwm_hip_ctx::BwEntry{bw, c->dist, buf, 0});
        slot = &c->bw_cache.back();
    } else {
        slot = &c->bw_cache[0];
        for (auto& e : c->bw_cache)
            if (e.used < slot->used) slot = &e;
    }
    slot->bw = NAN;   // not valid until the select succeeds
    slot->dist = c->dist;
    int rc = select_thresholds(c, bw, slot->buf);
    if (rc) return rc;
    slot->bw = bw;
    slot->used = ++c->bw_clock;
    *bw_pt = slot->buf;
    return GWM_OK;
}

}  // namespace detail
}  // namespace gwm
