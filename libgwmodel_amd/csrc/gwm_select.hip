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
// the candidates as among all n points.  For large bandwidths the workgroup
// select also takes a lower bound L: the mL-th smallest chunk lower bound,
// mL = ceil(floor(bw)/64), has fewer than floor(bw) points below it, so a
// chunk whose upper bound is below L holds only keys ranked below both
// wanted ones; such chunks are counted whole (their sizes lower the wanted
// ranks) and their keys never computed.
#include <algorithm>

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
__device__ __forceinline__ void key_bounds(const PTS& F, int f, GWM_GLOBAL const double* box, double* lb,
                                           double* ub) {
    if constexpr (DIST == kEuclidean) {
        const double x[2] = {F.px[f], F.py[f]};
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
        const double x[3] = {F.hx[f], F.hy[f], F.hz[f]};
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

// ---------------------------------------------------------------- wave helpers

constexpr int kWaveKeys = 1024;
constexpr int kWaveChunks = kWaveKeys / 64;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ unsigned int wave_scan(unsigned int x, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    return x;
}

// One 8-bit digit over a 256-bin histogram in LDS: the bin holding `rank`;
// updates prefix / rank and returns the bin's count.
__device__ __forceinline__ unsigned int wave_digit(const unsigned int* hist, int lane, int shift,
                                                   unsigned long long& prefix, long long& rank) {
    unsigned int c[4], sum = 0u;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        c[i] = hist[4 * lane + i];
        sum += c[i];
    }
    const unsigned int incl = wave_scan(sum, lane);
    const unsigned int excl = incl - sum;
    const bool mine = (long long)excl <= rank && rank < (long long)incl;
    const unsigned long long who = __ballot(mine);
    const int src = who ? __builtin_ctzll(who) : 63;
    int bin = 4 * lane;
    long long r = rank - (long long)excl;
    unsigned int cnt = c[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (r < (long long)c[i]) {
            cnt = c[i];
            break;
        }
        r -= c[i];
        ++bin;
    }
    bin = __shfl(bin, src);
    r = __shfl(r, src);
    cnt = __shfl(cnt, src);
    prefix |= (unsigned long long)bin << shift;
    rank = r;
    return cnt;
}

// One wave: the smallest chunk-bound key at 0-based rank `rank` among the
// first `nch` chunks (upper bounds when upper is set, else lower bounds), to
// 16 bits: returns the upper edge (upper) or lower edge (lower) of its bin.
template <int DIST, class PTS>
__device__ __forceinline__ double wave_chunk_bound(const PTS& F, int f, GWM_GLOBAL const double* boxes, int nch,
                                                   long long rank, bool upper, unsigned int* hist) {
    constexpr int BD = DIST == kEuclidean ? 4 : 6;
    const int lane = threadIdx.x & 63;
    unsigned long long prefix = 0ull;
    int shift = 64;
    for (int pass = 0; pass < 2; ++pass) {
        shift -= 8;
        const int hs = shift + 8;
#pragma unroll
        for (int i = 0; i < 4; ++i) hist[4 * lane + i] = 0u;
        wave_sync();
        for (int c = lane; c < nch; c += 64) {
            double lb, ub;
            key_bounds<DIST>(F, f, boxes + (size_t)c * BD, &lb, &ub);
            const unsigned long long key = key_of(upper ? ub : lb);
            if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & 255ull], 1u);
        }
        wave_sync();
        wave_digit(hist, lane, shift, prefix, rank);
        wave_sync();
    }
    return val_of(upper ? (prefix | ((1ull << shift) - 1ull)) : prefix);
}

// One wave: the keys at 0-based ranks `rank` and rank + 1 among keys[0..nk)
// in LDS (eight 8-bit radix digits over a 256-word LDS histogram); keys above
// the array start at `above`.
__device__ __forceinline__ void wave_rank_select(const unsigned long long* keys, int nk, long long rank,
                                                 unsigned long long above, unsigned int* hist,
                                                 unsigned long long* key0, unsigned long long* key1) {
    const int lane = threadIdx.x & 63;
    unsigned long long prefix = 0ull;
    int shift = 64;
    unsigned int cnt = 0u;
    while (shift > 0) {
        shift -= 8;
        const int hs = shift + 8;
#pragma unroll
        for (int i = 0; i < 4; ++i) hist[4 * lane + i] = 0u;
        wave_sync();
        for (int e = lane; e < nk; e += 64) {
            const unsigned long long key = keys[e];
            if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & 255ull], 1u);
        }
        wave_sync();
        cnt = wave_digit(hist, lane, shift, prefix, rank);
        wave_sync();
    }
    *key0 = prefix;
    if (rank + 1 < (long long)cnt) {
        *key1 = prefix;
        return;
    }
    unsigned long long best = above;
    for (int e = lane; e < nk; e += 64) {
        const unsigned long long key = keys[e];
        if (key > prefix && key < best) best = key;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const unsigned long long o = __shfl_xor(best, d);
        best = o < best ? o : best;
    }
    *key1 = best;
}

// One radix digit of a 256-thread select: hist[0..nb) holds the counts of
// the keys matching `prefix` above the digit; the bin holding 0-based rank
// `rank` is located by a block-wide scan of per-thread bin groups and its
// value appended to the prefix.  Results in s_prefix, s_rank (rank within
// the bin) and s_count (keys in the bin), valid after the final barrier.
__device__ __forceinline__ void resolve_digit(const unsigned int* hist, int nb, long long rank,
                                              unsigned long long prefix, int shift, unsigned long long* s_prefix,
                                              long long* s_rank, unsigned int* s_count) {
    __shared__ unsigned int s_wsum[4];
    __shared__ int s_group;
    __shared__ long long s_rem;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int per = nb / 256;
    unsigned int c = 0u;
    for (int i = 0; i < per; ++i) c += hist[tid * per + i];
    unsigned int x = c;   // inclusive scan over the wave
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wsum[wave] = x;
    if (tid == 0) {
        s_group = 255;
        s_rem = 0;
    }
    __syncthreads();
    long long off = 0;
    for (int w = 0; w < wave; ++w) off += s_wsum[w];
    const long long incl = off + x, excl = incl - c;
    if (excl <= rank && rank < incl) {
        s_group = tid;
        s_rem = rank - excl;
    }
    __syncthreads();
    if (tid == s_group) {
        long long r = s_rem;
        int bin = tid * per;
        for (int i = 0; i < per - 1; ++i, ++bin) {
            if (r < (long long)hist[bin]) break;
            r -= hist[bin];
        }
        *s_prefix = prefix | ((unsigned long long)bin << shift);
        *s_rank = r;
        *s_count = hist[bin];
    }
    __syncthreads();
}

// The interpolated threshold from the floor(bw)-th and floor(bw)+1-th
// smallest selection keys (bx = bw - floor(bw)); every path that resolves an
// adaptive bandwidth evaluates it through this one expression.
template <int DIST>
__device__ __forceinline__ double interp_threshold(unsigned long long key0, unsigned long long key1, double bx) {
    const double d0 = select_to_distance<DIST>(val_of(key0));
    const double d1 = select_to_distance<DIST>(val_of(key1));
    return d0 + (d1 - d0) * bx;
}

// Workgroup select of the keys at 0-based ranks floor(bw) - 1 and floor(bw)
// among the n selection keys of focus point f (1 <= bw < n); one thread
// calls emit(key0, key1).  All threads return together.
template <int DIST, class PTS, class EMIT>
__device__ __forceinline__ void select_keys(const PTS& F, const PTS& P, double bw, int f,
                                            GWM_GLOBAL const double* boxes, EMIT emit) {
    const int n = P.n;
    const int tid = threadIdx.x;
    constexpr int BD = DIST == kEuclidean ? 4 : 6;   // doubles per chunk box
    __shared__ unsigned int hist[kSelBins];
    __shared__ unsigned long long cand[kSelCap];
    __shared__ unsigned long long s_prefix, s_above, s_next;
    __shared__ long long s_rank;
    __shared__ unsigned int s_count, s_ncand, s_nclist, s_below;
    __shared__ double s_U, s_L;
    __shared__ int clist[kSelChunks];

    const double b0 = floor(bw);
    // pruning: candidate chunks (see the header)
    const int nch = (n + 63) / 64, nfull = n / 64;
    const long long need = (long long)b0 + 1;   // keys up to rank floor(bw) are needed
    const bool pruned = boxes && nch <= kSelChunks && need <= 64LL * nfull;
    int nclist = 0;
    long long below = 0;   // points of the chunks counted whole below the wanted keys
    double U = 0.0;
    if (pruned) {
        // U: upper edge of the 16-bit bin holding the m-th smallest chunk
        // upper bound over the full chunks (one wave)
        if (tid < 64) {
            const double u = wave_chunk_bound<DIST>(F, f, boxes, nfull, (need + 63) / 64 - 1, true, hist);
            if (tid == 0) s_U = u;
        }
        __syncthreads();
        U = s_U;
        // L: lower edge of the 16-bit bin holding the mL-th smallest chunk
        // lower bound over all chunks, mL = ceil(b0 / 64).  Fewer than
        // 64 (mL - 1) < b0 points have keys below L, so both wanted keys
        // (ranks b0 - 1 and b0) are >= L, and a chunk whose upper bound is
        // below L holds only keys ranked below them: it is counted whole.
        const long long mL = ((long long)b0 + 63) / 64;
        double L = -INFINITY;
        if (mL >= 2) {
            if (tid < 64) {
                const double l = wave_chunk_bound<DIST>(F, f, boxes, nch, mL - 1, false, hist);
                if (tid == 0) s_L = l;
            }
            __syncthreads();
            L = s_L;
        }
        if (tid == 0) {
            s_nclist = 0u;
            s_below = 0u;
        }
        __syncthreads();
        for (int c = tid; c < nch; c += 256) {
            double lb, ub;
            key_bounds<DIST>(F, f, boxes + (size_t)c * BD, &lb, &ub);
            if (ub < L) atomicAdd(&s_below, (unsigned int)min(64, n - 64 * c));
            else if (lb <= U) clist[atomicAdd(&s_nclist, 1u)] = c;
        }
        __syncthreads();
        nclist = (int)s_nclist;
        below = (long long)s_below;
    }
    // the selection keys of every candidate point j
    auto for_each_key = [&](auto&& fn) {
        if (pruned) {
            for (int e = tid; e < nclist * 64; e += 256) {
                const int j = clist[e >> 6] * 64 + (e & 63);
                if (j < n) fn(key_of(select_value<DIST>(F, f, P, j)));
            }
        } else {
            for (int j = tid; j < n; j += 256) fn(key_of(select_value<DIST>(F, f, P, j)));
        }
    };
    bool in_lds = false;
    int ncand = 0;
    if (pruned && nclist * 64 <= kSelCap) {
        // few candidates: hold all their keys in LDS (every key up to rank
        // floor(bw) is among them or counted below, so nothing lies "above" the
        // set that matters)
        if (tid == 0) {
            s_ncand = 0u;
            s_above = ~0ull;
        }
        __syncthreads();
        for_each_key([&](unsigned long long key) { cand[atomicAdd(&s_ncand, 1u)] = key; });
        __syncthreads();
        ncand = (int)s_ncand;
        in_lds = true;
    }
    // 0-based rank of d0 among the candidate keys (within the LDS keys, once gathered)
    long long rank0 = (long long)b0 - 1 - below;
    if (pruned && !in_lds && U > 0.0) {
        // many candidates: one pass of kSelBins linear bins over [0, U] (the
        // wanted keys are <= U; the map is monotone, larger keys share the last
        // bin), then gather the bin holding the rank when it fits in LDS
        const double scale = (double)kSelBins / U;
        auto lbin = [&](unsigned long long key) {
            const double b = val_of(key) * scale;
            return b < (double)(kSelBins - 1) ? (int)b : kSelBins - 1;
        };
        for (int i = tid; i < kSelBins; i += 256) hist[i] = 0u;
        __syncthreads();
        for_each_key([&](unsigned long long key) { atomicAdd(&hist[lbin(key)], 1u); });
        __syncthreads();
        resolve_digit(hist, kSelBins, rank0, 0ull, 0, &s_prefix, &s_rank, &s_count);
        const int bin = (int)s_prefix;
        if (s_count <= (unsigned int)kSelCap) {
            if (tid == 0) {
                s_ncand = 0u;
                s_above = ~0ull;
            }
            __syncthreads();
            unsigned long long above = ~0ull;
            for_each_key([&](unsigned long long key) {
                const int b = lbin(key);
                if (b == bin) cand[atomicAdd(&s_ncand, 1u)] = key;
                else if (b > bin && key < above) above = key;
            });
            atomicMin(&s_above, above);
            __syncthreads();
            ncand = (int)s_ncand;
            rank0 = s_rank;
            in_lds = true;
        }
        __syncthreads();
    }
    if (in_lds) {
        // the gathered keys: one wave resolves both order statistics
        if (tid < 64) {
            unsigned long long key0, key1;
            wave_rank_select(cand, ncand, rank0, s_above, hist, &key0, &key1);
            if (tid == 0) emit(key0, key1);
        }
        return;
    }
    if (tid == 0) {
        s_prefix = 0ull;
        s_rank = rank0;
    }
    int shift = 64;
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
                if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & mask], 1u);
            }
        } else {
            for_each_key([&](unsigned long long key) {
                if (hs == 64 || (key >> hs) == (prefix >> hs)) atomicAdd(&hist[(key >> shift) & mask], 1u);
            });
        }
        __syncthreads();
        resolve_digit(hist, nb, s_rank, prefix, shift, &s_prefix, &s_rank, &s_count);
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
    unsigned long long key1;
    if (tie) {
        key1 = key0;
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
        key1 = s_next;
    }
    if (tid == 0) emit(key0, key1);
}

template <int DIST, class PTS, class OUT>
__device__ __forceinline__ void adaptive_bw_one(const PTS& F, const PTS& P, double bw, OUT out, int f,
                                                GWM_GLOBAL const double* boxes) {
    const int n = P.n;
    const int tid = threadIdx.x;
    const double dn = bw / n;
    if (!(dn < 1)) {
        // (bw/n) * max(dist); distances are >= 0 and include the focus (0)
        __shared__ double s_red[256];
        double m = 0.0;
        for (int j = tid; j < n; j += 256) m = fmax(m, distance_f64<DIST>(F, f, P, j));
        s_red[tid] = m;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (tid < w) s_red[tid] = fmax(s_red[tid], s_red[tid + w]);
            __syncthreads();
        }
        if (tid == 0) out[f] = dn * s_red[0];
        return;
    }
    const double bx = bw - floor(bw);
    select_keys<DIST>(F, P, bw, f, boxes,
                      [&](unsigned long long key0, unsigned long long key1) {
                          out[f] = interp_threshold<DIST>(key0, key1, bx);
                      });
}

// ---------------------------------------------------------------- one wave
//
// Fast path for pruned selections with few candidates (the usual adaptive
// bandwidths, tens to hundreds of neighbours, and larger ones whose annulus
// of candidate chunks between L and U is small): one wave per focus point, no
// workgroup barriers.  U comes from two 8-bit radix digits of the chunk upper
// bounds (the upper edge of a 16-bit bin), the candidate chunks' keys (at most
// kWaveKeys, padded with the largest key) are gathered into the wave's LDS,
// and eight 8-bit radix digits resolve the two order statistics there.  A
// focus point outside this case is appended to a list that the workgroup
// select (adaptive_bw_one) then works through with a fixed grid.
// Returns false (nothing written) when the point is left to the workgroup select.
template <int DIST, class PTS, class OUT>
__device__ __forceinline__ bool adaptive_bw_wave(const PTS& F, const PTS& P, double bw, OUT out, int f,
                                                 GWM_GLOBAL const double* boxes, unsigned int* hist,
                                                 unsigned long long* keys, int* clist) {
    const int n = P.n;
    const int lane = threadIdx.x & 63;
    constexpr int BD = DIST == kEuclidean ? 4 : 6;
    const int nch = (n + 63) / 64, nfull = n / 64;
    const double b0 = floor(bw), bx = bw - b0;
    const long long need = (long long)b0 + 1;
    if (!(bw / n < 1) || !boxes || b0 < 1 || need > 64LL * nfull) return false;
    // U: upper edge of the 16-bit bin holding the m-th smallest chunk upper bound
    const double U = wave_chunk_bound<DIST>(F, f, boxes, nfull, (need + 63) / 64 - 1, true, hist);
    // L: the lower cut of adaptive_bw_one (chunks wholly below both wanted
    // keys are counted, not gathered), so that large bandwidths whose
    // candidate annulus is small also take this path
    const long long mL = ((long long)b0 + 63) / 64;
    const double L = mL >= 2 ? wave_chunk_bound<DIST>(F, f, boxes, nch, mL - 1, false, hist) : -INFINITY;
    // candidate chunks, in chunk order
    int ncl = 0, below = 0;
    bool overflow = false;
    for (int c0 = 0; c0 < nch; c0 += 64) {
        const int c = c0 + lane;
        bool take = false;
        if (c < nch) {
            double lb, ub;
            key_bounds<DIST>(F, f, boxes + (size_t)c * BD, &lb, &ub);
            if (ub < L) below += min(64, n - 64 * c);
            else take = lb <= U;
        }
        const unsigned long long m = __ballot(take);
        const int pos = ncl + __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
        if (take && pos < kWaveChunks) clist[pos] = c;
        ncl += __popcll(m);
        if (ncl > kWaveChunks) {
            overflow = true;
            break;
        }
    }
    if (overflow) return false;
    for (int m = 32; m >= 1; m >>= 1) below += __shfl_xor(below, m);
    wave_sync();
    // the candidates' keys; padding keys are the largest value
    const int nk = ncl * 64;
    for (int e = lane; e < nk; e += 64) {
        const int j = clist[e >> 6] * 64 + (e & 63);
        keys[e] = j < n ? key_of(select_value<DIST>(F, f, P, j)) : ~0ull;
    }
    wave_sync();
    unsigned long long key0, key1;
    wave_rank_select(keys, nk, (long long)b0 - 1 - below, ~0ull, hist, &key0, &key1);
    if (lane == 0) out[f] = interp_threshold<DIST>(key0, key1, bx);
    return true;
}

// LDS of the four waves of a fast-path workgroup.
struct WaveSelectShared {
    unsigned int hist[4][256];
    unsigned long long keys[4][kWaveKeys];
    int clist[4][kWaveChunks];
};

// Fast path over points 0..n-1 (entries 0..total-1 of a batch); the points
// it leaves are appended to left[] (*nleft must be zero beforehand).
template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_wave_kernel(Points F, Points P, double bw, double* out,
                                                               const double* boxes, int* left, unsigned int* nleft) {
    __shared__ WaveSelectShared sh;
    const int w = threadIdx.x >> 6;
    const int f = blockIdx.x * 4 + w;
    if (f >= F.n) return;
    if (!adaptive_bw_wave<DIST>(global_points(F), global_points(P), bw, gptr(out), f, gptr(boxes), sh.hist[w],
                                sh.keys[w], sh.clist[w]) &&
        (threadIdx.x & 63) == 0)
        left[atomicAdd(nleft, 1u)] = f;
}

template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_wave_batch_kernel(const SelectJob* jobs, const int* start,
                                                                     int njobs, int* left, unsigned int* nleft) {
    __shared__ WaveSelectShared sh;
    const int w = threadIdx.x >> 6;
    const int t = blockIdx.x * 4 + w;
    if (t >= start[njobs]) return;
    const int j = find_job(start, njobs, t);
    const SelectJob J = jobs[j];
    const GPoints P = global_points(J.P);
    if (!adaptive_bw_wave<DIST>(P, P, J.bw, gptr(J.out), J.f0 + t - start[j], gptr(J.boxes), sh.hist[w], sh.keys[w],
                                sh.clist[w]) &&
        (threadIdx.x & 63) == 0)
        left[atomicAdd(nleft, 1u)] = t;
}

// Workgroup select over the listed points (every point when the list is
// null), one point per workgroup iteration on a fixed grid.
template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_kernel(Points F, Points P, double bw, double* out,
                                                          const double* boxes, const int* left,
                                                          const unsigned int* nleft) {
    const unsigned int total = left ? *nleft : (unsigned int)F.n;
    for (unsigned int i = blockIdx.x; i < total; i += gridDim.x) {
        const int f = left ? left[i] : (int)i;
        adaptive_bw_one<DIST>(global_points(F), global_points(P), bw, gptr(out), f, gptr(boxes));
        __syncthreads();
    }
}

template <int DIST>
__global__ __launch_bounds__(256) void adaptive_bw_batch_kernel(const SelectJob* jobs, const int* start, int njobs,
                                                                const int* left, const unsigned int* nleft) {
    const unsigned int total = *nleft;
    for (unsigned int i = blockIdx.x; i < total; i += gridDim.x) {
        const int t = left[i];
        const int j = find_job(start, njobs, t);
        const SelectJob J = jobs[j];
        const GPoints P = global_points(J.P);
        adaptive_bw_one<DIST>(P, P, J.bw, gptr(J.out), J.f0 + t - start[j], gptr(J.boxes));
        __syncthreads();
    }
}

constexpr int kFallbackGrid = 2048;

// ---------------------------------------------------------------- neighbour table
//
// A golden-section search over adaptive bandwidths selects 20-40 neighbour
// counts on the same points (a multiscale backfit some hundreds), and each
// select streams the candidate keys of every focus point again.  The
// neighbour table keeps, per point, its K smallest selection keys in
// ascending order (K a power of two, at most kNbrMaxK): a select whose
// floor(bw) <= K - 1 then reads ranks floor(bw) - 1 and floor(bw) of the row.
// Row f is built by one workgroup: T = the key at 0-based rank K - 1 (the
// exact workgroup select), the keys below T gathered into LDS (fewer than K
// of them; the remaining slots hold T, the value of every rank up to K - 1
// that they do not fill), then a bitonic sort of the K slots.  The table holds
// the very keys the select returns, so lookups are bit-identical to it.
constexpr int kNbrMaxK = 8192;

// Ascending bitonic sort of a[0..K) in LDS (K a power of two), 256 threads.
__device__ __forceinline__ void lds_bitonic_sort(unsigned long long* a, int K) {
    for (int size = 2; size <= K; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = threadIdx.x; t < K / 2; t += 256) {
                const int lo = 2 * t - (t & (stride - 1));   // bit `stride` of lo is clear
                const int hi = lo + stride;
                const bool up = (lo & size) == 0;
                const unsigned long long x = a[lo], y = a[hi];
                if ((x > y) == up) {
                    a[lo] = y;
                    a[hi] = x;
                }
            }
            __syncthreads();
        }
    }
}

template <int DIST>
__global__ __launch_bounds__(256) void neighbour_table_kernel(Points Pt, const double* boxes_, int K,
                                                              unsigned long long* table) {
    __shared__ unsigned long long skeys[kNbrMaxK];
    __shared__ unsigned long long s_T;
    __shared__ unsigned int s_cnt;
    constexpr int BD = DIST == kEuclidean ? 4 : 6;
    const GPoints P = global_points(Pt);
    GWM_GLOBAL const double* boxes = gptr(boxes_);
    const int n = P.n, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nch = (n + 63) / 64;
    for (int f = blockIdx.x; f < n; f += gridDim.x) {
        select_keys<DIST>(P, P, (double)K, f, boxes, [&](unsigned long long key0, unsigned long long) { s_T = key0; });
        if (tid == 0) s_cnt = 0u;
        __syncthreads();
        const unsigned long long T = s_T;
        const double Tv = val_of(T);
        auto take = [&](int j) {
            const unsigned long long key = key_of(select_value<DIST>(P, f, P, j));
            if (key < T) {
                const unsigned int pos = atomicAdd(&s_cnt, 1u);
                if (pos < (unsigned int)K) skeys[pos] = key;
            }
        };
        if (boxes) {
            // one chunk per wave at a time; a chunk whose lower bound exceeds T holds no key below it
            for (int c = wave; c < nch; c += 4) {
                double lb, ub;
                key_bounds<DIST>(P, f, boxes + (size_t)c * BD, &lb, &ub);
                const int j = c * 64 + lane;
                if (lb <= Tv && j < n) take(j);
            }
        } else {
            for (int j = tid; j < n; j += 256) take(j);
        }
        __syncthreads();
        const int cnt = min((int)s_cnt, K);
        for (int i = cnt + tid; i < K; i += 256) skeys[i] = T;
        __syncthreads();
        lds_bitonic_sort(skeys, K);
        unsigned long long* row = table + (size_t)f * K;
        for (int i = tid; i < K; i += 256) row[i] = skeys[i];
        __syncthreads();
    }
}

template <int DIST>
__global__ __launch_bounds__(256) void neighbour_lookup_kernel(const unsigned long long* table, int K, int n,
                                                               double bw, double* out) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= n) return;
    const int r = (int)floor(bw);
    const unsigned long long* row = table + (size_t)f * K;
    out[f] = interp_threshold<DIST>(row[r - 1], row[r], bw - floor(bw));
}

// Largest distance from the focus points of workgroup column blockIdx.x to
// the data slice blockIdx.y (slice_len points); slices combine by an atomic
// maximum on the bit patterns, which order non-negative doubles as their
// values, so the result does not depend on the order.  block_max[] starts
// at zero.
template <int DIST>
__global__ __launch_bounds__(256) void max_distance_kernel(Points P, int slice_len, double* block_max) {
    __shared__ double s_red[256];
    const int n = P.n;
    const int f = blockIdx.x * 256 + threadIdx.x;
    const int j0 = blockIdx.y * slice_len, j1 = min(n, j0 + slice_len);
    double m = 0.0;
    if (f < n)
        for (int j = j0; j < j1; ++j) m = fmax(m, distance_f64<DIST>(P, f, j));
    s_red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_red[threadIdx.x] = fmax(s_red[threadIdx.x], s_red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double v = s_red[0] > 0.0 ? s_red[0] : 0.0;   // +0 for a -0 or NaN maximum
        atomicMax(reinterpret_cast<unsigned long long*>(block_max + blockIdx.x),
                  (unsigned long long)__double_as_longlong(v));
    }
}

template <int DIST>
__global__ void distance_row_kernel(Points P, int focus, double* out) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < P.n) out[j] = distance_f64<DIST>(P, focus, j);
}

}  // namespace

hipError_t launch_adaptive_bw(int dist, const Points& P, double bw, double* out, const double* boxes, int* left,
                              unsigned int* nleft, hipStream_t s, const Points* focus) {
    const Points F = focus ? *focus : P;
    const dim3 wgrid((F.n + 3) / 4), fgrid(std::min(F.n, kFallbackGrid));
    const int* lst = boxes ? left : nullptr;
    if (boxes) {
        hipError_t e = hipMemsetAsync(nleft, 0, sizeof(unsigned int), s);
        if (e != hipSuccess) return e;
    }
    if (dist == kEuclidean) {
        if (boxes)
            hipLaunchKernelGGL(adaptive_bw_wave_kernel<kEuclidean>, wgrid, dim3(256), 0, s, F, P, bw, out, boxes, left,
                               nleft);
        hipLaunchKernelGGL(adaptive_bw_kernel<kEuclidean>, fgrid, dim3(256), 0, s, F, P, bw, out, boxes, lst, nleft);
    } else {
        if (boxes)
            hipLaunchKernelGGL(adaptive_bw_wave_kernel<kGreatCircle>, wgrid, dim3(256), 0, s, F, P, bw, out, boxes,
                               left, nleft);
        hipLaunchKernelGGL(adaptive_bw_kernel<kGreatCircle>, fgrid, dim3(256), 0, s, F, P, bw, out, boxes, lst,
                           nleft);
    }
    return hipGetLastError();
}

hipError_t launch_adaptive_bw_batch(int dist, const SelectJob* jobs, const int* start, int njobs, int total,
                                    int* left, unsigned int* nleft, hipStream_t s) {
    if (total <= 0) return hipSuccess;
    // the fast path runs for every entry; jobs without boxes leave all their points to the workgroup select
    const dim3 wgrid((total + 3) / 4), fgrid(std::min(total, kFallbackGrid));
    hipError_t e = hipMemsetAsync(nleft, 0, sizeof(unsigned int), s);
    if (e != hipSuccess) return e;
    if (dist == kEuclidean) {
        hipLaunchKernelGGL(adaptive_bw_wave_batch_kernel<kEuclidean>, wgrid, dim3(256), 0, s, jobs, start, njobs, left,
                           nleft);
        hipLaunchKernelGGL(adaptive_bw_batch_kernel<kEuclidean>, fgrid, dim3(256), 0, s, jobs, start, njobs, left,
                           nleft);
    } else {
        hipLaunchKernelGGL(adaptive_bw_wave_batch_kernel<kGreatCircle>, wgrid, dim3(256), 0, s, jobs, start, njobs,
                           left, nleft);
        hipLaunchKernelGGL(adaptive_bw_batch_kernel<kGreatCircle>, fgrid, dim3(256), 0, s, jobs, start, njobs, left,
                           nleft);
    }
    return hipGetLastError();
}

hipError_t launch_neighbour_table(int dist, const Points& P, const double* boxes, int K, unsigned long long* table,
                                  hipStream_t s) {
    if (K < 2 || K > kNbrMaxK || (K & (K - 1)) || K >= P.n) return hipErrorInvalidValue;
    const dim3 grid(std::min(P.n, kFallbackGrid));
    if (dist == kEuclidean)
        hipLaunchKernelGGL(neighbour_table_kernel<kEuclidean>, grid, dim3(256), 0, s, P, boxes, K, table);
    else
        hipLaunchKernelGGL(neighbour_table_kernel<kGreatCircle>, grid, dim3(256), 0, s, P, boxes, K, table);
    return hipGetLastError();
}

hipError_t launch_neighbour_lookup(int dist, const unsigned long long* table, int K, int n, double bw, double* out,
                                   hipStream_t s) {
    if (!(bw >= 1.0) || !(floor(bw) <= K - 1) || K >= n) return hipErrorInvalidValue;
    const dim3 grid((n + 255) / 256);
    if (dist == kEuclidean)
        hipLaunchKernelGGL(neighbour_lookup_kernel<kEuclidean>, grid, dim3(256), 0, s, table, K, n, bw, out);
    else
        hipLaunchKernelGGL(neighbour_lookup_kernel<kGreatCircle>, grid, dim3(256), 0, s, table, K, n, bw, out);
    return hipGetLastError();
}

hipError_t launch_max_distance(int dist, const Points& P, double* block_max, int nblocks, hipStream_t s) {
    if (nblocks != (P.n + 255) / 256) return hipErrorInvalidValue;
    // about 2048 workgroups: the data split into slices of at least 256 points
    const int slices = std::max(1, std::min((P.n + 255) / 256, 2048 / std::max(nblocks, 1)));
    const int slice_len = (P.n + slices - 1) / slices;
    hipError_t e = hipMemsetAsync(block_max, 0, (size_t)nblocks * sizeof(double), s);
    if (e != hipSuccess) return e;
    const dim3 grid(nblocks, slices);
    if (dist == kEuclidean)
        hipLaunchKernelGGL(max_distance_kernel<kEuclidean>, grid, dim3(256), 0, s, P, slice_len, block_max);
    else
        hipLaunchKernelGGL(max_distance_kernel<kGreatCircle>, grid, dim3(256), 0, s, P, slice_len, block_max);
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
