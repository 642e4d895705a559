// Fused producer + contraction: for every focus location i and every column
// c of the per-point product table Z,
//     acc[i][c] = sum_j w(d(i, j); b)^p * Z[j][c]
// without ever materialising the n x n weight matrix.  This is the
// per-location X^T W X / X^T W y (and X^T W^2 X) of the reference's local
// solves (src/gwmodelpp/GWRBasic.cpp:204-207, 241-246, 278-282, 403-406;
// src/gwmodelpp/GWRMultiscale.cpp:372-375, 433-436), recast as one dense
// contraction W * Z.
//
// FP64 path: distances and weights in FP64 on the VALU, contraction on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64).
//
// Tiling (one workgroup = 4 waves of 64 lanes):
//   - a wave owns R tiles of 16 focus points; each lane produces one weight
//     per tile and MFMA step: focus = lane % 16, data = lane / 16 (4 per step);
//   - the data range is cut into nseg segments; in a single-problem launch
//     blockIdx % nseg = segment, so consecutive workgroups (which the
//     dispatcher spreads over the 8 XCDs round-robin) stream different slices
//     of Z and each XCD's L2 keeps one slice hot;
//   - Z rows and data coordinates of a 64-point chunk are staged in LDS and
//     shared by the 4 waves.
// A batched launch covers several problems: workgroup b takes the work item
// (job, tile, segment) items[3b..3b+2]; every job of a launch has the
// instance's tile counts (nta, ntb).
// Output: partial sums out[(seg * PT + c) * n_pad + i], PT = (nta + ntb) * 16,
// reduced in a fixed order by the solve kernels (deterministic, independent
// of scheduling and of the other problems of a batch).
#include "gwm_contract.h"

namespace gwm {

namespace {

constexpr int kWaves = 4;
constexpr int kChunk = 64;

template <int DIST>
struct CoordStage;

template <>
struct CoordStage<kEuclidean> {
    double x[kChunk], y[kChunk];
};
template <>
struct CoordStage<kGreatCircle> {
    double lon[kChunk], lat[kChunk], hx[kChunk], hy[kChunk], hz[kChunk], sl[kChunk];
};

template <int NTA, int DIST, int MODE>
struct SharedTile {
    double z[kChunk][NTA * 16];
    CoordStage<DIST> c;
    KConst64 kc[MODE == kModeQDiag ? kChunk : 1];
};

template <int DIST>
struct FocusCoords {
    double a, b, hx, hy, hz, sl;
};

template <int DIST>
__device__ __forceinline__ FocusCoords<DIST> load_focus(const GPoints& P, int f) {
    FocusCoords<DIST> fc;
    fc.a = P.px[f];
    fc.b = P.py[f];
    if (DIST == kGreatCircle) {
        fc.hx = P.hx[f];
        fc.hy = P.hy[f];
        fc.hz = P.hz[f];
        fc.sl = P.sl[f];
    }
    return fc;
}

// Kernel weight of the pair (focus fc, staged point r).  The Gaussian and
// the bisquare weigh the squared distance directly (no square root; d * d of
// the reference differs from it by an ulp at most), and the bisquare and
// boxcar decide the reference's d < bw on it exactly (KConst64::d2_lt).
// Great circle: near pairs take
// gc_d2_near_f64, the others (far, coincident, near-coincident) gcdist_f64,
// in a branch that costs a wave only when one of its lanes needs it.
// The squared distance of the pair (focus fc, staged point r); *d receives
// the distance itself where the far form computes it (else -1).
template <int DIST, class SH>
__device__ __forceinline__ double pair_d2(const FocusCoords<DIST>& fc, const SH& sh, int r, double* d) {
    double d2;
    *d = -1.0;
    if constexpr (DIST == kEuclidean) {
        d2 = d2_euclid(fc.a, fc.b, sh.c.x[r], sh.c.y[r]);
    } else {
        const double ex = sh.c.hx[r] - fc.hx, ey = sh.c.hy[r] - fc.hy, ez = sh.c.hz[r] - fc.hz;
        double S = ex * ex;
        S = S + ey * ey;
        S = S + ez * ez;
        d2 = gc_d2_near_f64(S, fc.sl, sh.c.sl[r]);
        if (!(S >= kNearSMin && S <= kNearSMax)) {
            *d = gcdist_f64(fc.a, fc.b, fc.hx, fc.hy, fc.hz, fc.sl, sh.c.lon[r], sh.c.lat[r], sh.c.hx[r],
                            sh.c.hy[r], sh.c.hz[r], sh.c.sl[r]);
            d2 = *d * *d;
        }
    }
    return d2;
}

// The weight from the pair's squared distance (and its distance when known:
// else sqrt(d2), which equals it -- the correctly rounded root of a
// correctly rounded square is the number itself).
__device__ __forceinline__ double weight_of_d2(int kernel, double d2, double d, const KConst64& k) {
    if (kernel == kGaussian) {
        const double e = d2 * k.gauss;
        return e < k.cut ? 0.0 : exp(e);
    }
    if (kernel == kBisquare || kernel == kBoxcar) {
        if (!(d2 < k.d2_lt)) return 0.0;
        if (kernel == kBoxcar) return 1.0;
        const double t = 1.0 - d2 * k.inv_bw2;
        return t * t;
    }
    return kernel_fast_f64(kernel, d >= 0.0 ? d : sqrt(d2), k);
}

template <int DIST, class SH>
__device__ __forceinline__ double pair_weight(int kernel, const FocusCoords<DIST>& fc, const SH& sh, int r,
                                              const KConst64& k) {
    double d;
    const double d2 = pair_d2<DIST>(fc, sh, r, &d);
    return weight_of_d2(kernel, d2, d, k);
}

// TAB: 0 the plain sweep; for the FP64 distance table (SweepJob::f64tab) of
// a single-job great-circle sweep at a fixed bandwidth: 3 counts each item's
// chunks within reach of job.bw, 1 writes their squared distances, 2 sweeps
// at job.bw <= job.tab_bw reading them (the chunks within reach of tab_bw
// keep their block index whether or not job.bw reaches them).
// KER: the kernel id as a constant (stage 2 of Gaussian sweeps), or -1 for the job's own.
template <int NTA, int NTB, int DIST, int MODE, int R, int TAB = 0, int KER = -1>
// Two waves per SIMD: the register budget is 256 per lane.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void contract_f64_kernel(SweepLaunch L) {
    __shared__ SharedTile<NTA, DIST, MODE> sh;
    // The job is copied into registers once (uniform scalar loads): a
    // reference into memory would be re-read after every LDS store.
    int tile, seg;
    SweepJob job;
    if (L.items) {
        const int* it = L.items + 3 * blockIdx.x;
        job = L.jobs[it[0]];
        tile = it[1];
        seg = it[2];
    } else {
        job = L.job0;
        if (L.order) {
            const int q = L.order[blockIdx.x];
            seg = q % L.job0.nseg;
            tile = q / L.job0.nseg;
        } else {
            seg = blockIdx.x % L.job0.nseg;
            const int ntiles = gridDim.x / L.job0.nseg;
            tile = (int)(((long long)(blockIdx.x / L.job0.nseg) * L.tile_stride) % ntiles);
        }
    }
    // a launch groups jobs of exactly the instance's tile counts
    constexpr int nta = NTA, ntb = NTB;
    const int kernel = KER >= 0 ? KER : job.kernel;
    const GPoints P = global_points(job.P);
    GWM_GLOBAL const double* bw_pt = gptr(job.bw_pt);
    const int ldz = nta * 16;
    const int PT = (nta + ntb) * 16;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int fbase = tile * (kWaves * 16 * R) + wave * 16 * R;
    const int n = job.n;
    GWM_GLOBAL const int* focus_idx = gptr(job.focus_idx);
    const bool qtab = job.Q.n > 0;   // focus points from a table of their own
    const GPoints FP = qtab ? global_points(job.Q) : P;
    const int nf = focus_idx ? job.nfocus : (qtab ? job.Q.n : n);   // focus points of this job

    // focus points owned by this lane (one per A tile)
    FocusCoords<DIST> fc[R];
    KConst64 fbw[R];
    int fidx[R];
    double fb = 0.0;   // largest focus bandwidth of this lane
    double fdn = 0.0;  // largest nearest-point distance of this lane's focus points (CV, predict)
    GWM_GLOBAL const double* dnn = MODE != kModeQDiag ? gptr(job.dnn) : nullptr;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int f = fbase + 16 * r + (lane & 15);
        int fl = f < nf ? f : nf - 1;
        if (focus_idx) fl = focus_idx[fl];
        fidx[r] = qtab ? -1 : (focus_idx ? (f < nf ? fl : -1) : f);   // self index for the CV exclusion
        fc[r] = load_focus<DIST>(FP, fl);
        const double b = (MODE != kModeQDiag && bw_pt) ? bw_pt[fl] : job.bw;
        if (dnn) {
            const double dn = dnn[fl];
            fbw[r] = make_kconst64_nn(b, kernel, dn);
            fdn = fmax(fdn, dn);
        } else {
            fbw[r] = make_kconst64(b);
        }
        fb = fmax(fb, b);
    }
    // pruning: the workgroup's focus box and the zero-weight threshold
    constexpr int BD = DIST == kEuclidean ? 2 : 3;
    const bool prune = job.boxes != nullptr;
    double flo[BD], fhi[BD], thr = INFINITY;
    if (prune) {
#pragma unroll
        for (int d = 0; d < BD; ++d) {
            flo[d] = INFINITY;
            fhi[d] = -INFINITY;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double x[3] = {DIST == kEuclidean ? fc[r].a : fc[r].hx, DIST == kEuclidean ? fc[r].b : fc[r].hy,
                                 fc[r].hz};
#pragma unroll
            for (int d = 0; d < BD; ++d) {
                flo[d] = fmin(flo[d], x[d]);
                fhi[d] = fmax(fhi[d], x[d]);
            }
        }
        workgroup_box<BD>(flo, fhi, fb, &fb, fdn, &fdn);
        // qDiag: the streamed points' bandwidths (per chunk when adaptive)
        if (MODE != kModeQDiag) thr = prune_threshold<DIST>(zero_weight_distance(kernel, fb, false, fdn));
        else if (!job.chunk_bw) thr = prune_threshold<DIST>(zero_weight_distance(kernel, job.bw, false));
    }
    // the table's reach (stage 2; the same expression as the stages that built it)
    double thr_tab = thr;
    if constexpr (TAB == 2) thr_tab = prune_threshold<DIST>(zero_weight_distance(kernel, job.tab_bw, false, fdn));
    const int item = tile * job.nseg + seg;
    constexpr long long kBlock = (long long)kWaves * 16 * R * 64;   // doubles per table block
    long long tb = 0;                                             // the item's next table block
    bool poison = false;   // stage 2: a block past the item's count (cannot occur): NaN outputs
    const long long tcnt = TAB == 1 || TAB == 2 ? (long long)job.tab_cnt[item] : 0;
    GWM_GLOBAL double* tab = TAB == 1 || TAB == 2 ? gptr(job.f64tab) + (TAB ? job.tab_off[item] : 0) * kBlock : nullptr;

    v4d accA[R][NTA];
    v4d accB[R][NTB > 0 ? NTB : 1];
#pragma unroll
    for (int r = 0; r < R; ++r) {
#pragma unroll
        for (int t = 0; t < NTA; ++t) accA[r][t] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < (NTB > 0 ? NTB : 1); ++t) accB[r][t] = v4d{0.0, 0.0, 0.0, 0.0};
    }

    const int jbeg = seg * job.seg_len;
    const int jend = min(n, jbeg + job.seg_len);
    const int kk = lane >> 4;

    int swept = 0;   // chunks processed (not pruned)
    for (int j0 = jbeg; j0 < jend; j0 += kChunk) {
        long long blk = 0;   // the chunk's table block (stages 1 and 2)
        if (prune) {
            const int ch = j0 / kChunk;
            double t = thr;
            if (MODE == kModeQDiag && job.chunk_bw)
                t = prune_threshold<DIST>(zero_weight_distance(kernel, gptr(job.chunk_bw)[ch], false));
            const double gap = box_gap2<DIST>(flo, fhi, gptr(job.boxes) + (size_t)ch * 2 * BD);
            if constexpr (TAB == 2) {
                if (__builtin_amdgcn_readfirstlane(gap > thr_tab)) continue;
                blk = tb++;
            }
            const int skip = gap > t;
            if (__builtin_amdgcn_readfirstlane(skip)) continue;
            if constexpr (TAB == 2) {
                if (blk >= tcnt) {   // uniform: nothing is read
                    poison = true;
                    continue;
                }
            }
            if constexpr (TAB == 1) {
                blk = tb++;
                if (blk >= tcnt) continue;   // uniform; cannot occur (stage 3 counted these chunks)
            }
        }
        ++swept;
        if constexpr (TAB == 3) continue;
        // stage 2: this lane's 16 x R squared distances of the chunk, loaded
        // into registers before the Z rows are staged (the loads overlap the
        // staging and the barrier); slots of steps past a partial last chunk
        // are in the block but were not written, and their weights are masked
        double tv[TAB == 2 ? 16 : 1][R];
        if constexpr (TAB == 2) {
            GWM_GLOBAL const double* src = tab + blk * kBlock + (size_t)wave * 16 * R * 64 + lane;
#pragma unroll
            for (int s = 0; s < 16; ++s)
#pragma unroll
                for (int r = 0; r < R; ++r) tv[s][r] = src[(s * R + r) * 64];
            // the Z rows straight from memory (cached): no staging, no barrier
            GWM_GLOBAL const double* zrow = gptr(job.Z) + (size_t)j0 * ldz + (lane & 15);
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int rr = 4 * s + kk;
                const int j = j0 + rr;
                double zb[NTA];
#pragma unroll
                for (int t = 0; t < NTA; ++t) zb[t] = t < nta ? zrow[(size_t)rr * ldz + 16 * t] : 0.0;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    // a slot past the chunk's end may hold anything: it is not weighed
                    double w = j < jend ? weight_of_d2(kernel, tv[s][r], -1.0, fbw[r]) : 0.0;
                    if (j >= jend) w = 0.0;
                    if (MODE == kModeCV && j == fidx[r]) w = 0.0;
#pragma unroll
                    for (int t = 0; t < NTA; ++t)
                        if (t < nta) accA[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(w, zb[t], accA[r][t], 0, 0, 0);
                }
            }
            continue;
        }
        __syncthreads();
        // stage Z rows [j0, j0 + kChunk) (zero padded beyond n in memory)
        if constexpr (TAB != 1) {
            GWM_GLOBAL const double* src = gptr(job.Z) + (size_t)j0 * ldz;
            for (int e = threadIdx.x; e < kChunk * ldz; e += 256) sh.z[e / ldz][e % ldz] = src[e];
        }
        if (TAB != 2 && threadIdx.x < kChunk) {
            int j = j0 + threadIdx.x;
            int jl = j < n ? j : n - 1;
            if constexpr (DIST == kEuclidean) {
                sh.c.x[threadIdx.x] = P.px[jl];
                sh.c.y[threadIdx.x] = P.py[jl];
            } else {
                sh.c.lon[threadIdx.x] = P.px[jl];
                sh.c.lat[threadIdx.x] = P.py[jl];
                sh.c.hx[threadIdx.x] = P.hx[jl];
                sh.c.hy[threadIdx.x] = P.hy[jl];
                sh.c.hz[threadIdx.x] = P.hz[jl];
                sh.c.sl[threadIdx.x] = P.sl[jl];
            }
            if constexpr (MODE == kModeQDiag) sh.kc[threadIdx.x] = make_kconst64(bw_pt ? bw_pt[jl] : job.bw);
        }
        __syncthreads();
        const int steps = min(kChunk, jend - j0 + 3) / 4;
        if constexpr (TAB == 1) {
            // one table block: entry ((wave * 16 + s) * R + r) * 64 + lane
            GWM_GLOBAL double* tb_out = tab + blk * kBlock + (size_t)wave * 16 * R * 64 + lane;
            for (int s = 0; s < steps; ++s) {
                const int rr = 4 * s + kk;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    double d;
                    tb_out[(s * R + r) * 64] = pair_d2<DIST>(fc[r], sh, rr, &d);
                }
            }
            continue;
        }
        for (int s = 0; s < steps; ++s) {
            const int rr = 4 * s + kk;    // row of the chunk this lane weighs
            const int j = j0 + rr;
            double zb[NTA];
#pragma unroll
            for (int t = 0; t < NTA; ++t) zb[t] = t < nta ? sh.z[rr][16 * t + (lane & 15)] : 0.0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                // qDiag with a fixed bandwidth: fbw holds the same constants as sh.kc
                double w = pair_weight<DIST>(kernel, fc[r], sh, rr,
                                             (MODE == kModeQDiag && bw_pt) ? sh.kc[rr] : fbw[r]);
                if (j >= jend) w = 0.0;
                if (MODE == kModeCV && j == fidx[r]) w = 0.0;
                double wa = MODE == kModeQDiag ? w * w : w;
#pragma unroll
                for (int t = 0; t < NTA; ++t)
                    if (t < nta) accA[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(wa, zb[t], accA[r][t], 0, 0, 0);
                if (NTB > 0) {
                    double w2 = w * w;
#pragma unroll
                    for (int t = 0; t < NTB; ++t)
                        if (t < ntb)
                            accB[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(w2, zb[t], accB[r][t], 0, 0, 0);
                }
            }
        }
    }

    if constexpr (TAB == 3) {
        if (threadIdx.x == 0) job.tab_cnt[item] = swept;
        return;
    }
    if constexpr (TAB == 1) return;
    if (job.pairs && threadIdx.x == 0 && swept)
        atomicAdd(job.pairs, (unsigned long long)swept * (unsigned long long)(kWaves * 16 * R) * 64ull);

    if (poison)
#pragma unroll
        for (int r = 0; r < R; ++r) {
#pragma unroll
            for (int t = 0; t < NTA; ++t) accA[r][t] = v4d{NAN, NAN, NAN, NAN};
        }
    // D layout of v_mfma_f64_16x16x4f64 (checked on gfx950): lane holds
    // column (lane % 16), rows (lane / 16) + 4 q, q = 0..3.
    const size_t np = (size_t)job.n_pad;
    GWM_GLOBAL double* out = gptr(job.out) + (size_t)seg * PT * np;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f0 = fbase + 16 * r + kk - job.out_off;
#pragma unroll
        for (int t = 0; t < NTA; ++t) {
            if (t >= nta) continue;
            GWM_GLOBAL double* o = out + (size_t)(16 * t + (lane & 15)) * np + f0;
#pragma unroll
            for (int q = 0; q < 4; ++q) o[4 * q] = accA[r][t][q];
        }
#pragma unroll
        for (int t = 0; t < NTB; ++t) {
            if (t >= ntb) continue;
            GWM_GLOBAL double* o = out + (size_t)(ldz + 16 * t + (lane & 15)) * np + f0;
#pragma unroll
            for (int q = 0; q < 4; ++q) o[4 * q] = accB[r][t][q];
        }
    }
}

template <int NTA, int NTB>
constexpr int r_of() {
    return (NTA + NTB) <= 8 ? 2 : 1;
}

template <int NTA, int NTB, int DIST, int MODE, int R>
hipError_t launch_f64_rr(const SweepLaunch& a, hipStream_t s) {
    int grid;
    if (a.items) {
        grid = a.nitems;
    } else {
        const int nf = a.job0.focus_idx ? a.job0.nfocus : (a.job0.Q.n > 0 ? a.job0.Q.n : a.job0.n);
        int tiles = (nf + kWaves * 16 * R - 1) / (kWaves * 16 * R);
        grid = tiles * a.job0.nseg;
    }
    if (grid <= 0) return hipSuccess;
    if (a.order && (a.items || a.norder != grid)) return hipErrorInvalidValue;
    SweepLaunch b = a;
    if (!a.items) b.tile_stride = a.linear_order ? 1 : tile_stride(grid / a.job0.nseg);
    hipLaunchKernelGGL((contract_f64_kernel<NTA, NTB, DIST, MODE, R>), dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
}

// The default shape, or one focus tile per wave when asked for (single jobs).
template <int NTA, int NTB, int DIST, int MODE>
hipError_t launch_f64_r(const SweepLaunch& a, hipStream_t s) {
    constexpr int RD = r_of<NTA, NTB>();
    if constexpr (RD == 2) {
        if (a.shape_r == 1) return launch_f64_rr<NTA, NTB, DIST, MODE, 1>(a, s);
    }
    return launch_f64_rr<NTA, NTB, DIST, MODE, RD>(a, s);
}

// Table stages of a single-job great-circle CV / AICc sweep (default shape).
template <int NTA, int MODE, int TAB>
hipError_t launch_f64_tab(const SweepLaunch& a, hipStream_t s) {
    constexpr int R = r_of<NTA, 0>();
    const int tiles = (a.job0.n + kWaves * 16 * R - 1) / (kWaves * 16 * R);
    const int grid = tiles * a.job0.nseg;
    if (grid <= 0) return hipSuccess;
    if (a.order && a.norder != grid) return hipErrorInvalidValue;
    SweepLaunch b = a;
    b.tile_stride = a.linear_order ? 1 : tile_stride(tiles);
    if (TAB == 2 && a.job0.kernel == kGaussian)
        hipLaunchKernelGGL((contract_f64_kernel<NTA, 0, kGreatCircle, MODE, R, TAB, TAB == 2 ? kGaussian : -1>),
                           dim3(grid), dim3(256), 0, s, b);
    else
        hipLaunchKernelGGL((contract_f64_kernel<NTA, 0, kGreatCircle, MODE, R, TAB>), dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
}

template <int MODE, int TAB>
hipError_t dispatch_tab(int nta, const SweepLaunch& a, hipStream_t s) {
    switch (nta) {
    case 1: return launch_f64_tab<1, MODE, TAB>(a, s);
    case 2: return launch_f64_tab<2, MODE, TAB>(a, s);
    case 3: return launch_f64_tab<3, MODE, TAB>(a, s);
    case 4: return launch_f64_tab<4, MODE, TAB>(a, s);
    case 5: return launch_f64_tab<5, MODE, TAB>(a, s);
    case 6: return launch_f64_tab<6, MODE, TAB>(a, s);
    case 7: return launch_f64_tab<7, MODE, TAB>(a, s);
    case 8: return launch_f64_tab<8, MODE, TAB>(a, s);
    case 9: return launch_f64_tab<9, MODE, TAB>(a, s);
    case 10: return launch_f64_tab<10, MODE, TAB>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int DIST, int MODE>
hipError_t dispatch_ab(int nta, int ntb, const SweepLaunch& a, hipStream_t s) {
#define GWM_CASE(A, B) \
    if (nta == A && ntb == B) return launch_f64_r<A, B, DIST, MODE>(a, s);
    if (MODE == kModeHat) {
        GWM_CASE(1, 1) GWM_CASE(2, 1) GWM_CASE(2, 2) GWM_CASE(3, 2) GWM_CASE(3, 3) GWM_CASE(4, 3)
        GWM_CASE(5, 4) GWM_CASE(5, 5) GWM_CASE(6, 5) GWM_CASE(7, 6) GWM_CASE(8, 7) GWM_CASE(9, 8)
        GWM_CASE(10, 9)
    } else {
        GWM_CASE(1, 0) GWM_CASE(2, 0) GWM_CASE(3, 0) GWM_CASE(4, 0) GWM_CASE(5, 0) GWM_CASE(6, 0)
        GWM_CASE(7, 0) GWM_CASE(8, 0) GWM_CASE(9, 0) GWM_CASE(10, 0)
    }
#undef GWM_CASE
    return hipErrorInvalidValue;
}

}  // namespace

int tile_stride(int tiles) {
    if (tiles <= 2) return 1;
    int s = (int)(tiles * 0.6180339887498949) | 1;
    auto gcd = [](int a, int b) {
        while (b) {
            const int t = a % b;
            a = b;
            b = t;
        }
        return a;
    };
    while (gcd(s, tiles) != 1) s += 2;
    return s % tiles;
}


namespace {

// Single-column sweeps of the multiscale path (GWRMultiscale.cpp:420-482,
// 560-636): per focus point the two sums sum_j w_ij x_j^2 and sum_j w_ij x_j
// y_j, in FP64 on the VALU.  Four lanes per focus point, 16 focus points per
// wave: lane group q = lane / 16 weighs the points q, q + 4, q + 8, ... of
// each chunk, and the four partial sums combine in a fixed order at the end.
// One lane per focus point would leave the chip short of waves at multiscale
// sizes (n = 50,000: 782 waves for 1,024 SIMDs, each stalled on its own FP64
// dependences); four give 3,125.  The waves are independent: each tests its
// own 16-point focus box against the boxes of the 16-chunk groups and of the
// chunks, stages its live chunks in its own LDS slice (wave-level syncs, no
// workgroup barrier), and loads the next live chunk into registers while it
// weighs the current one.  Same weights, pruning and data segments as
// contract_f64_kernel, whose 16-column matrix-core form would compute eight
// times the columns these sweeps use.
template <int DIST>
struct VarShared {
    CoordStage<DIST> c;
    double z0[kChunk], z1[kChunk];
};

// one staged point per lane: its coordinates and its two table entries
template <int DIST>
struct VarPoint {
    double c[DIST == kEuclidean ? 2 : 6];
    double z0, z1;
};

constexpr int kVarFocus = 64;   // focus points per workgroup (16 per wave)

__device__ __forceinline__ void var_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int DIST, int MODE, int KER>
__global__ __launch_bounds__(256) void contract_var_f64_kernel(SweepJob job) {
    static_assert(MODE == kModeCV || MODE == kModePlain, "single-column sweeps weigh with w");
    __shared__ VarShared<DIST> shw[4];
    const int seg = blockIdx.x % job.nseg;
    const int tile = blockIdx.x / job.nseg;
    const GPoints P = global_points(job.P);
    GWM_GLOBAL const double* bw_pt = gptr(job.bw_pt);
    GWM_GLOBAL const double* dnn = gptr(job.dnn);
    const int n = job.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4;
    VarShared<DIST>& sh = shw[wave];
    const int f = tile * kVarFocus + wave * 16 + (lane & 15);
    const int fl = f < n ? f : n - 1;
    const FocusCoords<DIST> fc = load_focus<DIST>(P, fl);
    const double b = bw_pt ? bw_pt[fl] : job.bw;
    double fdn = 0.0;
    KConst64 kc;
    if (dnn) {
        fdn = dnn[fl];
        kc = make_kconst64_nn(b, KER, fdn);
    } else {
        kc = make_kconst64(b);
    }
    // the wave's box: its 16 focus points sit in lanes 0-15 and again in each
    // group of 16, so a reduction over the low four lane bits covers them
    constexpr int BD = DIST == kEuclidean ? 2 : 3;
    const bool prune = job.boxes != nullptr;
    double wlo[BD], whi[BD], wthr = INFINITY;
    if (prune) {
        const double x[3] = {DIST == kEuclidean ? fc.a : fc.hx, DIST == kEuclidean ? fc.b : fc.hy, fc.hz};
        double wb = b, wdn = fdn;
#pragma unroll
        for (int d = 0; d < BD; ++d) wlo[d] = whi[d] = x[d];
#pragma unroll
        for (int m = 8; m >= 1; m >>= 1) {
#pragma unroll
            for (int d = 0; d < BD; ++d) {
                wlo[d] = fmin(wlo[d], __shfl_xor(wlo[d], m));
                whi[d] = fmax(whi[d], __shfl_xor(whi[d], m));
            }
            wb = fmax(wb, __shfl_xor(wb, m));
            wdn = fmax(wdn, __shfl_xor(wdn, m));
        }
        wthr = prune_threshold<DIST>(zero_weight_distance(KER, wb, false, wdn));
    }
    const int jbeg = seg * job.seg_len;
    const int jend = min(n, jbeg + job.seg_len);
    const int cbeg = jbeg / kChunk, cend = (jend + kChunk - 1) / kChunk;
    GWM_GLOBAL const double* boxes = gptr(job.boxes);
    // the boxes of the 16-chunk groups follow the n / 64 chunk boxes
    GWM_GLOBAL const double* gboxes = prune ? boxes + (size_t)((n + kChunk - 1) / kChunk) * 2 * BD : nullptr;
    GWM_GLOBAL const double* Z = gptr(job.Z);
    const int ldz = job.nta * 16;
    // the first chunk at or after c that the wave weighs (cend: none); the
    // tests are uniform over the wave
    auto next_live = [&](int c) {
        while (c < cend) {
            if (!prune) return c;
            if ((c & 15) == 0 || c == cbeg) {
                if (box_gap2<DIST>(wlo, whi, gboxes + (size_t)(c >> 4) * 2 * BD) > wthr) {
                    c = (c | 15) + 1;
                    continue;
                }
            }
            if (!(box_gap2<DIST>(wlo, whi, boxes + (size_t)c * 2 * BD) > wthr)) return c;
            ++c;
        }
        return cend;
    };
    auto load_point = [&](int c) {
        VarPoint<DIST> v;
        const int j = c * kChunk + lane;
        const int jl = j < n ? j : n - 1;
        v.c[0] = P.px[jl];
        v.c[1] = P.py[jl];
        if constexpr (DIST == kGreatCircle) {
            v.c[2] = P.hx[jl];
            v.c[3] = P.hy[jl];
            v.c[4] = P.hz[jl];
            v.c[5] = P.sl[jl];
        }
        // rows past n are zero in the table (padded to a multiple of 64)
        const size_t row = (size_t)j * ldz;
        v.z0 = Z[row];
        v.z1 = Z[row + 1];
        return v;
    };
    auto store_point = [&](const VarPoint<DIST>& v) {
        if constexpr (DIST == kEuclidean) {
            sh.c.x[lane] = v.c[0];
            sh.c.y[lane] = v.c[1];
        } else {
            sh.c.lon[lane] = v.c[0];
            sh.c.lat[lane] = v.c[1];
            sh.c.hx[lane] = v.c[2];
            sh.c.hy[lane] = v.c[3];
            sh.c.hz[lane] = v.c[4];
            sh.c.sl[lane] = v.c[5];
        }
        sh.z0[lane] = v.z0;
        sh.z1[lane] = v.z1;
    };
    double a0 = 0.0, a1 = 0.0;
    int swept = 0;
    int cur = next_live(cbeg);
    if (cur < cend) {
        store_point(load_point(cur));
        var_wave_sync();
    }
    while (cur < cend) {
        const int nxt = next_live(cur + 1);
        VarPoint<DIST> v;
        if (nxt < cend) v = load_point(nxt);   // in flight while the current chunk is weighed
        ++swept;
        const int j0 = cur * kChunk;
        // all 64 slots, uniform over the wave: segments are whole chunks, so
        // only the last chunk of the data is partial, and its slots past n
        // hold zero table rows (finite weights times zero)
#pragma unroll 4
        for (int st = 0; st < kChunk / 4; ++st) {
            const int e = 4 * st + q;
            double w = pair_weight<DIST>(KER, fc, sh, e, kc);
            if (MODE == kModeCV && j0 + e == f) w = 0.0;
            a0 = fma(w, sh.z0[e], a0);
            a1 = fma(w, sh.z1[e], a1);
        }
        var_wave_sync();
        if (nxt < cend) store_point(v);
        var_wave_sync();
        cur = nxt;
    }
    if (job.pairs && lane == 0 && swept)
        atomicAdd(job.pairs, (unsigned long long)swept * 16ull * 64ull);
    // (s_0 + s_1) + (s_2 + s_3) in every lane of the focus point (addition
    // commutes bit for bit, so both partners of each step hold the same sum)
    a0 += __shfl_xor(a0, 16);
    a1 += __shfl_xor(a1, 16);
    a0 += __shfl_xor(a0, 32);
    a1 += __shfl_xor(a1, 32);
    if (q == 0 && f < n) {
        GWM_GLOBAL double* out = gptr(job.out) + (size_t)seg * 2 * job.n_pad;
        out[f] = a0;
        out[job.n_pad + f] = a1;
    }
}

// the kernel id as a constant: the per-pair kernel switch folds away
template <int DIST, int MODE>
hipError_t launch_var_k(const SweepJob& j, hipStream_t s) {
    const dim3 grid((j.n + kVarFocus - 1) / kVarFocus * j.nseg);
    switch (j.kernel) {
    case kGaussian: hipLaunchKernelGGL((contract_var_f64_kernel<DIST, MODE, kGaussian>), grid, dim3(256), 0, s, j); break;
    case kExponential:
        hipLaunchKernelGGL((contract_var_f64_kernel<DIST, MODE, kExponential>), grid, dim3(256), 0, s, j);
        break;
    case kBisquare: hipLaunchKernelGGL((contract_var_f64_kernel<DIST, MODE, kBisquare>), grid, dim3(256), 0, s, j); break;
    case kTricube: hipLaunchKernelGGL((contract_var_f64_kernel<DIST, MODE, kTricube>), grid, dim3(256), 0, s, j); break;
    default: hipLaunchKernelGGL((contract_var_f64_kernel<DIST, MODE, kBoxcar>), grid, dim3(256), 0, s, j); break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_contract_var_f64(int dist, int mode, const SweepJob& j, hipStream_t s) {
    if (j.nta != 1 || j.ntb != 0 || j.focus_idx || j.Q.n > 0 || j.n <= 0 || j.nseg <= 0 || j.n_pad < j.n ||
        (mode != kModeCV && mode != kModePlain))
        return hipErrorInvalidValue;
    if (j.kernel < kGaussian || j.kernel > kBoxcar) return hipErrorInvalidValue;
    if (dist == kEuclidean) {
        if (mode == kModeCV) return launch_var_k<kEuclidean, kModeCV>(j, s);
        return launch_var_k<kEuclidean, kModePlain>(j, s);
    }
    if (mode == kModeCV) return launch_var_k<kGreatCircle, kModeCV>(j, s);
    return launch_var_k<kGreatCircle, kModePlain>(j, s);
}

hipError_t launch_contract_f64_table(int stage, int mode, int nta, const SweepLaunch& a, hipStream_t s,
                                     long long* block) {
    const SweepJob& j = a.job0;
    if (a.items || j.focus_idx || j.Q.n > 0 || j.bw_pt || !j.boxes || j.ntb != 0 || j.nta != nta || nta < 1 ||
        nta > 10 || (mode != kModeCV && mode != kModePlain) || stage < 1 || stage > 3)
        return hipErrorInvalidValue;
    if (!j.tab_cnt) return hipErrorInvalidValue;
    if (stage != 3 && (!j.f64tab || !j.tab_off)) return hipErrorInvalidValue;
    if (stage == 2 && !(j.bw <= j.tab_bw)) return hipErrorInvalidValue;
    if (block) *block = (long long)kWaves * 16 * (nta <= 8 ? 2 : 1) * 64;
    if (mode == kModeCV) {
        if (stage == 1) return dispatch_tab<kModeCV, 1>(nta, a, s);
        if (stage == 2) return dispatch_tab<kModeCV, 2>(nta, a, s);
        return dispatch_tab<kModeCV, 3>(nta, a, s);
    }
    if (stage == 1) return dispatch_tab<kModePlain, 1>(nta, a, s);
    if (stage == 2) return dispatch_tab<kModePlain, 2>(nta, a, s);
    return dispatch_tab<kModePlain, 3>(nta, a, s);
}

hipError_t launch_contract_f64(int dist, int mode, int nta, int ntb, const SweepLaunch& a, hipStream_t s) {
    if (dist == kEuclidean) {
        if (mode == kModeCV) return dispatch_ab<kEuclidean, kModeCV>(nta, ntb, a, s);
        if (mode == kModePlain) return dispatch_ab<kEuclidean, kModePlain>(nta, ntb, a, s);
        if (mode == kModeHat) return dispatch_ab<kEuclidean, kModeHat>(nta, ntb, a, s);
        return dispatch_ab<kEuclidean, kModeQDiag>(nta, ntb, a, s);
    } else {
        if (mode == kModeCV) return dispatch_ab<kGreatCircle, kModeCV>(nta, ntb, a, s);
        if (mode == kModePlain) return dispatch_ab<kGreatCircle, kModePlain>(nta, ntb, a, s);
        if (mode == kModeHat) return dispatch_ab<kGreatCircle, kModeHat>(nta, ntb, a, s);
        return dispatch_ab<kGreatCircle, kModeQDiag>(nta, ntb, a, s);
    }
}

}  // namespace gwm
