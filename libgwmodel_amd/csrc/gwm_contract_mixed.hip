// Mixed-precision fused producer + contraction on the FP16 matrix cores.
//
// Same contraction as gwm_contract.hip (acc[i][c] = sum_j w_ij^p Z[j][c]),
// with the arithmetic split as follows:
//   - coordinate differences and the squared distance (Euclidean) or the
//     half-chord S (great circle) in FP64 -- nearby points cancel there;
//   - the transcendental tail of the distance and the kernel weight in FP32;
//   - Z split into FP16 hi + lo parts, lo scaled by 2^11 so small values
//     stay normal: z = zh + zl * 2^-11 (each part carries 11 significant
//     bits, the pair ~22);
//   - AICc and CV sweeps (w^1): the weight is formed as W = 2^11 w (the
//     exponential kernels add 11 to the exponent) and split W = Wh + Wl with
//     Wl the unscaled FP16 residual, so one packed conversion gives Wh and a
//     subtraction and a second packed conversion Wl (Wh normal down to
//     w = 2^-25, Wl's absolute error 2^-22 W or 2^-25);
//     W z = Wh zh + Wl zh + 2^-11 (Wh zl + Wl zl), the first two in separate
//     FP32 accumulators, the 2^-11 terms in a third, all undone at the FP64
//     flush;
//   - hat and qDiag sweeps (w and w^2): w = wh + wl 2^-11 with the lo part
//     scaled, w z = wh zh + 2^-11 (wh zl + wl zh) + 2^-22 wl zl, the three
//     scales in separate accumulators;
//   - every 64 data points the FP32 accumulators are flushed into FP64 ones,
//     so FP32 rounding only ever acts on 64-term partial sums;
//   - Z columns carry a power-of-two scale (exact) that keeps them inside the
//     FP16 range; it is undone when the FP64 sums are written.
// Tiling: a wave owns R tiles of 16 focus points; per MFMA step a lane forms 8
// weights (focus = lane % 16, data = 8 * (lane / 16) + e).  Z planes of a
// 64-point chunk are staged in LDS with a padded column stride (40 halfs) so
// the 16-byte B-operand reads of a quarter-wave hit distinct banks.
#include "gwm_contract.h"

// Mixed precision tolerates (and profits from) fused multiply-adds: let the
// compiler contract a*b+c throughout this translation unit.
#pragma clang fp contract(fast)

namespace gwm {

namespace {

constexpr int kWaves = 4;
constexpr int kChunkM = 64;        // data points per chunk (2 MFMA steps)
constexpr int kColStride = 40;     // LDS halfs per column per 32-row block
constexpr float kLoScale = 2048.0f;

template <int DIST>
struct MCoords;
template <>
struct MCoords<kEuclidean> {
    double x[kChunkM], y[kChunkM];
};
template <>
struct MCoords<kGreatCircle> {
    double hx[kChunkM], hy[kChunkM], hz[kChunkM];
    float sl[kChunkM];   // sin(lat): only enters the flattening term
};
template <>
struct MCoords<kGcTable> {   // distances come from the precomputed table
    int unused;
};

// Kernel weight from the squared distance (Euclidean) or from the distance.
struct BwConst {
    float bw, inv_bw, bw2, inv_bw2, gauss;   // gauss = -log2(e) / (2 bw^2)
};

__device__ __forceinline__ BwConst make_bw(float b) {
    BwConst c;
    c.bw = b;
    c.inv_bw = 1.0f / b;
    c.bw2 = b * b;
    c.inv_bw2 = 1.0f / (b * b);
    c.gauss = -1.4426950408889634f * 0.5f * c.inv_bw2;
    return c;
}

template <int NTA, int DIST, int MODE>
struct alignas(16) MShared {
    // table sweeps stage the next chunk's planes while the current one is in
    // use: two buffers
    static constexpr int kBuf = DIST == kGcTable ? 2 : 1;
    _Float16 zh[kBuf][2][NTA * 16][kColStride];
    _Float16 zl[kBuf][2][NTA * 16][kColStride];
    MCoords<DIST> c;
    BwConst bwc[kBuf][MODE == kModeQDiag ? kChunkM : 1];   // qDiag: constants of each data point's bandwidth
    // table sweeps: the segment's chunks that pruning keeps (at most 128:
    // segments hold at most 8192 points), in order
    unsigned char live[DIST == kGcTable ? 128 : 1];
    int nlive;
};

template <int DIST>
struct MFocus {
    double a, b, c;
    float sl;
};

template <int DIST>
__device__ __forceinline__ MFocus<DIST> load_mfocus(const GPoints& P, int f) {
    MFocus<DIST> fc;
    if constexpr (DIST == kEuclidean) {
        fc.a = P.px[f];
        fc.b = P.py[f];
        fc.c = 0.0;
        fc.sl = 0.0f;
    } else {
        fc.a = P.hx[f];
        fc.b = P.hy[f];
        fc.c = P.hz[f];
        fc.sl = (float)P.sl[f];
    }
    return fc;
}

// sqrt(d2) inside one branch of a kernel switch.  The exponential and tricube
// branches both take the root of the same d2; without the opaque copy the
// compiler hoists that common root (with its denormal scaling, about ten
// instructions) above the switch, where every kernel pays for it.
__device__ __forceinline__ float branch_sqrt(float d2) {
    asm volatile("" : "+v"(d2));
    return __builtin_sqrtf(d2);
}

__device__ __forceinline__ float weight_from_d2(int kernel, float d2, const BwConst& c) {
    switch (kernel) {
    case kGaussian: return __builtin_amdgcn_exp2f(d2 * c.gauss);
    case kExponential: return __builtin_amdgcn_exp2f(-1.4426950408889634f * branch_sqrt(d2) * c.inv_bw);
    case kBisquare: {
        float t = 1.0f - d2 * c.inv_bw2;
        return d2 < c.bw2 ? t * t : 0.0f;
    }
    case kTricube: {
        float q = branch_sqrt(d2) * c.inv_bw;
        float t = 1.0f - q * q * q;
        return d2 < c.bw2 ? (t * t) * t : 0.0f;
    }
    default: return d2 < c.bw2 ? 1.0f : 0.0f;
    }
}

// Squared great-circle distance (km^2) from the half-chord S = sin^2(w),
// w = half the central angle, for S <= 0.25 (central angle <= 60 deg), in the
// identities of gcdist_f64:
//   w^2 = S q2(S),  q2 = asin(sqrt(S))^2 / S   (minimax, degree 5, 1.4e-8),
//   3R = 3 sinc(2w)                              (minimax in S, degree 3,
//        4.2e-6: R only enters the flattening term, f/8 (3R +- 1) < 2e-3),
//   d = 2 a w (1 + f/8 [(3R-1) sp^2 / C - (3R+1) sm^2 / S]),  C = 1 - S,
// with sp = sin lat_i + sin lat_j and sm = sin lat_j - sin lat_i in FP32
// (an error of 2^-24 sin lat in sm moves d by f 2^-24 / sqrt(S) relative,
// and the weight's exponent by that times (d / bw)^2); 1/C and 1/S come from
// one reciprocal of S*C, clamped so that coincident points (S = 0, sm = 0)
// give exactly 0.  No square root, two transcendentals per pair including
// the kernel's exponential.  Larger S takes gcdist_tail_f32 (the producer
// runs it in a second, wave-uniform pass).  Coefficients from
// mpmath.chebyfit on [0, 1/4].
__device__ __forceinline__ float gc_d2_near_f32(float S, float sp, float sm) {
    const float a2 = (float)(2.0 * 6378.137);                       // 2a
    const float c8 = (float)(2.0 * 6378.137 / 298.257223563 / 8.0);  // 2a f / 8
    const float q2 = fmaf(fmaf(fmaf(fmaf(fmaf(0.1182195246219635f, S, 0.06318631768226624f), S,
                                         0.11677248775959015f), S, 0.17762747406959534f), S,
                               0.33333659172058105f), S, 1.0f);
    const float R3 = fmaf(fmaf(fmaf(-0.1778002232313156f, S, -0.25808843970298767f), S, -2.0004422664642334f), S,
                          3.0000035762786865f);
    const float C = 1.0f - S;
    const float rsc = __builtin_amdgcn_rcpf(fmaxf(S * C, 1e-30f));
    // (3R - 1) sp^2 / C - (3R + 1) sm^2 / S = rsc [(3R - 1) S sp^2 - (3R + 1) C sm^2]
    const float inner = fmaf(fmaf(R3, S, -S), sp * sp, -(fmaf(R3, C, C) * (sm * sm)));
    const float g2a = fmaf(c8 * rsc, inner, a2);   // 2a (1 + f/8 rsc inner)
    return S * (q2 * (g2a * g2a));
}

// Squared great-circle distances of one focus point to the 8 data points
// row(0..7) (each (hx, hy, hz, sin lat)): the near form for S <= 1/4, then,
// in a pass the wave takes only when one of its pairs needs it, the general
// tail for pairs beyond 60 degrees of arc.  The distance table
// (d2_table_kernel) forms its entries here; the great-circle sweep writes the
// same expressions out in its own loop, so both give the same roundings.
struct GcRow {
    double hx, hy, hz;
    float sl;
};
template <class ROW>
__device__ __forceinline__ void gc_d2_8(double fa, double fb, double fcz, float fsl, ROW row, float* d2) {
    float smax = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const GcRow q = row(e);
        const double ex = q.hx - fa, ey = q.hy - fb, ez = q.hz - fcz;
        const float S = (float)(ex * ex + ey * ey + ez * ez);
        d2[e] = gc_d2_near_f32(S, fsl + q.sl, q.sl - fsl);
        smax = fmaxf(smax, S);
    }
    if (__any(smax > 0.25f)) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const GcRow q = row(e);
            const double ex = q.hx - fa, ey = q.hy - fb, ez = q.hz - fcz;
            const float S = (float)(ex * ex + ey * ey + ez * ez);
            if (S > 0.25f) {
                const float d = gcdist_tail_f32(S, fsl, q.sl);
                d2[e] = d * d;
            }
        }
    }
}

// Kernel weights of 8 squared distances (the kernel switch is taken once),
// times 2^11 when SCALED (added to the exponent of the exponential kernels).
template <bool SCALED>
__device__ __forceinline__ void weights8_from_d2(int kernel, const float* d2, const BwConst& c, float* w) {
    constexpr float e0 = SCALED ? 11.0f : 0.0f;
    constexpr float m0 = SCALED ? kLoScale : 1.0f;
    switch (kernel) {
    case kGaussian:
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = __builtin_amdgcn_exp2f(fmaf(d2[e], c.gauss, e0));
        break;
    case kExponential:
#pragma unroll
        for (int e = 0; e < 8; ++e)
            w[e] = __builtin_amdgcn_exp2f(fmaf(-1.4426950408889634f * branch_sqrt(d2[e]), c.inv_bw, e0));
        break;
    case kBisquare:
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float t = 1.0f - d2[e] * c.inv_bw2;
            w[e] = d2[e] < c.bw2 ? (t * t) * m0 : 0.0f;
        }
        break;
    case kTricube:
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float q = branch_sqrt(d2[e]) * c.inv_bw;
            float t = 1.0f - q * q * q;
            w[e] = d2[e] < c.bw2 ? ((t * t) * t) * m0 : 0.0f;
        }
        break;
    default:
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = d2[e] < c.bw2 ? m0 : 0.0f;
        break;
    }
}

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));

// W = Wh + Wl for two weights: one packed conversion to FP16 for the hi
// parts, the unscaled residuals from that same packed value (the empty
// register constraint keeps the compiler from forming the hi parts twice
// with different roundings), a second packed conversion for the lo parts.
__device__ __forceinline__ void split2_scaled(float w0, float w1, h2v& hi, h2v& lo) {
    hi = __builtin_convertvector((f2v){w0, w1}, h2v);
    asm volatile("" : "+v"(hi));
    const f2v back = __builtin_convertvector(hi, f2v);
    lo = __builtin_convertvector((f2v){w0 - back.x, w1 - back.y}, h2v);
}

__device__ __forceinline__ void split16(float w, _Float16& hi, _Float16& lo) {
    hi = (_Float16)w;
    lo = (_Float16)((w - (float)hi) * kLoScale);
}

// w * w rounded to FP32 once.  The instruction selector otherwise folds the
// product into one of its FP16 conversions (v_fma_mixlo_f16 rounds the exact
// product to FP16) but not into the other (v_cvt_pk_f16_f32 of the FP32
// product), so the hi part fed to the matrix core and the hi part the lo
// residual is taken against come from two different roundings, and hi + lo
// no longer represents w^2 (measured: up to 1e-3 relative error in sums of
// w^2 z).  The empty register constraint makes the FP32 product one opaque
// value that both conversions read.
__device__ __forceinline__ float square_rounded(float w) {
    float w2 = w * w;
    asm volatile("" : "+v"(w2));
    return w2;
}

// Table sweeps stream their distances from HBM: three waves per SIMD (at most
// 168 registers) hide more of that latency.  Asked for the AICc and CV
// sweeps of up to three column tiles, which fit that cap without spilling
// (164 and 167 registers at three tiles); the others keep the default.
template <int NTA, int NTB, int DIST, int MODE>
constexpr int mixed_min_waves() {
    return (DIST == kGcTable && (MODE == kModePlain || MODE == kModeCV) && NTB == 0 && NTA <= 3) ? 3 : 1;
}

template <int NTA, int NTB, int DIST, int MODE, int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(mixed_min_waves<NTA, NTB, DIST, MODE>())))
void contract_mixed_kernel(SweepLaunch L) {
    // the distance table sweeps great-circle data (boxes, focus coordinates)
    constexpr bool kTab = DIST == kGcTable;
    constexpr int GEO = kTab ? kGreatCircle : DIST;
    __shared__ MShared<NTA, DIST, MODE> sh;
    // The job is copied into registers once (uniform scalar loads): a
    // reference into memory would be re-read after every LDS store.
    int tile, seg;
    SweepJob args;
    if (L.items) {
        const int* it = L.items + 3 * blockIdx.x;
        args = L.jobs[it[0]];
        tile = it[1];
        seg = it[2];
    } else {
        args = L.job0;
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
    // AICc / CV sweeps weigh with w: the scaled split of W = 2^11 w
    constexpr bool kScaledSplit = (MODE == kModePlain || MODE == kModeCV) && NTB == 0;
    const GPoints P = global_points(args.P);
    GWM_GLOBAL const double* bw_pt = gptr(args.bw_pt);
    const int ldz = nta * 16;
    const int PT = (nta + ntb) * 16;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int fbase = tile * (kWaves * 16 * R) + wave * 16 * R;
    const int n = args.n;
    GWM_GLOBAL const int* focus_idx = gptr(args.focus_idx);
    const bool qtab = args.Q.n > 0;   // focus points from a table of their own
    const GPoints FP = qtab ? global_points(args.Q) : P;
    const int nf = focus_idx ? args.nfocus : (qtab ? args.Q.n : n);   // focus points of this job
    const int kq = lane >> 4;       // which 8-row group of the 32-row step
    const int col = lane & 15;

    MFocus<GEO> fc[R];
    BwConst fbw[R];
    int fidx[R];
    double fb = 0.0;   // largest focus bandwidth of this lane
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int f = fbase + 16 * r + col;
        int fl = f < nf ? f : nf - 1;
        if (focus_idx) fl = focus_idx[fl];
        fidx[r] = qtab ? -1 : (focus_idx ? (f < nf ? fl : -1) : f);   // self index for the CV exclusion
        fc[r] = load_mfocus<GEO>(FP, fl);
        const double b = (MODE != kModeQDiag && bw_pt) ? bw_pt[fl] : args.bw;
        fbw[r] = make_bw((float)b);
        fb = fmax(fb, b);
    }
    // pruning: the workgroup's focus box and the zero-weight threshold
    constexpr int BD = GEO == kEuclidean ? 2 : 3;
    const bool prune = args.boxes != nullptr;
    double flo[BD], fhi[BD], thr = INFINITY;
    if (prune) {
#pragma unroll
        for (int d = 0; d < BD; ++d) {
            flo[d] = INFINITY;
            fhi[d] = -INFINITY;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double x[3] = {fc[r].a, fc[r].b, fc[r].c};
#pragma unroll
            for (int d = 0; d < BD; ++d) {
                flo[d] = fmin(flo[d], x[d]);
                fhi[d] = fmax(fhi[d], x[d]);
            }
        }
        workgroup_box<BD>(flo, fhi, fb, &fb);
        // qDiag: the streamed points' bandwidths (per chunk when adaptive)
        if (MODE != kModeQDiag) thr = prune_threshold<GEO>(zero_weight_distance(args.kernel, fb, true));
        else if (!args.chunk_bw) thr = prune_threshold<GEO>(zero_weight_distance(args.kernel, args.bw, true));
    }

    v4f hh[R][NTA + NTB], xx[R][NTA + NTB], ll[R][NTA + NTB];
    v4d acc[R][NTA + NTB];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int t = 0; t < NTA + NTB; ++t) acc[r][t] = v4d{0.0, 0.0, 0.0, 0.0};

    const int jbeg = seg * args.seg_len;
    const int jend = min(n, jbeg + args.seg_len);

    int swept = 0;   // chunks processed (not pruned)
    // table sweeps: a chunk's two 16 x 32 distance blocks per focus tile (rows
    // are padded to 64, so both exist even in the last chunk)
    v4f tq[R][2][2], nq[R][2][2];
    // Table sweeps run a software pipeline.  The pruning tests of the whole
    // segment run once (wave 0, into the live list); then, per live chunk, the
    // next live chunk's distances and FP16 planes are loaded into registers
    // before the current chunk is computed, and stored into the other LDS
    // buffer after it: memory latency overlaps a chunk of work, with one
    // barrier per chunk.  (Vector-memory loads complete in issue order, so
    // loads issued just before a wait would hold back an earlier prefetch.)
    int nlive = 0;
    constexpr int kZv = (2 * NTA * 16 * 4 + 255) / 256;   // 16-byte vectors per thread and plane
    const int vecs = 2 * ldz * 4;
    v4u zhq[kZv], zlq[kZv];
    double bwq = 0.0;
    auto load_next = [&](int j) {
        GWM_GLOBAL const float* tab = gptr(args.d2tab);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const size_t fg = (size_t)((fbase + 16 * r) >> 4);
            GWM_GLOBAL const float* b = tab + (fg * args.d2_ndb + (size_t)(j >> 5)) * 512 + 4 * lane;
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
                nq[r][s2][0] = *reinterpret_cast<GWM_GLOBAL const v4f*>(b + 512 * s2);
                nq[r][s2][1] = *reinterpret_cast<GWM_GLOBAL const v4f*>(b + 512 * s2 + 256);
            }
        }
        const size_t base = (size_t)(j / 32) * ldz * 32;
        GWM_GLOBAL const v4u* gh = reinterpret_cast<GWM_GLOBAL const v4u*>(gptr(args.Zh) + base);
        GWM_GLOBAL const v4u* gl = reinterpret_cast<GWM_GLOBAL const v4u*>(gptr(args.Zl) + base);
#pragma unroll
        for (int v = 0; v < kZv; ++v) {
            const int e = threadIdx.x + 256 * v;
            if (e < vecs) {
                zhq[v] = gh[e];
                zlq[v] = gl[e];
            }
        }
        if constexpr (MODE == kModeQDiag) {
            if (threadIdx.x < kChunkM) {
                const int jj = j + threadIdx.x;
                bwq = bw_pt ? bw_pt[jj < n ? jj : n - 1] : args.bw;
            }
        }
    };
    auto store_next = [&](int zb) {
#pragma unroll
        for (int v = 0; v < kZv; ++v) {
            const int e = threadIdx.x + 256 * v;
            if (e < vecs) {
                const int blk = e / (ldz * 4);
                const int rem = e % (ldz * 4);
                const int c = rem >> 2, part = rem & 3;
                *reinterpret_cast<v4u*>(&sh.zh[zb][blk][c][part * 8]) = zhq[v];
                *reinterpret_cast<v4u*>(&sh.zl[zb][blk][c][part * 8]) = zlq[v];
            }
        }
        if constexpr (MODE == kModeQDiag) {
            if (threadIdx.x < kChunkM) sh.bwc[zb][threadIdx.x] = make_bw((float)bwq);
        }
    };
    if constexpr (kTab) {
        if (wave == 0) {
            const int c0 = jbeg / kChunkM, nch = (jend - jbeg + kChunkM - 1) / kChunkM;   // nch <= 128
            int cnt = 0;
            for (int b = 0; b < nch; b += 64) {
                const int ci = b + lane;
                int ok = ci < nch;
                if (ok && prune) {
                    double t = thr;
                    if (MODE == kModeQDiag && args.chunk_bw)
                        t = prune_threshold<GEO>(zero_weight_distance(args.kernel, gptr(args.chunk_bw)[c0 + ci], true));
                    ok = !(box_gap2<GEO>(flo, fhi, gptr(args.boxes) + (size_t)(c0 + ci) * 2 * BD) > t);
                }
                const unsigned long long m = __ballot(ok);
                if (ok) sh.live[cnt + __popcll(m & ((1ull << lane) - 1ull))] = (unsigned char)ci;
                cnt += __popcll(m);
            }
            if (lane == 0) sh.nlive = cnt;
        }
        __syncthreads();
        nlive = sh.nlive;
        if (nlive > 0) {
            load_next(jbeg + kChunkM * sh.live[0]);
            store_next(0);
        }
        __syncthreads();
    }
    int zb = 0;
    for (int i = 0;; ++i) {
        int j0;
        bool more = false;
        if constexpr (kTab) {
            if (i >= nlive) break;
            j0 = jbeg + kChunkM * sh.live[i];
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int s2 = 0; s2 < 2; ++s2) {
                    tq[r][s2][0] = nq[r][s2][0];
                    tq[r][s2][1] = nq[r][s2][1];
                }
            more = i + 1 < nlive;
            if (more) load_next(jbeg + kChunkM * sh.live[i + 1]);
        } else {
            j0 = jbeg + kChunkM * i;
            if (j0 >= jend) break;
            if (prune) {
                const int ch = j0 / kChunkM;
                double t = thr;
                if (MODE == kModeQDiag && args.chunk_bw)
                    t = prune_threshold<GEO>(zero_weight_distance(args.kernel, gptr(args.chunk_bw)[ch], true));
                const int skip = box_gap2<GEO>(flo, fhi, gptr(args.boxes) + (size_t)ch * 2 * BD) > t;
                if (__builtin_amdgcn_readfirstlane(skip)) continue;
            }
            __syncthreads();
            // stage the two 32-row blocks of both planes: global [blk][col][32]
            {
                const size_t base = (size_t)(j0 / 32) * ldz * 32;
                GWM_GLOBAL const v4u* gh = reinterpret_cast<GWM_GLOBAL const v4u*>(gptr(args.Zh) + base);
                GWM_GLOBAL const v4u* gl = reinterpret_cast<GWM_GLOBAL const v4u*>(gptr(args.Zl) + base);
                for (int e = threadIdx.x; e < vecs; e += 256) {
                    const int blk = e / (ldz * 4);
                    const int rem = e % (ldz * 4);
                    const int c = rem >> 2, part = rem & 3;
                    *reinterpret_cast<v4u*>(&sh.zh[0][blk][c][part * 8]) = gh[e];
                    *reinterpret_cast<v4u*>(&sh.zl[0][blk][c][part * 8]) = gl[e];
                }
            }
            if (threadIdx.x < kChunkM) {
                int j = j0 + threadIdx.x;
                int jl = j < n ? j : n - 1;
                if constexpr (DIST == kEuclidean) {
                    sh.c.x[threadIdx.x] = P.px[jl];
                    sh.c.y[threadIdx.x] = P.py[jl];
                } else {
                    sh.c.hx[threadIdx.x] = P.hx[jl];
                    sh.c.hy[threadIdx.x] = P.hy[jl];
                    sh.c.hz[threadIdx.x] = P.hz[jl];
                    sh.c.sl[threadIdx.x] = (float)P.sl[jl];
                }
                if constexpr (MODE == kModeQDiag)
                    sh.bwc[0][threadIdx.x] = make_bw((float)(bw_pt ? bw_pt[jl] : args.bw));
            }
            __syncthreads();
        }
        ++swept;
#pragma unroll
            for (int r = 0; r < R; ++r)
    #pragma unroll
                for (int t = 0; t < NTA + NTB; ++t) {
                    hh[r][t] = v4f{0.f, 0.f, 0.f, 0.f};
                    xx[r][t] = v4f{0.f, 0.f, 0.f, 0.f};
                    ll[r][t] = v4f{0.f, 0.f, 0.f, 0.f};
                }
            const int nsteps = (jend - j0) > 32 ? 2 : 1;
            for (int s = 0; s < nsteps; ++s) {
                // B operands: column col of each tile, rows 8*kq .. 8*kq+7 of block s
                v8h bh[NTA], bl[NTA];
    #pragma unroll
                for (int t = 0; t < NTA; ++t) {
                    if (t < nta) {
                        bh[t] = *reinterpret_cast<const v8h*>(&sh.zh[zb][s][16 * t + col][8 * kq]);
                        bl[t] = *reinterpret_cast<const v8h*>(&sh.zl[zb][s][16 * t + col][8 * kq]);
                    }
                }
    #pragma unroll
                for (int r = 0; r < R; ++r) {
                    v8h ah, al, a2h, a2l;
                    const int rb = 32 * s + 8 * kq;
                    float d2[8], w[8];
                    if constexpr (DIST == kEuclidean) {
    #pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const int rr = rb + e;
                            double dx = sh.c.x[rr] - fc[r].a, dy = sh.c.y[rr] - fc[r].b;
                            d2[e] = (float)(dx * dx + dy * dy);
                        }
                    } else if constexpr (DIST == kGreatCircle) {
                        // the arithmetic of gc_d2_8, written out: the shared form
                        // costs this kernel 14 VGPRs and its second wave per SIMD
                        float smax = 0.0f;   // any pair beyond 60 degrees of arc (S > 0.25)?
    #pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const int rr = rb + e;
                            double ex = sh.c.hx[rr] - fc[r].a, ey = sh.c.hy[rr] - fc[r].b, ez = sh.c.hz[rr] - fc[r].c;
                            const float slj = sh.c.sl[rr];
                            const float S = (float)(ex * ex + ey * ey + ez * ez);
                            d2[e] = gc_d2_near_f32(S, fc[r].sl + slj, slj - fc[r].sl);
                            smax = fmaxf(smax, S);
                        }
                        if (__any(smax > 0.25f)) {
    #pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const int rr = rb + e;
                                double ex = sh.c.hx[rr] - fc[r].a, ey = sh.c.hy[rr] - fc[r].b,
                                       ez = sh.c.hz[rr] - fc[r].c;
                                const float S = (float)(ex * ex + ey * ey + ez * ez);
                                if (S > 0.25f) {
                                    const float d = gcdist_tail_f32(S, fc[r].sl, sh.c.sl[rr]);
                                    d2[e] = d * d;
                                }
                            }
                        }
                    } else {
    #pragma unroll
                        for (int e = 0; e < 8; ++e) d2[e] = tq[r][s][e >> 2][e & 3];
                    }
                    // qDiag with a fixed bandwidth: fbw holds the same constants as sh.bwc
                    if (MODE == kModeQDiag && bw_pt) {
    #pragma unroll
                        for (int e = 0; e < 8; ++e) w[e] = weight_from_d2(args.kernel, d2[e], sh.bwc[zb][rb + e]);
                    } else {
                        weights8_from_d2<kScaledSplit>(args.kernel, d2, fbw[r], w);
                    }
                    // Rows past the segment end occur only in the last segment,
                    // at j >= n, where every Z table (design, G, single column)
                    // and its FP16 planes are zero padded: their products are
                    // exact zeros without masking the (finite) weights.
                    if constexpr (kScaledSplit) {
    #pragma unroll
                        for (int e = 0; e < 8; e += 2) {
                            float w0 = w[e], w1 = w[e + 1];
                            if (MODE == kModeCV) {
                                const int self = fidx[r] - (j0 + rb);
                                if (self == e) w0 = 0.0f;
                                if (self == e + 1) w1 = 0.0f;
                            }
                            h2v h, l;
                            split2_scaled(w0, w1, h, l);
                            ah[e] = h.x;
                            ah[e + 1] = h.y;
                            al[e] = l.x;
                            al[e + 1] = l.y;
                        }
                    } else {
    #pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            float we = w[e];
                            const float wa = MODE == kModeQDiag ? square_rounded(we) : we;
                            _Float16 h, l;
                            split16(wa, h, l);
                            ah[e] = h;
                            al[e] = l;
                            if (NTB > 0) {
                                split16(square_rounded(we), h, l);
                                a2h[e] = h;
                                a2l[e] = l;
                            }
                        }
                    }
    #pragma unroll
                    for (int t = 0; t < NTA; ++t) {
                        if (t >= nta) continue;
                        hh[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[t], hh[r][t], 0, 0, 0);
                        xx[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[t], xx[r][t], 0, 0, 0);
                        if constexpr (kScaledSplit) {
                            // Wl zh at scale 1 in its own accumulator, Wl zl with the 2^-11 terms
                            ll[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[t], ll[r][t], 0, 0, 0);
                            xx[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bl[t], xx[r][t], 0, 0, 0);
                        } else {
                            xx[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[t], xx[r][t], 0, 0, 0);
                            ll[r][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bl[t], ll[r][t], 0, 0, 0);
                        }
                    }
    #pragma unroll
                    for (int t = 0; t < NTB; ++t) {
                        if (t >= ntb) continue;
                        hh[r][NTA + t] =
                            __builtin_amdgcn_mfma_f32_16x16x32_f16(a2h, bh[t], hh[r][NTA + t], 0, 0, 0);
                        xx[r][NTA + t] =
                            __builtin_amdgcn_mfma_f32_16x16x32_f16(a2h, bl[t], xx[r][NTA + t], 0, 0, 0);
                        xx[r][NTA + t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a2l, bh[t], xx[r][NTA + t], 0, 0, 0);
                        ll[r][NTA + t] =
                            __builtin_amdgcn_mfma_f32_16x16x32_f16(a2l, bl[t], ll[r][NTA + t], 0, 0, 0);
                    }
                }
            }
            // flush the 64-point partial sums into FP64
    #pragma unroll
            for (int r = 0; r < R; ++r)
    #pragma unroll
                for (int t = 0; t < NTA + NTB; ++t)
    #pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        // hh + ll in FP32: one rounding per 64-point block, next to the
                        // 64 of the block's accumulation
                        if constexpr (kScaledSplit)
                            acc[r][t][q] += ((double)(hh[r][t][q] + ll[r][t][q]) +
                                             (double)xx[r][t][q] * (1.0 / 2048.0)) * (1.0 / 2048.0);
                        else
                            acc[r][t][q] += (double)hh[r][t][q] + (double)xx[r][t][q] * (1.0 / 2048.0) +
                                            (double)ll[r][t][q] * (1.0 / 4194304.0);
                    }
        if constexpr (kTab) {
            if (more) store_next(zb ^ 1);
            __syncthreads();
            zb ^= 1;
        }
    }

    if (args.pairs && threadIdx.x == 0 && swept)
        atomicAdd(args.pairs, (unsigned long long)swept * (unsigned long long)(kWaves * 16 * R) * 64ull);

    // D layout of v_mfma_f32_16x16x32_f16 (checked on gfx950): lane holds
    // column (lane % 16), rows 4 * (lane / 16) + q.
    const size_t np = (size_t)args.n_pad;
    GWM_GLOBAL double* out = gptr(args.out) + (size_t)seg * PT * np;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int f0 = fbase + 16 * r + 4 * kq;
#pragma unroll
        for (int t = 0; t < NTA + NTB; ++t) {
            const bool isB = t >= NTA;
            const int tt = isB ? t - NTA : t;
            if (tt >= (isB ? ntb : nta)) continue;
            const int zc = 16 * tt + col;
            const double unscale = gptr(args.colscale)[zc];
            GWM_GLOBAL double* o = out + (size_t)((isB ? ldz : 0) + zc) * np + (f0 - args.out_off);
#pragma unroll
            for (int q = 0; q < 4; ++q) o[q] = acc[r][t][q] * unscale;
        }
    }
}

template <int NTA, int NTB, int DIST, int MODE, int R>
hipError_t launch_mixed_rr(const SweepLaunch& a, hipStream_t s) {
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
    hipLaunchKernelGGL((contract_mixed_kernel<NTA, NTB, DIST, MODE, R>), dim3(grid), dim3(256), 0, s, b);
    return hipGetLastError();
}

// Default focus tiles per wave: two where the two-tile form fits in 256
// VGPRs (two waves per SIMD), i.e. up to three column tiles except the CV and
// qDiag sweeps on great-circle data with three; one elsewhere, where the
// two-tile form would run one wave per SIMD and the one-tile form two (e.g.
// Euclidean, four column tiles: 304 against 188 VGPRs; great-circle CV,
// three: 268 against 212).  Hat-mode sweeps keep two up to six column tiles:
// measured faster there despite one wave per SIMD (great circle, 3 + 3
// tiles: 5.0 against 5.6 ms).  Both forms give the same bits.
constexpr int mixed_default_r(int nta, int ntb, int mode, int dist) {
    // table sweeps hold a chunk of distances per focus tile and prefetch the
    // next: one tile per wave keeps 3 waves per SIMD (127 VGPRs at 3 column
    // tiles, AICc), where two would leave one
    if (dist == kGcTable) return 1;
    if (mode == kModeHat) return nta + ntb <= 6 ? 2 : 1;
    return (nta + ntb <= 3 && !(dist == kGreatCircle && (mode == kModeCV || mode == kModeQDiag) && nta >= 3)) ? 2
                                                                                                              : 1;
}
template <int NTA, int NTB, int DIST, int MODE>
constexpr int mixed_r() {
    return mixed_default_r(NTA, NTB, MODE, DIST);
}

// The default shape, or one focus tile per wave when asked for (single jobs).
template <int NTA, int NTB, int DIST, int MODE>
hipError_t launch_mixed_r(const SweepLaunch& a, hipStream_t s) {
    constexpr int RD = mixed_r<NTA, NTB, DIST, MODE>();
    if constexpr (RD == 2) {
        if (a.shape_r == 1) return launch_mixed_rr<NTA, NTB, DIST, MODE, 1>(a, s);
    }
    return launch_mixed_rr<NTA, NTB, DIST, MODE, RD>(a, s);
}

template <int DIST, int MODE>
hipError_t dispatch_mixed(int nta, int ntb, const SweepLaunch& a, hipStream_t s) {
#define GWM_CASE(A, B) \
    if (nta == A && ntb == B) return launch_mixed_r<A, B, DIST, MODE>(a, s);
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

// Column maxima of |Z| (as FP64 bit patterns, monotone for non-negative values).
__global__ void colmax_kernel(const double* Z, int rows, int ldz, unsigned long long* colmax) {
    const int c = blockIdx.y;
    unsigned long long m = 0ull;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < rows; j += gridDim.x * blockDim.x) {
        double v = fabs(Z[(size_t)j * ldz + c]);
        unsigned long long b = (unsigned long long)__double_as_longlong(v);
        m = b > m ? b : m;
    }
    __shared__ unsigned long long s[256];
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = s[threadIdx.x] > s[threadIdx.x + w] ? s[threadIdx.x] : s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(colmax + c, s[0]);
}

// Split Z (FP64, row-major [rows][ldz]) into the FP16 hi/lo planes laid out
// [rows/32][ldz][32] and the per-column unscale factors.
__global__ void split_kernel(const double* Z, int rows, int ldz, const unsigned long long* colmax, _Float16* Zh,
                             _Float16* Zl, double* unscale) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows * ldz) return;
    const int j = (int)(e / ldz), c = (int)(e % ldz);
    const double mx = __longlong_as_double((long long)colmax[c]);
    int ex = 0;
    if (mx > 0.0) ex = (int)floor(log2(16384.0 / mx));
    ex = ex > 60 ? 60 : (ex < -60 ? -60 : ex);
    const double s = ldexp(1.0, ex);
    const double v = Z[e] * s;
    const _Float16 hi = (_Float16)(float)v;
    const _Float16 lo = (_Float16)(float)((v - (double)(float)hi) * 2048.0);
    const size_t o = ((size_t)(j / 32) * ldz + c) * 32 + (j % 32);
    Zh[o] = hi;
    Zl[o] = lo;
    if (j == 0) unscale[c] = ldexp(1.0, -ex);
}

}  // namespace


int contract_r(int nta, int ntb, int precision_mixed, int mode, int dist) {
    if (precision_mixed) return mixed_default_r(nta, ntb, mode, dist);
    return (nta + ntb) <= 8 ? 2 : 1;
}

hipError_t launch_contract_mixed(int dist, int mode, int nta, int ntb, const SweepLaunch& a, hipStream_t s) {
    if (dist == kGcTable) {
        if (!a.job0.d2tab || a.items) return hipErrorInvalidValue;
        if (mode == kModeCV) return dispatch_mixed<kGcTable, kModeCV>(nta, ntb, a, s);
        if (mode == kModePlain) return dispatch_mixed<kGcTable, kModePlain>(nta, ntb, a, s);
        if (mode == kModeHat) return dispatch_mixed<kGcTable, kModeHat>(nta, ntb, a, s);
        return dispatch_mixed<kGcTable, kModeQDiag>(nta, ntb, a, s);
    }
    if (dist == kEuclidean) {
        if (mode == kModeCV) return dispatch_mixed<kEuclidean, kModeCV>(nta, ntb, a, s);
        if (mode == kModePlain) return dispatch_mixed<kEuclidean, kModePlain>(nta, ntb, a, s);
        if (mode == kModeHat) return dispatch_mixed<kEuclidean, kModeHat>(nta, ntb, a, s);
        return dispatch_mixed<kEuclidean, kModeQDiag>(nta, ntb, a, s);
    } else {
        if (mode == kModeCV) return dispatch_mixed<kGreatCircle, kModeCV>(nta, ntb, a, s);
        if (mode == kModePlain) return dispatch_mixed<kGreatCircle, kModePlain>(nta, ntb, a, s);
        if (mode == kModeHat) return dispatch_mixed<kGreatCircle, kModeHat>(nta, ntb, a, s);
        return dispatch_mixed<kGreatCircle, kModeQDiag>(nta, ntb, a, s);
    }
}

namespace {

// Distance table (layout: launch_d2_table).  A workgroup stages the
// coordinates of kTabPts consecutive data points in LDS; each of its waves
// takes one focus group of 16 points and writes that group's kTabPts / 32
// blocks, 2 KB each, contiguous in the table.
constexpr int kTabPts = 256;
__global__ __launch_bounds__(256) void d2_table_kernel(Points P0, int n, int nfg, int ndb, float* tab) {
    __shared__ double shx[kTabPts], shy[kTabPts], shz[kTabPts];
    __shared__ float ssl[kTabPts];
    const GPoints P = global_points(P0);
    const int nsb = (ndb * 32 + kTabPts - 1) / kTabPts;   // data super-blocks
    const int sb = (int)(blockIdx.x % (unsigned)nsb);
    const int fq = (int)(blockIdx.x / (unsigned)nsb);
    {
        const int j = min(sb * kTabPts + (int)threadIdx.x, n - 1);
        shx[threadIdx.x] = P.hx[j];
        shy[threadIdx.x] = P.hy[j];
        shz[threadIdx.x] = P.hz[j];
        ssl[threadIdx.x] = (float)P.sl[j];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int fg = fq * 4 + (threadIdx.x >> 6);
    if (fg >= nfg) return;
    const int col = lane & 15, kq = lane >> 4;
    const MFocus<kGreatCircle> fc = load_mfocus<kGreatCircle>(P, min(fg * 16 + col, n - 1));
    for (int b = 0; b < kTabPts / 32; ++b) {
        const int db = sb * (kTabPts / 32) + b;
        if (db >= ndb) break;
        const int rb = b * 32 + 8 * kq;
        float d2[8];
        gc_d2_8(fc.a, fc.b, fc.c, fc.sl,
                [&](int e) { return GcRow{shx[rb + e], shy[rb + e], shz[rb + e], ssl[rb + e]}; }, d2);
        GWM_GLOBAL float* out = gptr(tab) + ((size_t)fg * ndb + db) * 512 + 4 * lane;
        *reinterpret_cast<GWM_GLOBAL v4f*>(out) = v4f{d2[0], d2[1], d2[2], d2[3]};
        *reinterpret_cast<GWM_GLOBAL v4f*>(out + 256) = v4f{d2[4], d2[5], d2[6], d2[7]};
    }
}

__global__ void var_table_kernel(const double* __restrict__ x, const double* __restrict__ y, int n, int rows_pad,
                                 double* __restrict__ Z) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;   // one entry of the rows_pad x 16 table
    if (e >= rows_pad * 16) return;
    const int j = e >> 4, c = e & 15;
    double v = 0.0;
    if (j < n && c < 2) v = c == 0 ? x[j] * x[j] : x[j] * y[j];
    Z[e] = v;
}

}  // namespace

hipError_t launch_d2_table(const Points& P, int n, int n_pad, int rows, float* tab, hipStream_t s) {
    if (n <= 0 || n_pad % 16 || rows % 64 || n_pad < n || rows < n) return hipErrorInvalidValue;
    const int nfg = n_pad / 16, ndb = rows / 32;
    const long long grid = (long long)((nfg + 3) / 4) * ((rows + kTabPts - 1) / kTabPts);
    if (grid * 256 > 0xffffffffLL) return hipErrorInvalidValue;   // work items of one launch
    hipLaunchKernelGGL(d2_table_kernel, dim3((unsigned)grid), dim3(256), 0, s, P, n, nfg, ndb, tab);
    return hipGetLastError();
}

hipError_t launch_var_table(const double* x, const double* y, int n, int rows_pad, double* Z, hipStream_t s) {
    const int total = rows_pad * 16;
    hipLaunchKernelGGL(var_table_kernel, dim3((total + 255) / 256), dim3(256), 0, s, x, y, n, rows_pad, Z);
    return hipGetLastError();
}

hipError_t launch_split_fp16(const double* Z, int rows, int ldz, unsigned long long* colmax, _Float16* Zh,
                             _Float16* Zl, double* unscale, hipStream_t s) {
    hipError_t e = hipMemsetAsync(colmax, 0, ldz * sizeof(unsigned long long), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(colmax_kernel, dim3(std::min(64, (rows + 255) / 256), ldz), dim3(256), 0, s, Z, rows, ldz,
                       colmax);
    size_t total = (size_t)rows * ldz;
    hipLaunchKernelGGL(split_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, Z, rows, ldz, colmax, Zh,
                       Zl, unscale);
    return hipGetLastError();
}

}  // namespace gwm
