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
    _Float16 zh[2][NTA * 16][kColStride];
    _Float16 zl[2][NTA * 16][kColStride];
    MCoords<DIST> c;
    BwConst bwc[MODE == kModeQDiag ? kChunkM : 1];   // qDiag: constants of each data point's bandwidth
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

__device__ __forceinline__ float weight_from_d2(int kernel, float d2, const BwConst& c) {
    switch (kernel) {
    case kGaussian: return __builtin_amdgcn_exp2f(d2 * c.gauss);
    case kExponential: return __builtin_amdgcn_exp2f(-1.4426950408889634f * __builtin_sqrtf(d2) * c.inv_bw);
    case kBisquare: {
        float t = 1.0f - d2 * c.inv_bw2;
        return d2 < c.bw2 ? t * t : 0.0f;
    }
    case kTricube: {
        float q = __builtin_sqrtf(d2) * c.inv_bw;
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
            w[e] = __builtin_amdgcn_exp2f(fmaf(-1.4426950408889634f * __builtin_sqrtf(d2[e]), c.inv_bw, e0));
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
            float q = __builtin_sqrtf(d2[e]) * c.inv_bw;
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

template <int NTA, int NTB, int DIST, int MODE, int R>
__global__ __launch_bounds__(256) void contract_mixed_kernel(SweepLaunch L) {
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

    MFocus<DIST> fc[R];
    BwConst fbw[R];
    int fidx[R];
    double fb = 0.0;   // largest focus bandwidth of this lane
#pragma unroll
    for (int r = 0; r < R; ++r) {
        int f = fbase + 16 * r + col;
        int fl = f < nf ? f : nf - 1;
        if (focus_idx) fl = focus_idx[fl];
        fidx[r] = qtab ? -1 : (focus_idx ? (f < nf ? fl : -1) : f);   // self index for the CV exclusion
        fc[r] = load_mfocus<DIST>(FP, fl);
        const double b = (MODE != kModeQDiag && bw_pt) ? bw_pt[fl] : args.bw;
        fbw[r] = make_bw((float)b);
        fb = fmax(fb, b);
    }
    // pruning: the workgroup's focus box and the zero-weight threshold
    constexpr int BD = DIST == kEuclidean ? 2 : 3;
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
        if (MODE != kModeQDiag) thr = prune_threshold<DIST>(zero_weight_distance(args.kernel, fb, true));
        else if (!args.chunk_bw) thr = prune_threshold<DIST>(zero_weight_distance(args.kernel, args.bw, true));
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
    for (int j0 = jbeg; j0 < jend; j0 += kChunkM) {
        if (prune) {
            const int ch = j0 / kChunkM;
            double t = thr;
            if (MODE == kModeQDiag && args.chunk_bw)
                t = prune_threshold<DIST>(zero_weight_distance(args.kernel, gptr(args.chunk_bw)[ch], true));
            const int skip = box_gap2<DIST>(flo, fhi, gptr(args.boxes) + (size_t)ch * 2 * BD) > t;
            if (__builtin_amdgcn_readfirstlane(skip)) continue;
        }
        ++swept;
        __syncthreads();
        // stage the two 32-row blocks of both planes: global [blk][col][32]
        {
            const size_t base = (size_t)(j0 / 32) * ldz * 32;
            GWM_GLOBAL const v4u* gh = reinterpret_cast<GWM_GLOBAL const v4u*>(gptr(args.Zh) + base);
            GWM_GLOBAL const v4u* gl = reinterpret_cast<GWM_GLOBAL const v4u*>(gptr(args.Zl) + base);
            const int vecs = 2 * ldz * 4;   // 16-byte vectors per plane (8 halfs each)
            for (int e = threadIdx.x; e < vecs; e += 256) {
                const int blk = e / (ldz * 4);
                const int rem = e % (ldz * 4);
                const int c = rem >> 2, part = rem & 3;
                *reinterpret_cast<v4u*>(&sh.zh[blk][c][part * 8]) = gh[e];
                *reinterpret_cast<v4u*>(&sh.zl[blk][c][part * 8]) = gl[e];
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
            if constexpr (MODE == kModeQDiag) sh.bwc[threadIdx.x] = make_bw((float)(bw_pt ? bw_pt[jl] : args.bw));
        }
        __syncthreads();
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
                    bh[t] = *reinterpret_cast<const v8h*>(&sh.zh[s][16 * t + col][8 * kq]);
                    bl[t] = *reinterpret_cast<const v8h*>(&sh.zl[s][16 * t + col][8 * kq]);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v8h ah, al, a2h, a2l;
                const int rb = 32 * s + 8 * kq;
                float d2[8], w[8];
                float smax = 0.0f;   // great circle: any pair beyond 60 degrees of arc (S > 0.25)?
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int rr = rb + e;
                    if constexpr (DIST == kEuclidean) {
                        double dx = sh.c.x[rr] - fc[r].a, dy = sh.c.y[rr] - fc[r].b;
                        d2[e] = (float)(dx * dx + dy * dy);
                    } else {
                        double ex = sh.c.hx[rr] - fc[r].a, ey = sh.c.hy[rr] - fc[r].b, ez = sh.c.hz[rr] - fc[r].c;
                        const float slj = sh.c.sl[rr];
                        const float S = (float)(ex * ex + ey * ey + ez * ez);
                        d2[e] = gc_d2_near_f32(S, fc[r].sl + slj, slj - fc[r].sl);
                        smax = fmaxf(smax, S);
                    }
                }
                if constexpr (DIST != kEuclidean) {
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
                }
                // qDiag with a fixed bandwidth: fbw holds the same constants as sh.bwc
                if (MODE == kModeQDiag && bw_pt) {
#pragma unroll
                    for (int e = 0; e < 8; ++e) w[e] = weight_from_d2(args.kernel, d2[e], sh.bwc[rb + e]);
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
            GWM_GLOBAL double* o = out + (size_t)((isB ? ldz : 0) + zc) * np + f0;
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
        if (a.shape_r == 1 && !a.items) return launch_mixed_rr<NTA, NTB, DIST, MODE, 1>(a, s);
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
