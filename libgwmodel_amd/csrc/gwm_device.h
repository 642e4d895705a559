// Device-side building blocks shared by the GWR kernels (gfx950 / CDNA4).
//
// Every kernel here is compiled with -ffp-contract=off: a*b+c is never fused
// unless written as fma(), so the Euclidean distance and the kernel weights
// are evaluated with the same roundings as the reference's Armadillo
// expressions (include/gwmodelpp/spatialweight/CRSDistance.h:95-99,
// include/gwmodelpp/spatialweight/BandwidthWeight.h:68-121).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gwm {

enum KernelId { kGaussian = 0, kExponential = 1, kBisquare = 2, kTricube = 3, kBoxcar = 4 };
enum DistId { kEuclidean = 0, kGreatCircle = 1 };

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));

// Point table on the device, structure of arrays.  Euclidean uses px/py.
// Great-circle uses the raw lon/lat (degrees) for the coincident-point test
// of CRSDistance.cpp:26-37, the half unit vector (hx, hy, hz) =
// 0.5*(cos lat cos lon, cos lat sin lon, sin lat) and sl = sin lat.
struct Points {
    const double* px;   // x or lon (degrees)
    const double* py;   // y or lat (degrees)
    const double* hx;
    const double* hy;
    const double* hz;
    const double* sl;
    int n;
};

// Pointers into global memory.  Pointers read from a job table in memory are
// generic, and loads through generic pointers compile to flat loads, which
// also count against the LDS wait counter; casting them to the global address
// space once gives global loads.
#define GWM_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GWM_GLOBAL T* gptr(T* p) {
    return (GWM_GLOBAL T*)p;
}

struct GPoints {
    GWM_GLOBAL const double* px;
    GWM_GLOBAL const double* py;
    GWM_GLOBAL const double* hx;
    GWM_GLOBAL const double* hy;
    GWM_GLOBAL const double* hz;
    GWM_GLOBAL const double* sl;
    int n;
};

__device__ __forceinline__ GPoints global_points(const Points& P) {
    return GPoints{gptr(P.px), gptr(P.py), gptr(P.hx), gptr(P.hy), gptr(P.hz), gptr(P.sl), P.n};
}

// ---------------------------------------------------------------- FP64

// Squared Euclidean distance, evaluated as (dx*dx) + (dy*dy) without
// contraction; sqrt of it is bit-identical to CRSDistance::EuclideanDistance.
__device__ __forceinline__ double d2_euclid(double xf, double yf, double xj, double yj) {
    double dx = xj - xf;
    double dy = yj - yf;
    double s = dx * dx;
    return s + dy * dy;
}

// WGS-84 ellipsoid-corrected great-circle distance in km
// (CRSDistance::SpGcdist, CRSDistance.cpp:15-69), restated through exact
// identities so that one FP64 reciprocal replaces the reference's three
// divisions and two of its square roots:
//   S = sin^2(G)cos^2(L) + cos^2(F)sin^2(L) = |u_i - u_j|^2 / 4 (u: unit vectors),
//   C = 1 - S,  w = atan(sqrt(S/C)) = asin(sqrt(S)),
//   R = sqrt(S C) / w = sin(2w) / (2w)  (a series in (2w)^2 for S <= 1/2),
//   sin F cos G = (sin lat1 + sin lat2)/2,  cos F sin G = (sin lat1 - sin lat2)/2,
//   H1 / (2C) and H2 / (2S) from 1/(S C).
// R only enters through the flattening term (f ~ 3.4e-3), so its truncation
// (< 1e-13) stays below 1e-15 relative in the distance.
__device__ __forceinline__ double sinc2w_f64(double u) {
    // sum_{m=0}^{10} (-u)^m / (2m+1)!,  u = (2w)^2 <= pi^2/4
    double r = -1.9572941063391263e-20;
    r = fma(r, u, 8.2206352466243297e-18);
    r = fma(r, u, -2.8114572543455206e-15);
    r = fma(r, u, 7.6471637318198164e-13);
    r = fma(r, u, -1.6059043836821613e-10);
    r = fma(r, u, 2.5052108385441719e-08);
    r = fma(r, u, -2.7557319223985891e-06);
    r = fma(r, u, 1.9841269841269841e-04);
    r = fma(r, u, -8.3333333333333333e-03);
    r = fma(r, u, 1.6666666666666667e-01);
    return fma(-r, u, 1.0);
}

__device__ __forceinline__ double gcdist_f64(double lonf, double latf, double hxf, double hyf, double hzf,
                                             double slf, double lonj, double latj, double hxj, double hyj,
                                             double hzj, double slj) {
    const double a = 6378.137;
    const double f = 1.0 / 298.257223563;
    if (fabs(latj - latf) < 1e-8) {
        if (fabs(lonj - lonf) < 1e-8) return 0.0;
        if (fabs((fabs(lonj) + fabs(lonf)) - 360.0) < 1e-8) return 0.0;
    }
    double ex = hxj - hxf, ey = hyj - hyf, ez = hzj - hzf;
    double S = ex * ex;
    S = S + ey * ey;
    S = S + ez * ez;
    if (S <= 0.0) return 0.0;
    double C = 1.0 - S;
    double w = asin(sqrt(S));
    double R = S <= 0.5 ? sinc2w_f64(4.0 * w * w) : sqrt(S * C) / w;
    double inv = 1.0 / (S * C);
    double H1 = (3 * R - 1) * (0.5 * S * inv);
    double H2 = (3 * R + 1) * (0.5 * C * inv);
    double sp = 0.5 * (slj + slf);
    double sm = 0.5 * (slj - slf);
    double D = 2 * w * a;
    return D * (1 + f * H1 * (sp * sp) - f * H2 * (sm * sm));
}

// Squared great-circle distance (km^2) of a near pair, 1e-19 <= S <= 1/8
// (central angle up to about 41 degrees), from the same identities as
// gcdist_f64 without the arcsine and the square root: on [0, 1/8]
//   w^2 / S = asin(sqrt(S))^2 / S     (minimax, degree 10, 3e-18 relative),
//   R = sinc(2w)                      (minimax in S, degree 7, 3e-14 absolute;
//                                      R only enters the flattening term,
//                                      f (3R +- 1) / 2 < 1e-2),
// fitted with mpmath.chebyfit; d^2 = 4 a^2 w^2 (1 + f H1 sp^2 - f H2 sm^2)^2
// with one FP64 reciprocal.  Pairs outside the range (coincident and
// near-coincident points, the reference's lon / lat equality tests, and far
// pairs) take gcdist_f64.
constexpr double kNearSMin = 1e-19, kNearSMax = 0.125;

__device__ __forceinline__ double gc_d2_near_f64(double S, double slf, double slj) {
    const double a = 6378.137;
    const double f = 1.0 / 298.257223563;
    double q = 0.04621192004569709;
    q = fma(q, S, 0.021186963403752236);
    q = fma(q, S, 0.03455657591647529);
    q = fma(q, S, 0.039645299221225);
    q = fma(q, S, 0.04872256259518243);
    q = fma(q, S, 0.06156765311448536);
    q = fma(q, S, 0.08126985231940705);
    q = fma(q, S, 0.11428571411189105);
    q = fma(q, S, 0.17777777777917667);
    q = fma(q, S, 0.33333333333332893);
    q = fma(q, S, 1.0);
    const double w2 = S * q;
    double R = -0.016035700859708638;
    R = fma(R, S, -0.012188560686387616);
    R = fma(R, S, -0.01812589061232504);
    R = fma(R, S, -0.025953150420487386);
    R = fma(R, S, -0.04232830192546632);
    R = fma(R, S, -0.0888888847995917);
    R = fma(R, S, -0.6666666666911097);
    R = fma(R, S, 1.000000000000024);
    const double C = 1.0 - S;
    const double inv = 1.0 / (S * C);
    const double H1 = (3 * R - 1) * (0.5 * S * inv);
    const double H2 = (3 * R + 1) * (0.5 * C * inv);
    const double sp = 0.5 * (slj + slf);
    const double sm = 0.5 * (slj - slf);
    const double g = 1 + f * H1 * (sp * sp) - f * H2 * (sm * sm);
    return (4.0 * a * a) * w2 * (g * g);
}

// Kernel weights with the CPU boundary (d < bw): BandwidthWeight.h:68-121.
//
// Gaussian and exponential weights whose exponent lies more than kExpCut
// below the exponent of the focus point's largest weight are taken as zero
// (e^-40 = 4.2e-18 of that weight, below 2^-57: less than half an ulp of any
// term the largest weight forms).  The largest weight is the focus point's
// own (1, at d = 0) in fits and AICc sweeps; in CV sweeps, whose own weight
// is zeroed, and at prediction locations it is the weight of the nearest
// other point (distance dnn), so the cut sits at exponent e(dnn) - 40.  The
// reference keeps every weight until exp underflows (e^-745); dropping them
// moves every entry of X^T W X and X^T W y by at most 4.2e-18 times the
// largest weight times the unweighted sum of its |z_j|, and lets the sweeps
// skip points beyond 8.9 bandwidths (sqrt(dnn^2 + 80 b^2) in general) instead
// of 38.6 (the exponent of e^-745).
constexpr double kExpCut = 40.0;

__device__ __forceinline__ double exp_cut(double e) { return e < -kExpCut ? 0.0 : exp(e); }

// Exponent of the kernel weight at distance d (Gaussian and exponential), the
// reference's expression (BandwidthWeight.h:68-79).
__device__ __forceinline__ double kernel_exponent(int kernel, double d, double bw) {
    return kernel == kGaussian ? (d * d) / ((-2.0) * (bw * bw)) : -d / bw;
}

// Mixed precision: a focus point whose largest weight is below this lies so
// far from every other point that its FP16-split weights lose precision
// (FP16 hi parts turn subnormal below 2^-14); such points are recomputed in
// FP64.
constexpr double kMixedMinWeight = 1.0 / 4096.0;

__device__ __forceinline__ double kernel_f64(int kernel, double d, double bw) {
    switch (kernel) {
    case kGaussian: return exp_cut((d * d) / ((-2.0) * (bw * bw)));
    case kExponential: return exp_cut(-d / bw);
    case kBisquare: {
        double t = 1.0 - (d * d) / (bw * bw);
        return d < bw ? t * t : 0.0;
    }
    case kTricube: {
        double t = 1.0 - (d * d * d) / (bw * bw * bw);
        return d < bw ? (t * t) * t : 0.0;
    }
    default: return d < bw ? 1.0 : 0.0;
    }
}

// Smallest double x whose correctly rounded square root is >= bw (0 when bw
// is not positive or NaN).  RN(sqrt(.)) is non-decreasing over the doubles,
// so for every double d2 >= 0: d2 < x exactly when RN(sqrt(d2)) < bw -- the
// reference's boundary test d < bw on the distance d = sqrt(d2), decided on
// the squared distance without the square root.  bw^2 lies within a few ulps
// of x when it is a normal number; a bandwidth below 2^-500 is scaled by
// 2^1000 first (squares by 2^2000: exact for every x up to the boundary, and
// RN(sqrt(x s^2)) = s RN(sqrt(x)) since both roots are normal numbers), and
// the scaled boundary is mapped back rounding up.
// the double `steps` representable values above a non-negative x (below for steps < 0)
__host__ __device__ inline double ulp_step(double x, long long steps) {
    return __builtin_bit_cast(double, __builtin_bit_cast(long long, x) + steps);
}

__host__ __device__ inline double sqrt_boundary_scaled(double bw) {
    double x = bw * bw;   // bw >= 2^-500: normal, or inf for bw > 2^512
    if (sqrt(x) >= bw) {
        for (int i = 0; i < 64; ++i) {
            const double p = ulp_step(x, -1);   // x > 0 here
            if (!(sqrt(p) >= bw)) break;
            x = p;
        }
    } else {
        for (int i = 0; i < 64 && !(sqrt(x) >= bw); ++i) x = ulp_step(x, 1);
    }
    return x;
}

__host__ __device__ inline double sqrt_boundary(double bw) {
    if (!(bw > 0.0)) return 0.0;
    if (bw >= 0x1p-500) return sqrt_boundary_scaled(bw);
    const double t = sqrt_boundary_scaled(bw * 0x1p1000);   // t < 2^1000
    double q = (t * 0x1p-1000) * 0x1p-1000;                // rounds only in the last step
    if ((q * 0x1p1000) * 0x1p1000 < t) q = ulp_step(q, 1);
    return q;
}

// Per-bandwidth constants so that the sweep weighs pairs without divisions
// (multiplication by the reciprocal: within an ulp of the reference's quotient;
// the boundary tests d < bw are unchanged).
struct KConst64 {
    double bw, gauss, inv_bw, inv_bw2, inv_bw3;   // gauss = 1 / (-2 bw^2)
    double cut;    // exponents below cut give zero weights
    double d2_lt;  // sqrt_boundary(bw): d < bw exactly when d^2 < d2_lt
};

__device__ __forceinline__ KConst64 make_kconst64(double bw) {
    KConst64 k;
    k.bw = bw;
    k.d2_lt = sqrt_boundary(bw);
    k.gauss = 1.0 / ((-2.0) * (bw * bw));
    k.inv_bw = 1.0 / bw;
    k.inv_bw2 = 1.0 / (bw * bw);
    k.inv_bw3 = 1.0 / (bw * bw * bw);
    k.cut = -kExpCut;
    return k;
}

// The constants of a focus point whose largest weight is that of a point at
// distance dnn (CV sweeps, prediction locations): the exponent cut moves down
// by the exponent of that weight.
__device__ __forceinline__ KConst64 make_kconst64_nn(double bw, int kernel, double dnn) {
    KConst64 k = make_kconst64(bw);
    if (dnn > 0.0 && (kernel == kGaussian || kernel == kExponential))
        k.cut = kernel_exponent(kernel, dnn, bw) - kExpCut;
    return k;
}

// Largest kernel weight of a focus point whose nearest point lies at dnn,
// without any cut.
__device__ __forceinline__ double nearest_weight(int kernel, double dnn, double bw) {
    switch (kernel) {
    case kGaussian:
    case kExponential: return exp(kernel_exponent(kernel, dnn, bw));
    case kBisquare: {
        double t = 1.0 - (dnn * dnn) / (bw * bw);
        return dnn < bw ? t * t : 0.0;
    }
    case kTricube: {
        double t = 1.0 - (dnn * dnn * dnn) / (bw * bw * bw);
        return dnn < bw ? (t * t) * t : 0.0;
    }
    default: return dnn < bw ? 1.0 : 0.0;
    }
}

__device__ __forceinline__ double kernel_fast_f64(int kernel, double d, const KConst64& k) {
    switch (kernel) {
    case kGaussian: {
        const double e = (d * d) * k.gauss;
        return e < k.cut ? 0.0 : exp(e);
    }
    case kExponential: {
        const double e = -d * k.inv_bw;
        return e < k.cut ? 0.0 : exp(e);
    }
    case kBisquare: {
        double t = 1.0 - (d * d) * k.inv_bw2;
        return d < k.bw ? t * t : 0.0;
    }
    case kTricube: {
        double t = 1.0 - (d * d * d) * k.inv_bw3;
        return d < k.bw ? (t * t) * t : 0.0;
    }
    default: return d < k.bw ? 1.0 : 0.0;
    }
}

// Distance from point f of the focus table F to point j of the data table P
// (F is P itself for fits and criteria, the prediction locations for predict).
template <int DIST, class PTS>
__device__ __forceinline__ double distance_f64(const PTS& F, int f, const PTS& P, int j) {
    if (DIST == kEuclidean) {
        return sqrt(d2_euclid(F.px[f], F.py[f], P.px[j], P.py[j]));
    } else {
        return gcdist_f64(F.px[f], F.py[f], F.hx[f], F.hy[f], F.hz[f], F.sl[f], P.px[j], P.py[j], P.hx[j],
                          P.hy[j], P.hz[j], P.sl[j]);
    }
}
template <int DIST, class PTS>
__device__ __forceinline__ double distance_f64(const PTS& P, int f, int j) {
    return distance_f64<DIST>(P, f, P, j);
}

// Selection key: a monotone image of the distance.  Euclidean uses the
// squared distance (sqrt is monotone, so order statistics commute with it);
// great-circle uses the distance itself.
template <int DIST, class PTS>
__device__ __forceinline__ double select_value(const PTS& F, int f, const PTS& P, int j) {
    if (DIST == kEuclidean) return d2_euclid(F.px[f], F.py[f], P.px[j], P.py[j]);
    return distance_f64<DIST>(F, f, P, j);
}
template <int DIST>
__device__ __forceinline__ double select_to_distance(double v) {
    return DIST == kEuclidean ? sqrt(v) : v;
}

// ---------------------------------------------------------------- FP32

// FP32 pieces of the mixed-precision producer: the same restatement as
// gcdist_f64 with hardware square root / reciprocal and an FP32 arcsine
// (odd polynomial on [0, 0.5], reflected as pi/2 - 2 asin(sqrt((1-x)/2))).
__device__ __forceinline__ float asin_f32(float x) {
    // x in [0, 1]
    bool big = x > 0.5f;
    float z = big ? 0.5f * (1.0f - x) : x * x;
    float s = big ? __builtin_sqrtf(z) : x;
    // asin(s) = s + s*z*p(z) on |s| <= 0.5 with the published Cephes asinf
    // coefficients (S. L. Moshier, Cephes Math Library), ~2.5e-7 relative
    float p = fmaf(fmaf(fmaf(fmaf(4.2163199048e-2f, z, 2.4181311049e-2f), z, 4.5470025998e-2f), z,
                        7.4953002686e-2f), z, 1.6666752422e-1f);
    float r = fmaf(s * z, p, s);
    return big ? 1.5707963267948966f - 2.0f * r : r;
}

// S = sin^2(sigma/2) is formed in FP64 by the caller (the chord difference
// of nearby points cancels); the transcendental tail runs in FP32.
__device__ __forceinline__ float gcdist_tail_f32(float S, float slf, float slj) {
    const float a = 6378.137f;
    const float f = (float)(1.0 / 298.257223563);
    if (!(S > 0.0f)) return 0.0f;
    float C = 1.0f - S;
    float s = __builtin_sqrtf(S);
    float w = asin_f32(s);
    float R = s * __builtin_sqrtf(C) * __builtin_amdgcn_rcpf(w);
    float H1 = (3.0f * R - 1.0f) * 0.5f * __builtin_amdgcn_rcpf(C);
    float H2 = (3.0f * R + 1.0f) * 0.5f * __builtin_amdgcn_rcpf(S);
    float sp = 0.5f * (slj + slf);
    float sm = 0.5f * (slj - slf);
    float corr = fmaf(f * H1, sp * sp, -(f * H2) * (sm * sm));
    return 2.0f * a * w * (1.0f + corr);
}

__device__ __forceinline__ float kernel_f32(int kernel, float d, float bw, float inv_bw) {
    switch (kernel) {
    case kGaussian: {
        float t = d * inv_bw;
        return __expf(-0.5f * t * t);
    }
    case kExponential: return __expf(-d * inv_bw);
    case kBisquare: {
        float q = d * inv_bw;
        float t = 1.0f - q * q;
        return d < bw ? t * t : 0.0f;
    }
    case kTricube: {
        float q = d * inv_bw;
        float t = 1.0f - q * q * q;
        return d < bw ? (t * t) * t : 0.0f;
    }
    default: return d < bw ? 1.0f : 0.0f;
    }
}

// ---------------------------------------------------------------- pruning
//
// The points are stored in a spatial order and cut into chunks of 64 with a
// bounding box each.  A chunk whose points are all farther from every focus
// point of a workgroup than the distance at which the kernel weight is exactly
// zero contributes exactly nothing, and the sweep skips it; results are the
// same bits as without skipping.

// Distance beyond which the kernel weight is exactly zero: in FP64 (the
// exponent cut at -kExpCut of kernel_f64; compact kernels vanish for
// d >= bw), or in the mixed path, whose FP32 weight is split into FP16 hi + lo
// parts that are both zero below 2^-36 (exp2 arguments below -37 keep a
// margin).  The FP64 distances carry a relative margin of 1e-9 over the cut
// for the rounding of the box gaps and of the exponent.  Never zero: infinity.
// dnn: the largest distance to a focus point's nearest other point among the
// focus points considered (FP64 CV sweeps and predictions; 0 elsewhere), which
// moves the FP64 exponent cut outwards (make_kconst64_nn).
__device__ __forceinline__ double zero_weight_distance(int kernel, double bw, bool mixed, double dnn = 0.0) {
    switch (kernel) {
    case kGaussian:   // d^2 / (2 b^2) > 26.3 (mixed) / kExpCut + dnn^2 / (2 b^2) (FP64)
        return mixed ? 7.25 * bw : sqrt(dnn * dnn + 2.0 * kExpCut * (bw * bw)) * (1.0 + 1e-9);
    case kExponential: return mixed ? 26.0 * bw : (dnn + kExpCut * bw) * (1.0 + 1e-9);
    default: return bw * (mixed ? 1.0 + 1e-5 : 1.0 + 1e-9);
    }
}

// Threshold on the squared box gap (box_gap2) beyond which every pair is
// farther than `cut`.  Euclidean: cut^2.  Great circle: the gap is measured
// between boxes of the half unit vectors h, whose squared difference is
// S = sin^2(theta/2) (theta the central angle), and the ellipsoidal distance
// is at least 0.98 a theta (the flattening terms change it by less than 4f).
template <int DIST>
__device__ __forceinline__ double prune_threshold(double cut) {
    if (DIST == kEuclidean) return cut * cut;
    const double a = 6378.137;
    const double theta = cut / (0.98 * a);
    if (!(theta < 3.14159)) return INFINITY;
    const double s = sin(0.5 * theta);
    return s * s;
}

// Squared gap between a focus box (lo[d], hi[d]) and a chunk box stored as
// [lo_0, hi_0, lo_1, hi_1, ...].
template <int DIST>
__device__ __forceinline__ double box_gap2(const double* lo, const double* hi, GWM_GLOBAL const double* box) {
    constexpr int D = DIST == kEuclidean ? 2 : 3;
    double g2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double g = fmax(0.0, fmax(box[2 * d] - hi[d], lo[d] - box[2 * d + 1]));
        g2 += g * g;
    }
    return g2;
}

// Workgroup-wide box of the focus points, their largest bandwidth and their
// largest nearest-point distance: every thread of the 256-thread workgroup
// passes the box lo/hi of its own focus points, their largest bandwidth b and
// distance dn, and receives the workgroup's.
template <int D>
__device__ __forceinline__ void workgroup_box(double* lo, double* hi, double b, double* bmax, double dn = 0.0,
                                              double* dnmax = nullptr) {
    __shared__ double s_part[4][2 * D + 2];
    double v[2 * D + 2];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        v[2 * d] = lo[d];
        v[2 * d + 1] = hi[d];
    }
    v[2 * D] = b;
    v[2 * D + 1] = dn;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            v[2 * d] = fmin(v[2 * d], __shfl_xor(v[2 * d], m));
            v[2 * d + 1] = fmax(v[2 * d + 1], __shfl_xor(v[2 * d + 1], m));
        }
        v[2 * D] = fmax(v[2 * D], __shfl_xor(v[2 * D], m));
        v[2 * D + 1] = fmax(v[2 * D + 1], __shfl_xor(v[2 * D + 1], m));
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int e = 0; e < 2 * D + 2; ++e) s_part[wave][e] = v[e];
    __syncthreads();
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = fmin(fmin(s_part[0][2 * d], s_part[1][2 * d]), fmin(s_part[2][2 * d], s_part[3][2 * d]));
        hi[d] = fmax(fmax(s_part[0][2 * d + 1], s_part[1][2 * d + 1]), fmax(s_part[2][2 * d + 1], s_part[3][2 * d + 1]));
    }
    *bmax = fmax(fmax(s_part[0][2 * D], s_part[1][2 * D]), fmax(s_part[2][2 * D], s_part[3][2 * D]));
    if (dnmax)
        *dnmax = fmax(fmax(s_part[0][2 * D + 1], s_part[1][2 * D + 1]), fmax(s_part[2][2 * D + 1], s_part[3][2 * D + 1]));
}

// Job of batch entry t, given the prefix sums start[0..njobs] of the jobs'
// entry counts (start[0] = 0): the largest j with start[j] <= t.
__device__ __forceinline__ int find_job(const int* start, int njobs, int t) {
    int lo = 0, hi = njobs;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (start[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

}  // namespace gwm
