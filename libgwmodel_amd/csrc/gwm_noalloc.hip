// Allocation-free variant of the criterion entry point (gwm_hip_criterion_noalloc).
//
// Every device buffer beyond the caller's inputs -- the product table Z, the
// great-circle point tables, adaptive thresholds, nearest-point distances,
// the partial sums of the sweep, the per-focus results and the reduction --
// lives in one file-scope device array, `scratch`, of kScratchBytes per
// device.  Its device address is looked up once per device at context
// creation (noalloc_init, called by gwm_hip_create with the device
// current) and only that address is ever handed to kernels.  The entry point
// calls no hipMalloc / hipFree: inputs are the caller's device arrays in the
// caller's order (no spatial sort, so no pruning), and the sweep uses at most
// four data segments to keep the partial sums small.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "gwm_ctx.h"

using namespace gwm;
using namespace gwm::detail;

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;                        // guards g_scratch
double* g_scratch[kMaxDevices] = {};    // device address of `scratch` on each device
std::mutex g_dev_mu[kMaxDevices];       // one call per device at a time

// Z = [x_a x_b (a <= b) | x_a y] in rows of ldz doubles (FP64 sweep layout),
// rows n .. rows_pad-1 zero.
__global__ void ztable_kernel(const double* X, const double* y, int n, int k, int ldz, int rows_pad, double* Z) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= rows_pad) return;
    double* z = Z + (size_t)j * ldz;
    int p = 0;
    if (j < n) {
        for (int a = 0; a < k; ++a) {
            const double xa = X[j + (size_t)a * n];
            for (int b = a; b < k; ++b) z[p++] = xa * X[j + (size_t)b * n];
        }
        const double yj = y[j];
        for (int a = 0; a < k; ++a) z[p++] = X[j + (size_t)a * n] * yj;
    }
    for (; p < ldz; ++p) z[p] = 0.0;
}

// Great circle: half unit vectors and sin(lat) of points given in degrees.
__global__ void sphere_kernel(const double* lon, const double* lat, int n, double* hx, double* hy, double* hz,
                              double* sl) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double de2ra = 3.14159265358979323846 / 180.0;
    const double lo = lon[j] * de2ra, la = lat[j] * de2ra;
    hx[j] = 0.5 * cos(la) * cos(lo);
    hy[j] = 0.5 * cos(la) * sin(lo);
    hz[j] = 0.5 * sin(la);
    sl[j] = sin(la);
}

__global__ void zero_kernel(double* p, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = 0.0;
}

// Carves consecutive 256-byte-aligned pieces out of the scratch array.
struct Carve {
    char* base;
    size_t used = 0;
    template <class T>
    T* take(size_t count) {
        T* p = reinterpret_cast<T*>(base + used);
        used += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

}  // namespace

namespace gwm {
namespace detail {

int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return GWM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_scratch[device]) return GWM_OK;
    double* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return GWM_ERR_HIP;
    g_scratch[device] = ptr;
    return GWM_OK;
}

}  // namespace detail
}  // namespace gwm

extern "C" {

int gwm_hip_criterion_noalloc(gwm_hip_ctx* c, int n, int k, const double* coords, const double* X, const double* y,
                              int distance, int kernel, int adaptive, double bw, int crit, double* value) {
    if (!c || !value) return GWM_ERR_INVALID;
    *value = DBL_MAX;
    if (n < 2 || k < 1 || k > 16 || !coords || !X || !y || (distance != GWM_EUCLIDEAN && distance != GWM_GREAT_CIRCLE) ||
        kernel < 0 || kernel > 4 || (crit != GWM_CRIT_AIC && crit != GWM_CRIT_CV)) {
        c->err = "noalloc: invalid arguments (need n >= 2, 1 <= k <= 16, device arrays, known distance/kernel/criterion)";
        return GWM_ERR_INVALID;
    }
    if (adaptive ? !(bw >= 1.0) : !(bw > 0.0)) {
        c->err = "noalloc: bandwidth must be >= 1 neighbour (adaptive) or > 0 (fixed)";
        return GWM_ERR_INVALID;
    }
    if (c->device < 0 || c->device >= kMaxDevices) return GWM_ERR_INVALID;
    double* base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = g_scratch[c->device];
    }
    if (!base) {
        c->err = "noalloc: the scratch array of this device was not set up (gwm_hip_create failed to look it up)";
        return GWM_ERR_HIP;
    }
    // layout: at most 4 data segments of a multiple of 64 points
    const int nta = nta_of(k), ldz = nta * 16, PT = ldz;
    const int rows_pad = (n + 63) / 64 * 64, n_pad = (n + 127) / 128 * 128;
    const int seg_len = std::max(64, ((n + 3) / 4 + 63) / 64 * 64);
    const int nseg = (n + seg_len - 1) / seg_len;
    Carve cv{reinterpret_cast<char*>(base)};
    double* Z = cv.take<double>((size_t)rows_pad * ldz);
    double* part = cv.take<double>((size_t)nseg * PT * n_pad);
    double* res2 = cv.take<double>(n);
    double* sii = cv.take<double>(n);
    double* ssi = cv.take<double>(n);
    double* bw_pt = adaptive ? cv.take<double>(n) : nullptr;
    double* dnn = crit == GWM_CRIT_CV ? cv.take<double>(n) : nullptr;
    double* sph = distance == GWM_GREAT_CIRCLE ? cv.take<double>((size_t)4 * n) : nullptr;
    double* red = cv.take<double>(8);            // 4 sums + fail flag slot
    int* fail = reinterpret_cast<int*>(red + 4);
    unsigned int* nleft = reinterpret_cast<unsigned int*>(red + 5);
    if (cv.used > kScratchBytes) {
        c->err = "noalloc: n = " + std::to_string(n) + ", k = " + std::to_string(k) + " needs " +
                 std::to_string((cv.used + (1 << 20) - 1) >> 20) + " MiB of scratch, more than the " +
                 std::to_string(kScratchBytes >> 20) + " MiB each device holds";
        return GWM_ERR_INVALID;
    }
    // one call per device at a time: every call shares the device's array
    std::lock_guard<std::mutex> dlk(g_dev_mu[c->device]);
    if (hipSetDevice(c->device) != hipSuccess) return GWM_ERR_HIP;
    hipStream_t s = c->stream;
    auto fail_hip = [&](hipError_t e) {
        c->err = std::string("noalloc: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? GWM_ERR_OOM : GWM_ERR_HIP;
    };
    hipError_t e;
    Points P{};
    P.px = coords;
    P.py = coords + n;
    P.n = n;
    if (sph) {
        P.hx = sph;
        P.hy = sph + n;
        P.hz = sph + 2 * (size_t)n;
        P.sl = sph + 3 * (size_t)n;
        hipLaunchKernelGGL(sphere_kernel, dim3((n + 255) / 256), dim3(256), 0, s, coords, coords + n, n, sph, sph + n,
                           sph + 2 * (size_t)n, sph + 3 * (size_t)n);
    }
    hipLaunchKernelGGL(ztable_kernel, dim3((rows_pad + 255) / 256), dim3(256), 0, s, X, y, n, k, ldz, rows_pad, Z);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, s, red, 8);   // sums, fail flag, counter
    if ((e = hipGetLastError()) != hipSuccess) return fail_hip(e);
    if (adaptive && (e = launch_adaptive_bw(distance, P, bw, bw_pt, nullptr, nullptr, nleft, s)) != hipSuccess)
        return fail_hip(e);
    if (dnn && (e = launch_adaptive_bw(distance, P, 2.0, dnn, nullptr, nullptr, nleft, s)) != hipSuccess)
        return fail_hip(e);
    const int mode = crit == GWM_CRIT_CV ? kModeCV : kModePlain;
    SweepLaunch L;
    SweepJob& j = L.job0;
    j.P = P;
    j.Z = Z;
    j.Zh = nullptr;
    j.Zl = nullptr;
    j.colscale = nullptr;
    j.bw_pt = bw_pt;
    j.bw = bw;
    j.kernel = kernel;
    j.n = n;
    j.n_pad = n_pad;
    j.seg_len = seg_len;
    j.nseg = nseg;
    j.nta = nta;
    j.ntb = 0;
    j.out = part;
    j.dnn = dnn;
    L.jobs = nullptr;
    L.items = nullptr;
    L.nitems = 0;
    if ((e = launch_contract_f64(distance, mode, nta, 0, L, s)) != hipSuccess) return fail_hip(e);
    SolveArgs a;
    a.part = part;
    a.nseg = nseg;
    a.PT = PT;
    a.colB = 0;
    a.n = n;
    a.n_pad = n_pad;
    a.X = X;
    a.y = y;
    a.bw_pt = bw_pt;
    a.bw = bw;
    a.kernel = kernel;
    a.res2 = res2;
    a.sii = sii;
    a.ssi = ssi;
    a.betas = nullptr;
    a.se_raw = nullptr;
    a.G = nullptr;
    a.ldg = 0;
    a.gvec = nullptr;
    a.fail = fail;
    const int smode = crit == GWM_CRIT_CV ? kSolveCV : kSolveAIC;
    if ((e = launch_solve(k, smode, a, s)) != hipSuccess) return fail_hip(e);
    if ((e = launch_reduce(res2, crit == GWM_CRIT_CV ? nullptr : sii, nullptr, nullptr, n, red, s)) != hipSuccess)
        return fail_hip(e);
    double h[5];
    if ((e = hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail_hip(e);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail_hip(e);
    int failed = 0;
    std::memcpy(&failed, &h[4], sizeof(int));
    if (failed) return GWM_ERR_NOT_SPD;
    const double v = crit == GWM_CRIT_CV ? h[0] : aicc(h[0], (double)n, h[1]);
    if (!std::isfinite(v)) return GWM_ERR_NOT_SPD;
    *value = v;
    return GWM_OK;
}

}  // extern "C"
