// Host-side helpers shared by the translation units of the single-problem
// API: gwm_problem.cpp (problem set-up, device buffers, adaptive bandwidths,
// the distance table), gwm_sweep.cpp (the sweep launcher, its autotuner and
// the FP64 recomputation) and gwm_api.cpp (criteria, searches, fits,
// multiscale, prediction, variable selection and the C entry points).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "gwm_ctx.h"

#define GWM_HIP_CHECK(ctx, expr)                                                         \
    do {                                                                                 \
        hipError_t e__ = (expr);                                                         \
        if (e__ != hipSuccess) {                                                         \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e__);             \
            return e__ == hipErrorOutOfMemory ? GWM_ERR_OOM : GWM_ERR_HIP;               \
        }                                                                                \
    } while (0)

namespace gwm {
namespace detail {

constexpr double kPi = 3.14159265358979323846;

// Context buffers come from dalloc and go back through cfree: stream-ordered
// (hipMallocAsync / hipFreeAsync) on server contexts, plain otherwise.
template <class T>
void cfree(gwm_hip_ctx* c, T*& p) {
    if (p) {
        if (c->async_mem) hipFreeAsync(p, c->stream);
        else hipFree(p);
    }
    p = nullptr;
}

template <class T>
int dalloc(gwm_hip_ctx* c, T** p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (c->async_mem) GWM_HIP_CHECK(c, hipMallocAsync((void**)p, bytes, c->stream));
    else GWM_HIP_CHECK(c, hipMalloc((void**)p, bytes));
    return GWM_OK;
}

void free_problem(gwm_hip_ctx* c);
// Download a sorted-order device array (n x cols) into a caller-order buffer.
int download_caller(gwm_hip_ctx* c, const double* dev, double* host, int cols);

template <class F>
void parallel_for(size_t n, const F& f) {
    const size_t hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = std::min<size_t>(std::min<size_t>(hw, 16), (n + 16383) / 16384);
    if (T <= 1) {
        for (size_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (size_t i = n * t / T; i < n * (t + 1) / T; ++i) f(i);
        });
    for (auto& x : th) x.join();
}

// Great circle: the half unit vectors (hx, hy, hz) and sin(lat) of points
// given in degrees.
void sphere_tables(size_t N, const double* lon, const double* lat, double* hx, double* hy, double* hz, double* sl);
// Permutation of the n points along a Hilbert curve over (xs, ys).
std::vector<int> spatial_order(int n, const double* xs, const double* ys);
// Bounding boxes of the 64-point chunks of sorted points ([lo, hi] per dimension).
std::vector<double> chunk_boxes(int n, int dims, const double* const* xs);

// The caller's stop callback (ITelegram::stop) asks to cancel.
bool stopped(gwm_hip_ctx* c);
int ensure_part(gwm_hip_ctx* c, size_t bytes);
int ensure_plan(gwm_hip_ctx* c, size_t items);
void timer_begin(gwm_hip_ctx* c);
int timer_end(gwm_hip_ctx* c, const char* name);
int ensure_d2_table(gwm_hip_ctx* c);

// Scope of an outermost call whose mixed great-circle sweeps share one
// distance table; the table is released when that call returns.
struct D2Scope {
    gwm_hip_ctx* c;
    explicit D2Scope(gwm_hip_ctx* ctx) : c(ctx) {
        if (c->d2_scope++ == 0) c->d2_valid = false;
    }
    ~D2Scope() {
        if (--c->d2_scope == 0) {
            c->d2_valid = false;
            if (c->d2tab) {
                hipStreamSynchronize(c->stream);
                hipFree(c->d2tab);
            }
            c->d2tab = nullptr;
            c->d2tab_bytes = 0;
        }
    }
    D2Scope(const D2Scope&) = delete;
    D2Scope& operator=(const D2Scope&) = delete;
};

// Scope of a mixed-precision fixed-bandwidth search whose FP64 great-circle
// certification sweeps share one FP64 distance table (ensure_cert_table);
// the table is released when the search returns.
void release_cert_table(gwm_hip_ctx* c);
struct CertScope {
    gwm_hip_ctx* c;
    bool on;
    CertScope(gwm_hip_ctx* ctx, bool enable) : c(ctx), on(enable) {
        if (on && c->cert_scope++ == 0) {
            c->cert_hi = 0.0;
            c->cert.failed_bw = 0.0;
        }
    }
    ~CertScope() {
        if (on && --c->cert_scope == 0) release_cert_table(c);
    }
    CertScope(const CertScope&) = delete;
    CertScope& operator=(const CertScope&) = delete;
};

// Adaptive bandwidth -> per-focus thresholds (a vector of the bandwidth cache,
// filled by the select on a miss); fixed -> *bw_pt = nullptr.
int resolve_bandwidth(gwm_hip_ctx* c, int adaptive, double bw, const double** bw_pt);

// A product table in both representations.
struct ZPlanes {
    const double* z64;
    const _Float16* zh;
    const _Float16* zl;
    const double* scale;
};
const char* mode_name(int mode);
ZPlanes design_planes(gwm_hip_ctx* c);
ZPlanes g_planes(gwm_hip_ctx* c);
ZPlanes var_planes(gwm_hip_ctx* c);
// One sweep of the fused contraction into c->part (gwm_sweep.cpp).
int sweep(gwm_hip_ctx* c, int mode, int nta, int ntb, const ZPlanes& zp, const double* bw_pt, double bw, int kernel,
          int precision, int* PT_out, int* nseg_out, const double* dnn = nullptr);
SolveArgs solve_args(gwm_hip_ctx* c, int PT, int nseg, int colB, const double* bw_pt, double bw, int kernel);

// Device memory owned by one API call.
struct DevBuf {
    void* p = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() {
        if (p) hipFree(p);
    }
};

// Focus points of a prediction (gwm_hip_predict): their table in spatial
// order and the device arrays the solves write.
struct PredictSet {
    gwm::Points Q;
    double* cond;   // [m] conditioning indicator
    int* rlist;     // [m] refinement list
    double* betas;  // m x k column-major
};
// Grids below this many workgroups (4 per CU) are under-occupied: the
// autotuner compares launch shapes on them; larger pruned sweeps run their
// items in the work order of launch_sweep_plan.
constexpr int kTuneGridMax = 1024;

int refine_fp64(gwm_hip_ctx* c, int mode, int smode, int ntb, int colB, const double* bw_pt, double bw, int kernel,
                double limit, int* count, const PredictSet* q = nullptr, const double* dnn = nullptr);

}  // namespace detail
}  // namespace gwm
