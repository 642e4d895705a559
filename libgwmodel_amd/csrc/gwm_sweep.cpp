// The single-problem sweep launcher (fused distance -> weight -> W.Z
// contraction), its online autotuner and the FP64 recomputation of
// ill-conditioned focus points.
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>

#include "gwm_internal.h"

namespace {

using namespace gwm;

// ---------------------------------------------------------------- autotuner
//
// Launch shapes of a single-problem sweep that leave every result bit unchanged
// (each focus point sums its data in the same order whatever the shape):
//   0: default focus tiles per wave, tiles visited by the coprime stride
//   1: default tiles per wave, tiles visited in order
//   2: one focus tile per wave (where the default is two), stride order
//   3: one focus tile per wave, in order
// The shape matters where the default grid leaves the chip under-occupied:
// fewer than kTuneGridMax workgroups (4 per CU).  The first sweep of such a
// class (distance, mode, precision, column tiles, n to a power of two) runs
// every candidate kTuneSamples times on the same work, timed with HIP events,
// and the fastest is used from then on; the table is process-wide.  Larger
// grids keep the default, whose stride spreads pruned work over the XCDs.
struct TuneKey {
    int dist, mode, prec, nta, ntb, nlog2;
    bool operator<(const TuneKey& o) const {
        return std::tie(dist, mode, prec, nta, ntb, nlog2) < std::tie(o.dist, o.mode, o.prec, o.nta, o.ntb, o.nlog2);
    }
};
struct TuneState {
    double best_ms[4] = {INFINITY, INFINITY, INFINITY, INFINITY};
    int ncand = 1;
    int chosen = -1;
};
constexpr int kTuneSamples = 2;
std::mutex g_tune_mu;
std::map<TuneKey, TuneState> g_tune;

}  // namespace

namespace gwm {
namespace detail {

const char* mode_name(int mode) {
    switch (mode) {
    case kModeCV: return "cv";
    case kModeHat: return "hat";
    case kModeQDiag: return "qdiag";
    default: return "plain";
    }
}

void release_cert_table(gwm_hip_ctx* c) {
    auto& t = c->cert;
    if (t.tab || t.off || t.cnt) hipStreamSynchronize(c->stream);
    if (t.tab) hipFree(t.tab);
    if (t.off) hipFree(t.off);
    if (t.cnt) hipFree(t.cnt);
    t = gwm_hip_ctx::CertTable{};
    c->cert_hi = 0.0;
}

// The FP64 certification sweep of a mixed-precision fixed-bandwidth
// great-circle search (L: the sweep's single job at bandwidth bw) through
// the search's FP64 distance table, built here for the bracket's upper end
// when there is none for this sweep's mode, kernel and column tiles.  *used
// is 0 when the table does not fit (the caller then sweeps as usual).
static int cert_sweep(gwm_hip_ctx* c, SweepLaunch L, int mode, int nta, bool cv, int* used) {
    *used = 0;
    auto& t = c->cert;
    SweepJob& j = L.job0;
    const int r = nta <= 8 ? 2 : 1;   // the FP64 sweep's default shape
    const int tiles = (c->n + 64 * r - 1) / (64 * r);
    const size_t items = (size_t)tiles * j.nseg;
    const bool fits = t.valid && t.mode == mode && t.kernel == j.kernel && t.nta == nta && t.cv_dnn == cv &&
                      j.bw <= t.bw;
    if (!fits) {
        const double B = c->cert_hi;
        if (t.failed_bw > 0.0 && !(B < t.failed_bw)) return GWM_OK;
        t.valid = false;
        if (items > t.items_cap) {
            if (t.off) hipFree(t.off);
            if (t.cnt) hipFree(t.cnt);
            t.off = nullptr;
            t.cnt = nullptr;
            t.items_cap = 0;
            GWM_HIP_CHECK(c, hipMalloc((void**)&t.off, items * sizeof(long long)));
            GWM_HIP_CHECK(c, hipMalloc((void**)&t.cnt, items * sizeof(int)));
            t.items_cap = items;
        }
        SweepLaunch T = L;
        T.job0.bw = B;
        T.job0.pairs = nullptr;
        T.job0.tab_cnt = t.cnt;
        long long block = 0;
        GWM_HIP_CHECK(c, launch_contract_f64_table(3, mode, nta, T, c->stream, &block));
        std::vector<int> cnt(items);
        std::vector<long long> off(items);
        GWM_HIP_CHECK(c, hipMemcpyAsync(cnt.data(), t.cnt, items * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        long long blocks = 0;
        for (size_t q = 0; q < items; ++q) {
            off[q] = blocks;
            blocks += cnt[q];
        }
        const size_t bytes = (size_t)blocks * (size_t)block * sizeof(double);
        size_t free_b = 0, total_b = 0;
        if (bytes > c->cert_cap || hipMemGetInfo(&free_b, &total_b) != hipSuccess ||
            free_b + t.bytes < bytes + (size_t(8) << 30)) {
            (void)hipGetLastError();
            t.failed_bw = B;
            return GWM_OK;
        }
        if (t.bytes < bytes) {
            if (t.tab) hipFree(t.tab);
            t.tab = nullptr;
            t.bytes = 0;
            if (hipMalloc((void**)&t.tab, bytes) != hipSuccess) {
                (void)hipGetLastError();
                t.failed_bw = B;
                return GWM_OK;
            }
            t.bytes = bytes;
        }
        GWM_HIP_CHECK(c, hipMemcpyAsync(t.off, off.data(), items * sizeof(long long), hipMemcpyHostToDevice,
                                        c->stream));
        T.job0.f64tab = t.tab;
        T.job0.tab_off = t.off;
        timer_begin(c);
        GWM_HIP_CHECK(c, launch_contract_f64_table(1, mode, nta, T, c->stream));
        int rc = timer_end(c, "f64_table");
        if (rc) return rc;
        // the host offsets go out of scope
        GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        t.valid = true;
        t.bw = B;
        t.mode = mode;
        t.kernel = j.kernel;
        t.nta = nta;
        t.cv_dnn = cv;
    }
    j.f64tab = t.tab;
    j.tab_off = t.off;
    j.tab_cnt = t.cnt;
    j.tab_bw = t.bw;
    L.linear_order = 0;
    L.shape_r = 0;
    L.order = nullptr;
    if (tiles * j.nseg >= kTuneGridMax && j.nseg <= kPlanMaxSeg) {
        // longest first, as the other large pruned sweeps
        int rc;
        if ((rc = ensure_plan(c, items))) return rc;
        GWM_HIP_CHECK(c, launch_sweep_plan(c->dist, c->points(), c->n, c->boxes, nullptr, j.bw, nullptr, j.kernel, 0,
                                           64 * r, j.seg_len, j.nseg, tiles, c->plan_counts, c->plan_order,
                                           c->stream));
        L.order = c->plan_order;
        L.norder = tiles * j.nseg;
    }
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_contract_f64_table(2, mode, nta, L, c->stream));
    if (c->timing) {
        std::string name = std::string("contract_fp64_") + mode_name(mode);
        int rc;
        if ((rc = timer_end(c, name.c_str()))) return rc;
        unsigned long long pairs = 0;
        GWM_HIP_CHECK(c, hipMemcpyAsync(&pairs, c->pairs_dev, sizeof(pairs), hipMemcpyDeviceToHost, c->stream));
        GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        c->stats[name].pairs += (double)pairs;
    }
    *used = 1;
    return GWM_OK;
}

// One sweep of the fused contraction.
int sweep(gwm_hip_ctx* c, int mode, int nta, int ntb, const ZPlanes& zp, const double* bw_pt, double bw, int kernel,
          int precision, int* PT_out, int* nseg_out, const double* dnn) {
    const bool mixed = precision == GWM_MIXED;
    int nseg, seg_len;
    data_segments(c->n, nseg, seg_len);
    int rc;
    // FP64 single-column sweeps (multiscale steps): the two used columns of
    // the single-column table on the VALU, partial sums [nseg][2][n_pad]
    if (!mixed && nta == 1 && ntb == 0 && zp.z64 == c->vZ && (mode == kModeCV || mode == kModePlain)) {
        if ((rc = ensure_part(c, (size_t)nseg * 2 * c->n_pad * sizeof(double)))) return rc;
        SweepJob j;
        j.P = c->points();
        j.Z = c->vZ;
        j.Zh = nullptr;
        j.Zl = nullptr;
        j.colscale = nullptr;
        j.bw_pt = bw_pt;
        j.bw = bw;
        j.kernel = kernel;
        j.n = c->n;
        j.n_pad = c->n_pad;
        j.seg_len = seg_len;
        j.nseg = nseg;
        j.nta = 1;
        j.ntb = 0;
        j.out = c->part;
        j.dnn = dnn;
        if (c->prune) j.boxes = c->boxes;
        if (c->timing) {
            if (!c->pairs_dev) GWM_HIP_CHECK(c, hipMalloc((void**)&c->pairs_dev, sizeof(unsigned long long)));
            GWM_HIP_CHECK(c, hipMemsetAsync(c->pairs_dev, 0, sizeof(unsigned long long), c->stream));
            j.pairs = c->pairs_dev;
        }
        timer_begin(c);
        GWM_HIP_CHECK(c, launch_contract_var_f64(c->dist, mode, j, c->stream));
        if (c->timing) {
            std::string name = std::string("contract_fp64_") + mode_name(mode);
            if ((rc = timer_end(c, name.c_str()))) return rc;
            unsigned long long pairs = 0;
            GWM_HIP_CHECK(c, hipMemcpyAsync(&pairs, c->pairs_dev, sizeof(pairs), hipMemcpyDeviceToHost, c->stream));
            GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
            c->stats[name].pairs += (double)pairs;
        }
        *PT_out = 2;
        *nseg_out = nseg;
        return GWM_OK;
    }
    const int PT = (nta + ntb) * 16;
    if ((rc = ensure_part(c, (size_t)nseg * PT * c->n_pad * sizeof(double)))) return rc;
    SweepLaunch L;
    SweepJob& j = L.job0;
    j.P = c->points();
    j.Z = zp.z64;
    j.Zh = zp.zh;
    j.Zl = zp.zl;
    j.colscale = zp.scale;
    j.bw_pt = bw_pt;
    j.bw = bw;
    j.kernel = kernel;
    j.n = c->n;
    j.n_pad = c->n_pad;
    j.seg_len = seg_len;
    j.nseg = nseg;
    j.nta = nta;
    j.ntb = ntb;
    j.out = c->part;
    j.dnn = dnn;
    if (c->prune) {
        j.boxes = c->boxes;
        if (mode == kModeQDiag && bw_pt) {
            GWM_HIP_CHECK(c, launch_chunk_max(bw_pt, c->n, c->chunk_bw, c->stream));
            j.chunk_bw = c->chunk_bw;
        }
    }
    L.jobs = nullptr;
    L.items = nullptr;
    L.nitems = 0;
    // mixed great-circle sweeps inside a search / fit / backfit call read the
    // pairs' squared distances from the table (same values, computed once)
    int ldist = c->dist;
    if (mixed && c->dist == kGreatCircle) {
        if ((rc = ensure_d2_table(c))) return rc;
        if (c->d2_valid) {
            j.d2tab = c->d2tab;
            j.d2_ndb = c->rows_pad / 32;
            ldist = kGcTable;
        }
    }
    if (c->timing) {
        if (!c->pairs_dev) GWM_HIP_CHECK(c, hipMalloc((void**)&c->pairs_dev, sizeof(unsigned long long)));
        GWM_HIP_CHECK(c, hipMemsetAsync(c->pairs_dev, 0, sizeof(unsigned long long), c->stream));
        j.pairs = c->pairs_dev;
    }
    // FP64 certification sweeps of a mixed-precision fixed-bandwidth
    // great-circle search: through the search's FP64 distance table
    if (!mixed && c->cert_scope > 0 && c->dist == kGreatCircle && !bw_pt && c->prune && ntb == 0 && nta >= 1 &&
        nta <= 10 && (mode == kModeCV || mode == kModePlain) && bw <= c->cert_hi && c->force_shape < 0) {
        int used = 0;
        if ((rc = cert_sweep(c, L, mode, nta, dnn != nullptr, &used))) return rc;
        if (used) {
            *PT_out = PT;
            *nseg_out = nseg;
            return GWM_OK;
        }
    }
    // launch shape: forced, or from the autotuner (results are the same for every shape)
    if (c->force_shape >= 0) {
        L.linear_order = c->force_shape & 1;
        L.shape_r = c->force_shape >= 2 ? 1 : 0;
    } else if (c->autotune) {
        TuneKey key{ldist, mode, mixed ? 1 : 0, nta, ntb, 0};
        while ((2 << key.nlog2) <= c->n) ++key.nlog2;
        int chosen;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune.find(key);
            chosen = it == g_tune.end() ? -1 : it->second.chosen;
        }
        if (chosen < 0) {
            const int r = contract_r(nta, ntb, mixed ? 1 : 0, mode, ldist);
            const int grid = (c->n + 64 * r - 1) / (64 * r) * nseg;
            TuneState st;
            st.ncand = grid >= kTuneGridMax ? 1 : (r == 2 ? 4 : 2);
            if (st.ncand > 1) {
                if (!c->tev0) GWM_HIP_CHECK(c, hipEventCreate(&c->tev0));
                if (!c->tev1) GWM_HIP_CHECK(c, hipEventCreate(&c->tev1));
                SweepLaunch T = L;
                T.job0.pairs = nullptr;
                for (int q = 0; q < st.ncand; ++q)
                    for (int rep = 0; rep < kTuneSamples; ++rep) {
                        T.linear_order = q & 1;
                        T.shape_r = q >= 2 ? 1 : 0;
                        float ms = 0.0f;
                        GWM_HIP_CHECK(c, hipEventRecord(c->tev0, c->stream));
                        if (mixed) GWM_HIP_CHECK(c, launch_contract_mixed(ldist, mode, nta, ntb, T, c->stream));
                        else GWM_HIP_CHECK(c, launch_contract_f64(c->dist, mode, nta, ntb, T, c->stream));
                        GWM_HIP_CHECK(c, hipEventRecord(c->tev1, c->stream));
                        GWM_HIP_CHECK(c, hipEventSynchronize(c->tev1));
                        GWM_HIP_CHECK(c, hipEventElapsedTime(&ms, c->tev0, c->tev1));
                        st.best_ms[q] = std::min(st.best_ms[q], (double)ms);
                    }
            }
            int best = 0;
            for (int q = 1; q < st.ncand; ++q)
                if (st.best_ms[q] < st.best_ms[best]) best = q;
            st.chosen = best;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune.emplace(key, st);
            chosen = g_tune[key].chosen;
        }
        L.linear_order = chosen & 1;
        L.shape_r = chosen >= 2 ? 1 : 0;
        // Large pruned sweeps run their (tile, segment) items longest first
        // (gwm_plan.hip): narrow bandwidths leave a few busy items per tile,
        // and in plain order the sweep ends on a tail of them.
        const int r = contract_r(nta, ntb, mixed ? 1 : 0, mode, ldist);
        const int tiles = (c->n + 64 * r - 1) / (64 * r);
        if (c->prune && chosen == 0 && tiles * nseg >= kTuneGridMax && nseg <= kPlanMaxSeg) {
            if ((rc = ensure_plan(c, (size_t)tiles * nseg))) return rc;
            // qDiag thresholds come from the streamed points' bandwidths (chunk_bw)
            const double* focus_bw = mode == kModeQDiag ? nullptr : bw_pt;
            GWM_HIP_CHECK(c, launch_sweep_plan(c->dist, c->points(), c->n, c->boxes, focus_bw, bw, j.chunk_bw, kernel,
                                               mixed ? 1 : 0, 64 * r, seg_len, nseg, tiles, c->plan_counts,
                                               c->plan_order, c->stream));
            L.order = c->plan_order;
            L.norder = tiles * nseg;
        }
    }
    timer_begin(c);
    if (mixed) GWM_HIP_CHECK(c, launch_contract_mixed(ldist, mode, nta, ntb, L, c->stream));
    else GWM_HIP_CHECK(c, launch_contract_f64(c->dist, mode, nta, ntb, L, c->stream));
    if (c->timing) {
        std::string name = std::string("contract_") + (mixed ? "mixed_" : "fp64_") + mode_name(mode);
        if ((rc = timer_end(c, name.c_str()))) return rc;
        unsigned long long pairs = 0;
        GWM_HIP_CHECK(c, hipMemcpyAsync(&pairs, c->pairs_dev, sizeof(pairs), hipMemcpyDeviceToHost, c->stream));
        GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        c->stats[name].pairs += (double)pairs;
    }
    *PT_out = PT;
    *nseg_out = nseg;
    return GWM_OK;
}

ZPlanes design_planes(gwm_hip_ctx* c) { return ZPlanes{c->Z, c->Zh, c->Zl, c->zscale}; }
ZPlanes g_planes(gwm_hip_ctx* c) { return ZPlanes{c->G, c->Gh, c->Gl, c->gscale}; }
ZPlanes var_planes(gwm_hip_ctx* c) { return ZPlanes{c->vZ, c->vZh, c->vZl, c->vscale}; }

SolveArgs solve_args(gwm_hip_ctx* c, int PT, int nseg, int colB, const double* bw_pt, double bw, int kernel) {
    SolveArgs s;
    s.part = c->part;
    s.nseg = nseg;
    s.PT = PT;
    s.colB = colB;
    s.n = c->n;
    s.n_pad = c->n_pad;
    s.X = c->X;
    s.y = c->y;
    s.bw_pt = bw_pt;
    s.bw = bw;
    s.kernel = kernel;
    s.res2 = c->res2;
    s.sii = c->sii;
    s.ssi = c->ssi;
    s.betas = c->betas;
    s.se_raw = c->se;
    s.G = c->G;
    s.ldg = c->ntb * 16;
    s.gvec = c->gvec;
    s.fail = c->fail;
    s.cond = c->cond;
    return s;
}

// After a mixed-precision sweep and solve over all points: recompute in FP64
// the focus points whose local design is ill conditioned (indicator above
// limit, or not SPD), so that their results are those of the FP64 path (same
// kernel, same data segments, so the same bits).  The failure flag is left as
// the FP64 solves set it: every point the mixed solves failed on is in the
// list.  *count receives the number of recomputed points; the indicator
// maximum of the mixed pass goes to c->last_cond.  q: the focus points are
// prediction locations rather than the data points.
int refine_fp64(gwm_hip_ctx* c, int mode, int smode, int ntb, int colB, const double* bw_pt, double bw, int kernel,
                double limit, int* count, const PredictSet* q, const double* dnn) {
    const int nf = q ? q->Q.n : c->n;
    double* cond = q ? q->cond : c->cond;
    int* rlist = q ? q->rlist : c->rlist;
    GWM_HIP_CHECK(c, launch_reduce(nullptr, nullptr, nullptr, cond, nf, c->red, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(&c->last_cond, c->red + 3, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipMemsetAsync(c->rcount, 0, sizeof(int), c->stream));
    GWM_HIP_CHECK(c, launch_compact(cond, nf, limit, rlist, c->rcount, c->stream));
    GWM_HIP_CHECK(c, hipMemcpyAsync(count, c->rcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    GWM_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->last_refined = *count;
    if (*count == 0) return GWM_OK;
    const int nta = c->nta;
    int nseg, seg_len;
    data_segments(c->n, nseg, seg_len);
    const int PT = (nta + ntb) * 16;
    const int npad = (*count + 127) / 128 * 128;
    const size_t bytes = (size_t)nseg * PT * npad * sizeof(double);
    if (bytes > c->rpart_bytes) {
        cfree(c, c->rpart);
        c->rpart_bytes = 0;
        int rc = dalloc(c, reinterpret_cast<char**>(&c->rpart), bytes);
        if (rc) return rc;
        c->rpart_bytes = bytes;
    }
    SweepLaunch L;
    SweepJob& j = L.job0;
    j.P = c->points();
    j.Z = c->Z;
    j.Zh = nullptr;
    j.Zl = nullptr;
    j.colscale = nullptr;
    j.bw_pt = bw_pt;
    j.bw = bw;
    j.kernel = kernel;
    j.n = c->n;
    j.n_pad = npad;
    j.seg_len = seg_len;
    j.nseg = nseg;
    j.nta = nta;
    j.ntb = ntb;
    j.out = c->rpart;
    j.focus_idx = rlist;
    j.nfocus = *count;
    j.dnn = dnn;
    if (q) j.Q = q->Q;
    if (c->prune) j.boxes = c->boxes;
    L.jobs = nullptr;
    L.items = nullptr;
    L.nitems = 0;
    timer_begin(c);
    GWM_HIP_CHECK(c, launch_contract_f64(c->dist, mode, nta, ntb, L, c->stream));
    int rc = timer_end(c, "refine");
    if (rc) return rc;
    GWM_HIP_CHECK(c, hipMemsetAsync(c->fail, 0, sizeof(int), c->stream));
    SolveArgs s = solve_args(c, PT, nseg, colB, bw_pt, bw, kernel);
    s.part = c->rpart;
    s.n_pad = npad;
    s.focus_idx = rlist;
    s.nfocus = *count;
    if (q) {
        s.n = q->Q.n;
        s.X = nullptr;
        s.y = nullptr;
        s.betas = q->betas;
        s.cond = q->cond;
    }
    GWM_HIP_CHECK(c, launch_solve(c->k, smode, s, c->stream));
    return GWM_OK;
}

}  // namespace detail
}  // namespace gwm

extern "C" {

int gwm_hip_autotune_report(char* buf, int cap) {
    if (!buf || cap <= 0) return GWM_ERR_INVALID;
    static const char* modes[] = {"cv", "hat", "qdiag", "plain"};   // ContractMode order
    std::string out;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const auto& kv : g_tune) {
            const TuneKey& k = kv.first;
            const TuneState& st = kv.second;
            char line[256];
            int m = std::snprintf(line, sizeof(line), "%s %s %s nta=%d ntb=%d n<2^%d chosen=%d ms=",
                                  k.dist == kGcTable ? "great_circle_table"
                                                     : (k.dist == GWM_GREAT_CIRCLE ? "great_circle" : "euclidean"),
                                  k.mode >= 0 && k.mode < 4 ? modes[k.mode] : "?", k.prec ? "mixed" : "fp64", k.nta,
                                  k.ntb, k.nlog2 + 1, st.chosen);
            for (int q = 0; q < st.ncand && m < (int)sizeof(line) - 16; ++q)
                m += std::snprintf(line + m, sizeof(line) - m, q ? ",%.4f" : "%.4f",
                                   std::isfinite(st.best_ms[q]) ? st.best_ms[q] : -1.0);
            out += line;
            out += "\n";
        }
    }
    std::snprintf(buf, cap, "%s", out.c_str());
    return GWM_OK;
}

}  // extern "C"
