// Continuous-batching scheduler: many GWR problems in flight on one device.
//
// The reference runs one problem per fit() call (src/gwmodelpp/GWRBasic.cpp:36-128)
// and has no batching or serving layer; this file adds one.  Each request (a
// GWRBasic fit at a given bandwidth, or a golden-section bandwidth search
// followed by the fit) runs the unchanged host algorithm -- the golden section
// of BandwidthSelector::optimize (BandwidthSelector.cpp:6-81) and the fit /
// diagnostics of GWRBasic::fit -- on a worker thread of its own.  Every
// criterion or fit evaluation the worker needs is posted to the scheduler and
// the worker blocks.  The scheduler runs in rounds: once every in-flight
// request has posted (or finished), it executes a slice of every posted
// evaluation as batched launches -- one adaptive-bandwidth launch, one sweep
// launch per kernel class, one solve launch per k, one reduction launch -- on
// its stream, and hands back the evaluations that completed.  A round has a
// work budget (about round_ms of device time): the posted evaluations share it
// shortest-remaining-first, each getting at least one chunk of focus points,
// so an evaluation of a small problem completes in one short round while a
// large fit next to it advances chunk by chunk.  Between rounds the scheduler
// retires finished requests and admits waiting ones while their device
// footprint (O(n k^2): coordinates, design, product table and its FP16 planes,
// partial sums; never n^2) fits the memory budget.
//
// Each request's arithmetic depends only on its own data: its data segments
// are a function of its n alone (those of the single-problem API), a focus
// point's partial sums and solve do not depend on which round computes them,
// and every reduction has a fixed order, so a request returns the same bits
// alone, in any mix, or through the single-problem API.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <tuple>
#include <mutex>
#include <thread>
#include <vector>

#include "gwm_ctx.h"

using namespace gwm;
using namespace gwm::detail;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum { kEvalCrit = 0, kEvalFit = 1 };

// Phases of an evaluation.  Mixed precision adds a check at the end of the
// main sweep (which focus points have an ill-conditioned local design?) and,
// when there are any, an FP64 recomputation of those points.
enum Phase {
    kMain = 1,     // adaptive bandwidths, sweep and solves over the focus points
    kQDiag = 2,    // fits: second sweep and qDiag
    kDone = 3,
    kRefine = 4,   // FP64 sweep and solves over the listed focus points
    kCheck = 5     // main sweep done; the list is counted at the end of the round
};

// Slicing of evaluations across rounds.  Work is counted in units of one data
// point x one 16-column tile of one focus point in the mixed sweep; an FP64
// sweep tile costs about 3.5 units and an adaptive-bandwidth selection about 4
// units per data point.  The device does about 1.5e9 units per ms.
constexpr int kChunk = 128;              // focus points; a multiple of every sweep's 64 R
constexpr double kUnitsPerMs = 1.5e9;
constexpr double kFp64Factor = 3.5;
constexpr double kGatherMs = 2.0;   // longest wait for the last posts before a round starts
constexpr double kGatherBusyMs = 0.5;   // the same while another round is on the device
constexpr double kSelectUnits = 4.0;

struct Eval {
    int type = kEvalCrit;
    double bw = 0.0;
    int precision = GWM_FP64;
    int crit = GWM_CRIT_CV;
    int rc = GWM_OK;
    double value = DBL_MAX;
    double cond = 0.0;   // conditioning indicator maximum (criterion evaluations)
    double bound = 0.0;  // error bound of a mixed criterion value (crit_error_bound)
    double shat[2] = {0.0, 0.0};
};

struct Request {
    long long id = 0;
    int n = 0, k = 0, dist = 0, kernel = 0, adaptive = 0, select_bw = 0, crit = 0;
    double bw = 0.0, lower = 0.0, upper = 0.0;
    int has_lower = 0, has_upper = 0;
    std::vector<double> coords, X, y;
    double *o_betas = nullptr, *o_se = nullptr, *o_shat = nullptr, *o_qdiag = nullptr, *o_diag = nullptr,
           *o_bw = nullptr;
    // device state (from admission to retirement)
    gwm_hip_ctx* ctx = nullptr;
    double* part = nullptr;
    int* dfail = nullptr;        // non-SPD flag of the evaluation in progress
    int nseg = 1, seg_len = 64;
    // partial sums hold one round's slice: at most s_cap focus points (or
    // list entries), a multiple of kChunk
    int s_cap = 128;
    size_t footprint = 0;
    double maxdist = 0.0;
    // timing and status
    double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
    int status = GWM_OK;
    bool admitted = false, finished = false, done = false;
    long long solves = 0;
    // worker handshake
    std::thread worker;
    bool posted = false;
    bool in_round = false;       // its posted evaluation has a slice in a round not yet finished
    std::condition_variable wcv;   // the worker waits here for its posted evaluation
    Eval ev;
    // progress of the posted evaluation (Phase): focus points, list entries or
    // qDiag points [0, next) of the current phase are done
    int phase = 1, next = 0;
    int nrefine = 0;             // focus points recomputed in FP64 (mixed)
    double refine_crit = 10.0, refine_fit = 2.0;
    // host staging of fit outputs
    std::vector<double> hbetas, hse, hq;
};


size_t footprint_bytes(int n, int k, int dist, int nseg, int s_cap) {
    const size_t rows = (size_t)(n + 63) / 64 * 64, np = (size_t)s_cap;
    const size_t nta = nta_of(k), ntb = ntb_of(k);
    size_t b = 0;
    b += (size_t)n * (dist == GWM_GREAT_CIRCLE ? 6 : 2) * 8;        // coordinates
    b += (size_t)n * (k + 1) * 8;                                       // X, y
    b += rows * nta * 16 * (8 + 4) + rows * ntb * 16 * (8 + 4);        // Z, G and FP16 planes
    b += rows * 16 * (8 + 4) + (size_t)n * 16;                          // single-column tables
    b += (size_t)n * 8 * 8 + (size_t)n * k * 3 * 8;                     // per-focus vectors
    b += (size_t)nseg * (nta + ntb) * 16 * np * 8;                      // partial sums of a slice
    return b;
}

// Focus points a round may compute for one request: all of them (padded to
// 128) while their partial sums take at most `budget` bytes, else as many as
// fit it (whole chunks of 128, at least one), so that a request's footprint
// stays O(n k^2) with a bounded partial-sum term (default 256 MiB: at
// n = 200,000, k = 16, 4,352 points of 25 segments x 304 columns, 265 MB
// instead of 12.2 GB).
int slice_cap(int n, int k, int nseg, size_t budget) {
    const size_t per = (size_t)nseg * (nta_of(k) + ntb_of(k)) * 16 * sizeof(double);
    const size_t all = (size_t)(n + 127) / 128 * 128;
    const size_t fit = std::max<size_t>(1, budget / per / 128) * 128;
    return (int)std::min(all, fit);
}

// Device arena for the per-round job tables.
struct Arena {
    char* base = nullptr;
    size_t cap = 0;
};

struct Request;

// Online tuning of batched sweep launches.  A class (distance, mode,
// precision, column tiles) whose default is two focus tiles per wave is
// launched alternately with two and with one for its first kTuneRounds
// rounds each (launches of at least kTuneMinItems workgroups), timed with
// events and normalised by the pairs the launch swept (counted on the
// device); the faster shape is kept for the server's lifetime.  Every
// focus point sums its data in the same order in either shape, so results do
// not depend on the choice.
struct TuneClass {
    int dist, mode, prec, nta, ntb;
    bool operator<(const TuneClass& o) const {
        return std::tie(dist, mode, prec, nta, ntb) < std::tie(o.dist, o.mode, o.prec, o.nta, o.ntb);
    }
};
struct SrvTune {
    double ms[3] = {0.0, 0.0, 0.0}, units[3] = {0.0, 0.0, 0.0};
    int samples[3] = {0, 0, 0};
    int chosen = -1;   // focus tiles per wave once decided
};
constexpr int kTuneRounds = 3;
// launches smaller than this (workgroups) are dominated by their fixed
// latency, whatever the shape: they are not samples
constexpr int kTuneMinItems = 512;
struct TimedGroup {
    TuneClass key{};
    int R = 2;
    size_t off_pairs = 0;                      // its swept-pair count in the slot's arena
    unsigned long long* pairs_host = nullptr;  // pinned read-back of that count
    hipEvent_t a = nullptr, b = nullptr;
};

// One round in flight.  Two slots alternate: while the device runs one
// round, the host finishes the previous one and launches the next, so the
// stream always holds work.  Tables go up from pinned host memory and results
// come back after the round's end event on a stream of their own, so neither
// copy waits behind the round queued after it.
struct RoundSlot {
    Arena arena;                       // job tables and compaction counts on the device
    char* host = nullptr;              // pinned copy of the tables
    size_t host_cap = 0;
    double* red = nullptr;             // device reductions of the completed evaluations
    int red_cap = 0;
    double* red_host = nullptr;        // pinned
    int* cnt_host = nullptr;           // pinned compaction counts
    int cnt_cap = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool busy = false;
    std::vector<Request*> batch, checks, finished;
    size_t off_cnt = 0;
    std::vector<TimedGroup> timed;     // sweep launches timed for the tuner (first ntimed)
    int ntimed = 0;
};

}  // namespace

struct gwm_hip_server {
    int device = 0;
    int precision = GWM_FP64;
    size_t budget = 0;
    int max_inflight = 1;
    double certify_rel = 1e-7;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Request*> pending;
    std::vector<Request*> active;
    std::map<long long, Request*> reqs;
    long long next_id = 1;
    bool stopping = false;
    std::thread sched;
    size_t inflight_bytes = 0, peak_bytes = 0;
    int peak_inflight = 0;
    long long rounds = 0, solves = 0;
    double busy_ms = 0.0;
    std::string err;
    RoundSlot slots[2];
    int slot_next = 0;
    std::map<TuneClass, SrvTune> tune;   // the batched-sweep tuner
    std::mutex tune_mu;
    hipStream_t rb_stream = nullptr;   // read-back of finished rounds
    int* sel_left = nullptr;           // scratch of the batched adaptive-bandwidth select
    unsigned int* sel_nleft = nullptr;
    int sel_cap = 0;
    double round_ms = 25.0;      // work budget of a round; <= 0: whole evaluations
    size_t part_budget = size_t(256) << 20;   // partial-sum bytes of one request's round slice
    double refine_crit = 10.0, refine_fit = 2.0;   // FP64 recomputation limits (mixed)
    // Request set-ups (sort, boxes, uploads, the first select) run a few at a
    // time, the costliest waiting one first (n^2 k): a burst of admissions
    // then has requests ready for the device early instead of all of them
    // finishing together, and the long searches start first.
    std::mutex setup_mu;
    std::condition_variable setup_cv;
    std::vector<std::pair<double, Request*>> setup_waiting;
    int setup_running = 0;
};

namespace {

#define SRV_CHECK(S, expr)                                                 \
    do {                                                                   \
        hipError_t e__ = (expr);                                           \
        if (e__ != hipSuccess) {                                           \
            (S)->err = std::string(#expr) + ": " + hipGetErrorString(e__); \
            return e__ == hipErrorOutOfMemory ? GWM_ERR_OOM : GWM_ERR_HIP; \
        }                                                                  \
    } while (0)

// Host-side job tables of one round, uploaded into the arena in one pass.
struct Tables {
    std::vector<char> host;
    size_t push(const void* p, size_t bytes) {
        size_t off = (host.size() + 255) / 256 * 256;
        host.resize(off + bytes);
        std::memcpy(host.data() + off, p, bytes);
        return off;
    }
};

struct SweepGroup {
    int dist, mode, precision, nta, ntb;
    std::vector<SweepJob> jobs;
    std::vector<int> items;
    size_t off_jobs = 0, off_items = 0;
    int R = 1;            // focus tiles per wave of this launch (the tuner's choice)
    bool timed = false;   // a launch the tuner times: its jobs count the pairs they sweep
    size_t off_pairs = 0; // that count in the arena
};

struct SolveGroup {
    int k, mode;
    std::vector<SolveArgs> jobs;
    std::vector<int> start;
    size_t off_jobs = 0, off_start = 0;
};

struct QdGroup {
    int k;
    std::vector<QDiagJob> jobs;
    std::vector<int> start;
    size_t off_jobs = 0, off_start = 0;
};

// f_off: the first focus point (or list entry) of the round's slice, the
// first column of the request's partial-sum buffer
SweepJob sweep_job(Request* r, int mode, int nta, int ntb, const double* bw_pt, double bw, int f_off) {
    gwm_hip_ctx* c = r->ctx;
    SweepJob j;
    j.P = c->points();
    if (mode == kModeQDiag) {
        j.Z = c->G;
        j.Zh = c->Gh;
        j.Zl = c->Gl;
        j.colscale = c->gscale;
    } else {
        j.Z = c->Z;
        j.Zh = c->Zh;
        j.Zl = c->Zl;
        j.colscale = c->zscale;
    }
    j.bw_pt = bw_pt;
    j.bw = bw;
    j.kernel = r->kernel;
    j.n = c->n;
    j.n_pad = r->s_cap;
    j.out_off = f_off;
    j.seg_len = r->seg_len;
    j.nseg = r->nseg;
    j.nta = nta;
    j.ntb = ntb;
    j.out = r->part;
    if (mode == kModeCV) j.dnn = c->dnn;   // weights relative to the nearest other point
    if (c->prune) {
        j.boxes = c->boxes;
        if (mode == kModeQDiag && bw_pt) j.chunk_bw = c->chunk_bw;
    }
    return j;
}

// Sweep items of focus points [f0, f1) (f0 a multiple of 64 R), tiles in the
// spread order of tile_stride.
void add_items(SweepGroup& g, int job, int f0, int f1, int nseg, int R) {
    const int t0 = f0 / (64 * R), t1 = (f1 + 64 * R - 1) / (64 * R);
    const int cnt = t1 - t0, stride = tile_stride(cnt);
    for (int u = 0; u < cnt; ++u) {
        const int t = t0 + (int)(((long long)u * stride) % cnt);
        for (int s = 0; s < nseg; ++s) {
            g.items.push_back(job);
            g.items.push_back(t);
            g.items.push_back(s);
        }
    }
}

// Focus tiles per wave of a batched launch of class (dist, mode, prec, nta,
// ntb): the tuner's choice, or while it is undecided the shape with fewer
// timed rounds so far (the default first).
int tuned_r(gwm_hip_server* S, int dist, int mode, int prec, int nta, int ntb, bool* timed) {
    *timed = false;
    const int rd = contract_r(nta, ntb, prec == GWM_MIXED, mode, dist);
    if (rd != 2) return rd;
    std::lock_guard<std::mutex> lk(S->tune_mu);
    const SrvTune& t = S->tune[TuneClass{dist, mode, prec, nta, ntb}];
    if (t.chosen > 0) return t.chosen;
    *timed = true;
    return t.samples[2] <= t.samples[1] ? 2 : 1;
}

SweepGroup& group_of(gwm_hip_server* S, std::vector<SweepGroup>& gs, int dist, int mode, int prec, int nta,
                     int ntb) {
    for (auto& g : gs)
        if (g.dist == dist && g.mode == mode && g.precision == prec && g.nta == nta && g.ntb == ntb) return g;
    gs.push_back(SweepGroup{dist, mode, prec, nta, ntb, {}, {}, 0, 0});
    gs.back().R = tuned_r(S, dist, mode, prec, nta, ntb, &gs.back().timed);
    return gs.back();
}

template <class G>
G& kgroup_of(std::vector<G>& gs, int k) {
    for (auto& g : gs)
        if (g.k == k) return g;
    gs.push_back(G{});
    gs.back().k = k;
    return gs.back();
}

SolveGroup& sgroup_of(std::vector<SolveGroup>& gs, int k, int mode) {
    for (auto& g : gs)
        if (g.k == k && g.mode == mode) return g;
    gs.push_back(SolveGroup{k, mode, {}, {}, 0, 0});
    return gs.back();
}

// Launches a round's sweep groups; groups of classes still being tuned are
// timed with events of slot R (read by finish_round).
int launch_sweeps(gwm_hip_server* S, std::vector<SweepGroup>& gs, const char* dev, RoundSlot& R) {
    for (auto& g : gs) {
        SweepLaunch L;
        std::memset(&L.job0, 0, sizeof(L.job0));
        L.jobs = reinterpret_cast<const SweepJob*>(dev + g.off_jobs);
        L.items = reinterpret_cast<const int*>(dev + g.off_items);
        L.nitems = (int)g.items.size() / 3;
        const int rd = contract_r(g.nta, g.ntb, g.precision == GWM_MIXED, g.mode, g.dist);
        L.shape_r = g.R < rd ? 1 : 0;
        const TuneClass key{g.dist, g.mode, g.precision, g.nta, g.ntb};
        TimedGroup* tg = nullptr;
        if (g.timed && L.nitems >= kTuneMinItems) {
            if (R.ntimed == (int)R.timed.size()) {
                R.timed.emplace_back();
                SRV_CHECK(S, hipEventCreate(&R.timed.back().a));
                SRV_CHECK(S, hipEventCreate(&R.timed.back().b));
                SRV_CHECK(S, hipHostMalloc((void**)&R.timed.back().pairs_host, sizeof(unsigned long long),
                                           hipHostMallocDefault));
            }
            tg = &R.timed[R.ntimed++];
            tg->key = key;
            tg->R = g.R;
            tg->off_pairs = g.off_pairs;
            SRV_CHECK(S, hipEventRecord(tg->a, S->stream));
        }
        if (g.precision == GWM_MIXED) SRV_CHECK(S, launch_contract_mixed(g.dist, g.mode, g.nta, g.ntb, L, S->stream));
        else SRV_CHECK(S, launch_contract_f64(g.dist, g.mode, g.nta, g.ntb, L, S->stream));
        if (tg) SRV_CHECK(S, hipEventRecord(tg->b, S->stream));
    }
    return GWM_OK;
}

// Work units per focus point of the main sweep (selection, sweep, solve), per
// listed point of the FP64 recomputation, and per point of the qDiag sweep of
// r's posted evaluation.
double unit_cost(const Request* r, int phase) {
    const Eval& ev = r->ev;
    const double tiles = ev.type == kEvalCrit ? nta_of(r->k) : nta_of(r->k) + ntb_of(r->k);
    if (phase == kMain) {
        const double pf = ev.precision == GWM_FP64 ? kFp64Factor : 1.0;
        return r->n * (tiles * pf + (r->adaptive ? kSelectUnits : 0.0));
    }
    if (phase == kRefine) return r->n * tiles * kFp64Factor;
    const double pf = ev.precision == GWM_FP64 ? kFp64Factor : 1.0;
    return (double)r->n * ntb_of(r->k) * pf;
}

double remaining_cost(const Request* r) {
    const double q = r->ev.type == kEvalFit ? r->n * unit_cost(r, kQDiag) : 0.0;
    if (r->phase == kMain) return (r->n - r->next) * unit_cost(r, kMain) + q;
    if (r->phase == kRefine) return (r->nrefine - r->next) * unit_cost(r, kRefine) + q;
    return (r->n - r->next) * unit_cost(r, kQDiag);
}

// What one round does for one evaluation.
struct Slice {
    Request* r = nullptr;
    int a0 = 0, a1 = 0;   // main-sweep focus range
    int e0 = 0, e1 = 0;   // range of list entries recomputed in FP64
    int q0 = 0, q1 = 0;   // qDiag point range
    bool split = false;   // qDiag starts: the mixed path splits G into FP16 planes first
    bool check = false;   // main sweep completes (mixed): count the ill-conditioned points
    bool done = false;    // the evaluation completes in this round (after a check: if none)
    double cost = 0.0;
};

// Advances r by about `units` of work, at least one chunk (or the rest).
Slice plan_slice(Request* r, double units, bool unlimited) {
    Slice sl;
    sl.r = r;
    const bool fit = r->ev.type == kEvalFit;
    auto count = [&](int phase, int total) {
        const int left = total - r->next;
        if (unlimited) return std::min(left, r->s_cap);
        const long long want = (long long)std::ceil(std::max(units, 0.0) / unit_cost(r, phase));
        const long long chunks = std::max<long long>(1, (want + kChunk - 1) / kChunk);
        return (int)std::min<long long>(std::min<long long>(left, chunks * kChunk), r->s_cap);
    };
    auto take = [&](int phase, int total, int& lo, int& hi) {
        const int c = count(phase, total);
        lo = r->next;
        hi = r->next + c;
        sl.cost += c * unit_cost(r, phase);
        units -= c * unit_cost(r, phase);
        r->next = hi;
        return r->next == total;
    };
    if (r->phase == kMain) {
        if (!take(kMain, r->n, sl.a0, sl.a1)) return sl;
        if (r->ev.precision == GWM_MIXED) {
            // the refinement list is known after this round; a criterion
            // completes now if it is empty
            sl.check = true;
            sl.done = !fit;
            r->phase = kCheck;
            return sl;
        }
        if (!fit) {
            r->phase = kDone;
            sl.done = true;
            return sl;
        }
        r->phase = kQDiag;
        r->next = 0;
        if (!unlimited && units <= 0.0) return sl;
    } else if (r->phase == kRefine) {
        if (!take(kRefine, r->nrefine, sl.e0, sl.e1)) return sl;
        if (!fit) {
            r->phase = kDone;
            sl.done = true;
            return sl;
        }
        r->phase = kQDiag;
        r->next = 0;
        if (!unlimited && units <= 0.0) return sl;
    }
    sl.split = r->next == 0;
    if (take(kQDiag, r->n, sl.q0, sl.q1)) {
        r->phase = kDone;
        sl.done = true;
    }
    return sl;
}

SolveArgs solve_job(Request* r, int PT, int colB, const double* bw_pt, int f0) {
    gwm_hip_ctx* c = r->ctx;
    SolveArgs sa;
    sa.part = r->part;
    sa.nseg = r->nseg;
    sa.PT = PT;
    sa.colB = colB;
    sa.n = c->n;
    sa.n_pad = r->s_cap;
    sa.part_off = f0;
    sa.X = c->X;
    sa.y = c->y;
    sa.bw_pt = bw_pt;
    sa.bw = r->ev.bw;
    sa.kernel = r->kernel;
    sa.res2 = c->res2;
    sa.sii = c->sii;
    sa.ssi = c->ssi;
    sa.betas = c->betas;
    sa.se_raw = c->se;
    sa.G = c->G;
    sa.ldg = ntb_of(r->k) * 16;
    sa.gvec = c->gvec;
    sa.fail = r->dfail;
    sa.f0 = f0;
    sa.cond = c->cond;
    return sa;
}

// Executes one round: a slice of every posted evaluation.  Evaluations that
// complete get their results and phase kDone.
// Plans and launches one round into slot R (the slot is idle); its results
// are read back by finish_round.
int launch_round(gwm_hip_server* S, std::vector<Request*>& batch, RoundSlot& R) {
    const int ne = (int)batch.size();
    if (ne > R.red_cap) {
        if (R.red) hipFree(R.red);
        if (R.red_host) hipHostFree(R.red_host);
        R.red = nullptr;
        R.red_host = nullptr;
        R.red_cap = 0;
        SRV_CHECK(S, hipMalloc((void**)&R.red, (size_t)ne * 2 * 5 * sizeof(double)));
        SRV_CHECK(S, hipHostMalloc((void**)&R.red_host, (size_t)ne * 2 * 5 * sizeof(double), hipHostMallocDefault));
        R.red_cap = ne * 2;
    }
    if (ne > R.cnt_cap) {
        if (R.cnt_host) hipHostFree(R.cnt_host);
        R.cnt_host = nullptr;
        R.cnt_cap = 0;
        SRV_CHECK(S, hipHostMalloc((void**)&R.cnt_host, (size_t)ne * 2 * sizeof(int), hipHostMallocDefault));
        R.cnt_cap = ne * 2;
    }
    if (!R.ev0) SRV_CHECK(S, hipEventCreate(&R.ev0));
    if (!R.ev1) SRV_CHECK(S, hipEventCreate(&R.ev1));
    // share the round's budget, shortest remaining work first
    std::vector<int> order(ne);
    std::vector<double> rem(ne);
    for (int e = 0; e < ne; ++e) {
        order[e] = e;
        rem[e] = remaining_cost(batch[e]);
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return rem[a] < rem[b]; });
    const bool unlimited = !(S->round_ms > 0.0);
    double left = unlimited ? 0.0 : S->round_ms * kUnitsPerMs;
    std::vector<Slice> slices;
    for (int i = 0; i < ne; ++i) {
        Slice sl = plan_slice(batch[order[i]], left / (ne - i), unlimited);
        left -= sl.cost;
        slices.push_back(sl);
    }

    Tables T;
    std::vector<SelectJob> sel[2];
    std::vector<int> sel_start[2] = {{0}, {0}};
    std::vector<SweepGroup> sweeps, qsweeps;
    std::vector<SolveGroup> solves;
    std::vector<QdGroup> qds;
    std::vector<ReduceJob> reds;
    std::vector<Request*> splits, qstarts, finished, checks, refine_starts;
    for (const Slice& sl : slices) {
        Request* r = sl.r;
        const Eval& ev = r->ev;
        gwm_hip_ctx* c = r->ctx;
        const int k = r->k, nta = nta_of(k), ntb = ntb_of(k);
        const int prec = ev.precision;
        const double* bw_pt = r->adaptive ? c->bwi : nullptr;
        int mode, smode, PT, colB = 0;
        if (ev.type == kEvalCrit) {
            mode = ev.crit == GWM_CRIT_CV ? kModeCV : kModePlain;
            smode = ev.crit == GWM_CRIT_CV ? kSolveCV : kSolveAIC;
            PT = nta * 16;
        } else {
            mode = kModeHat;
            smode = kSolveFit;
            PT = (nta + ntb) * 16;
            colB = nta * 16;
        }
        const int tb = mode == kModeHat ? ntb : 0;
        if (sl.a1 > sl.a0) {
            if (r->adaptive) {
                const int d = r->dist;
                sel[d].push_back(SelectJob{c->points(), ev.bw, c->bwi, sl.a0, c->prune ? c->boxes : nullptr});
                sel_start[d].push_back(sel_start[d].back() + sl.a1 - sl.a0);
            }
            SweepGroup& g = group_of(S, sweeps, r->dist, mode, prec, nta, tb);
            g.jobs.push_back(sweep_job(r, mode, nta, tb, bw_pt, ev.bw, sl.a0));
            add_items(g, (int)g.jobs.size() - 1, sl.a0, sl.a1, r->nseg, g.R);
            SolveGroup& sg = sgroup_of(solves, k, smode);
            if (sg.start.empty()) sg.start.push_back(0);
            SolveArgs sa = solve_job(r, PT, colB, bw_pt, sl.a0);
            // mixed CV: focus points far from every other point go to the FP64 list
            if (prec == GWM_MIXED && mode == kModeCV) sa.dnn = c->dnn;
            sg.jobs.push_back(sa);
            sg.start.push_back(sg.start.back() + sl.a1 - sl.a0);
        }
        if (sl.e1 > sl.e0) {
            // FP64 recomputation of listed focus points: FP64 kernel, the
            // request's own segments, compact partial-sum columns in part
            SweepGroup& g = group_of(S, sweeps, r->dist, mode, GWM_FP64, nta, tb);
            SweepJob j = sweep_job(r, mode, nta, tb, bw_pt, ev.bw, sl.e0);
            j.Z = c->Z;
            j.focus_idx = c->rlist;
            j.nfocus = r->nrefine;
            g.jobs.push_back(j);
            add_items(g, (int)g.jobs.size() - 1, sl.e0, sl.e1, r->nseg, g.R);
            SolveGroup& sg = sgroup_of(solves, k, smode);
            if (sg.start.empty()) sg.start.push_back(0);
            SolveArgs sa = solve_job(r, PT, colB, bw_pt, sl.e0);
            sa.focus_idx = c->rlist;
            sa.nfocus = r->nrefine;
            sg.jobs.push_back(sa);
            sg.start.push_back(sg.start.back() + sl.e1 - sl.e0);
            if (sl.e0 == 0) refine_starts.push_back(r);
        }
        if (sl.check) checks.push_back(r);
        if (sl.split && prec == GWM_MIXED) splits.push_back(r);
        if (sl.split && r->adaptive && c->prune) qstarts.push_back(r);
        if (sl.q1 > sl.q0) {
            SweepGroup& q = group_of(S, qsweeps, r->dist, kModeQDiag, prec, ntb, 0);
            q.jobs.push_back(sweep_job(r, kModeQDiag, ntb, 0, bw_pt, ev.bw, sl.q0));
            add_items(q, (int)q.jobs.size() - 1, sl.q0, sl.q1, r->nseg, q.R);
            QdGroup& qg = kgroup_of(qds, k);
            if (qg.start.empty()) qg.start.push_back(0);
            QDiagJob qj{r->part, r->nseg, ntb * 16, c->n, r->s_cap, c->X, c->sii, c->qd};
            qj.f0 = sl.q0;
            qj.part_off = sl.q0;
            qg.jobs.push_back(qj);
            qg.start.push_back(qg.start.back() + sl.q1 - sl.q0);
        }
        if (sl.done) {
            ReduceJob rj;
            rj.n = c->n;
            rj.out = R.red + 5 * finished.size();
            rj.fail = r->dfail;
            if (ev.type == kEvalCrit) {
                rj.a = c->res2;
                rj.b = ev.crit == GWM_CRIT_CV ? nullptr : c->sii;
                rj.c = c->ssi;   // (res yhat)^2 of the error bound
                rj.m = c->cond;
            } else {
                rj.a = c->sii;
                rj.b = c->ssi;
                rj.c = c->res2;
            }
            reds.push_back(rj);
            finished.push_back(r);
        }
    }
    // pack every table, upload once
    size_t off_sel[2] = {0, 0}, off_sels[2] = {0, 0};
    for (int d = 0; d < 2; ++d)
        if (!sel[d].empty()) {
            off_sel[d] = T.push(sel[d].data(), sel[d].size() * sizeof(SelectJob));
            off_sels[d] = T.push(sel_start[d].data(), sel_start[d].size() * sizeof(int));
        }
    for (auto* gv : {&sweeps, &qsweeps})
        for (auto& g : *gv) {
            g.off_jobs = T.push(g.jobs.data(), g.jobs.size() * sizeof(SweepJob));
            g.off_items = T.push(g.items.data(), g.items.size() * sizeof(int));
            if (g.timed) {
                const unsigned long long zero = 0ull;
                g.off_pairs = T.push(&zero, sizeof(zero));
            }
        }
    for (auto& g : solves) {
        g.off_jobs = T.push(g.jobs.data(), g.jobs.size() * sizeof(SolveArgs));
        g.off_start = T.push(g.start.data(), g.start.size() * sizeof(int));
    }
    for (auto& g : qds) {
        g.off_jobs = T.push(g.jobs.data(), g.jobs.size() * sizeof(QDiagJob));
        g.off_start = T.push(g.start.data(), g.start.size() * sizeof(int));
    }
    size_t off_red = reds.empty() ? 0 : T.push(reds.data(), reds.size() * sizeof(ReduceJob));
    // checks: one compaction launch for every checked request, counts in the
    // arena (zero from the upload), read back with one copy
    std::vector<CompactJob> cjobs;
    int cmax = 0;
    for (Request* r : checks) {
        const double limit = r->ev.type == kEvalCrit ? r->refine_crit : r->refine_fit;
        cjobs.push_back(CompactJob{r->ctx->cond, r->n, limit, r->ctx->rlist, nullptr});
        cmax = std::max(cmax, r->n);
    }
    const std::vector<int> czero(checks.size(), 0);
    const size_t off_cnt = checks.empty() ? 0 : T.push(czero.data(), czero.size() * sizeof(int));
    const size_t off_cj = checks.empty() ? 0 : T.push(cjobs.data(), cjobs.size() * sizeof(CompactJob));
    // failure flags of the FP64 recomputations that start this round
    std::vector<int*> zeros;
    for (Request* r : refine_starts) zeros.push_back(r->dfail);
    const size_t off_zero = zeros.empty() ? 0 : T.push(zeros.data(), zeros.size() * sizeof(int*));
    if (T.host.size() > R.arena.cap) {
        if (R.arena.base) hipFree(R.arena.base);
        if (R.host) hipHostFree(R.host);
        R.arena.base = nullptr;
        R.host = nullptr;
        R.arena.cap = 0;
        R.host_cap = 0;
        size_t cap = std::max<size_t>(T.host.size() * 2, 1 << 20);
        SRV_CHECK(S, hipMalloc((void**)&R.arena.base, cap));
        SRV_CHECK(S, hipHostMalloc((void**)&R.host, cap, hipHostMallocDefault));
        R.arena.cap = cap;
        R.host_cap = cap;
    }
    const char* dev = R.arena.base;
    // jobs of timed launches count their swept pairs (the tuner's work measure)
    for (auto* gv : {&sweeps, &qsweeps})
        for (auto& g : *gv)
            if (g.timed)
                for (size_t i = 0; i < g.jobs.size(); ++i) {
                    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(R.arena.base + g.off_pairs);
                    std::memcpy(T.host.data() + g.off_jobs + i * sizeof(SweepJob) + offsetof(SweepJob, pairs), &cnt,
                                sizeof(cnt));
                }
    for (size_t i = 0; i < cjobs.size(); ++i) {
        CompactJob j = cjobs[i];
        j.count = reinterpret_cast<int*>(R.arena.base + off_cnt) + i;
        std::memcpy(T.host.data() + off_cj + i * sizeof(CompactJob), &j, sizeof(CompactJob));
    }
    SRV_CHECK(S, hipEventRecord(R.ev0, S->stream));
    if (!T.host.empty()) {
        std::memcpy(R.host, T.host.data(), T.host.size());
        SRV_CHECK(S, hipMemcpyAsync(R.arena.base, R.host, T.host.size(), hipMemcpyHostToDevice, S->stream));
    }
    // the failure flag of an FP64 recomputation is that of its own solves
    // (every point the mixed solves failed on is listed)
    SRV_CHECK(S, launch_zero_ints(reinterpret_cast<int* const*>(dev + off_zero), (int)zeros.size(), S->stream));
    // 1. adaptive bandwidths of the main-sweep focus points
    const int sel_need = std::max(sel_start[0].back(), sel_start[1].back());
    if (sel_need > S->sel_cap) {
        if (S->sel_left) hipFree(S->sel_left);
        S->sel_left = nullptr;
        S->sel_cap = 0;
        SRV_CHECK(S, hipMalloc((void**)&S->sel_left, (size_t)sel_need * 2 * sizeof(int)));
        S->sel_cap = sel_need * 2;
    }
    if (!S->sel_nleft) SRV_CHECK(S, hipMalloc((void**)&S->sel_nleft, 2 * sizeof(unsigned int)));
    for (int d = 0; d < 2; ++d)
        if (!sel[d].empty())
            SRV_CHECK(S, launch_adaptive_bw_batch(d, reinterpret_cast<const SelectJob*>(dev + off_sel[d]),
                                                  reinterpret_cast<const int*>(dev + off_sels[d]), (int)sel[d].size(),
                                                  sel_start[d].back(), S->sel_left, S->sel_nleft + d, S->stream));
    // 2. criterion and hat sweeps
    R.ntimed = 0;
    int rc = launch_sweeps(S, sweeps, dev, R);
    if (rc) return rc;
    // 3. local solves
    for (auto& g : solves)
        SRV_CHECK(S, launch_solve_batch(g.k, g.mode, reinterpret_cast<const SolveArgs*>(dev + g.off_jobs),
                                        reinterpret_cast<const int*>(dev + g.off_start), (int)g.jobs.size(),
                                        g.start.back(), S->stream));
    // 3b. mixed evaluations whose main sweep completed: list the focus points
    //     with an ill-conditioned local design
    if (!checks.empty())
        SRV_CHECK(S, launch_compact_batch(reinterpret_cast<const CompactJob*>(dev + off_cj), (int)checks.size(), cmax,
                                          S->stream));
    // 4. qDiag second sweeps of the fits (G complete: every solve of the fit precedes on the stream)
    for (Request* r : qstarts)
        SRV_CHECK(S, launch_chunk_max(r->ctx->bwi, r->n, r->ctx->chunk_bw, S->stream));
    for (Request* r : splits)
        SRV_CHECK(S, launch_split_fp16(r->ctx->G, r->ctx->rows_pad, ntb_of(r->k) * 16, r->ctx->colmax, r->ctx->Gh,
                                       r->ctx->Gl, r->ctx->gscale, S->stream));
    if ((rc = launch_sweeps(S, qsweeps, dev, R))) return rc;
    for (auto& g : qds)
        SRV_CHECK(S, launch_qdiag_batch(g.k, reinterpret_cast<const QDiagJob*>(dev + g.off_jobs),
                                        reinterpret_cast<const int*>(dev + g.off_start), (int)g.jobs.size(),
                                        g.start.back(), S->stream));
    // 5. reductions of the completed evaluations
    if (!finished.empty())
        SRV_CHECK(S, launch_reduce_batch(reinterpret_cast<const ReduceJob*>(dev + off_red), (int)finished.size(),
                                         S->stream));
    SRV_CHECK(S, hipEventRecord(R.ev1, S->stream));
    R.batch = batch;
    R.checks = checks;
    R.finished = finished;
    R.off_cnt = off_cnt;
    R.busy = true;
    return GWM_OK;
}

// Waits for slot R's round, reads its results back and advances the phases of
// its evaluations.
int finish_round(gwm_hip_server* S, RoundSlot& R) {
    R.busy = false;
    SRV_CHECK(S, hipEventSynchronize(R.ev1));
    const std::vector<Request*>& checks = R.checks;
    const std::vector<Request*>& finished = R.finished;
    const int nd = (int)finished.size();
    if (!checks.empty())
        SRV_CHECK(S, hipMemcpyAsync(R.cnt_host, R.arena.base + R.off_cnt, checks.size() * sizeof(int),
                                    hipMemcpyDeviceToHost, S->rb_stream));
    if (nd)
        SRV_CHECK(S, hipMemcpyAsync(R.red_host, R.red, (size_t)nd * 5 * sizeof(double), hipMemcpyDeviceToHost,
                                    S->rb_stream));
    for (Request* r : finished) {
        if (r->ev.type != kEvalFit) continue;
        const size_t nk = (size_t)r->n * r->k;
        r->hbetas.resize(nk);
        r->hse.resize(nk);
        r->hq.resize(r->n);
        SRV_CHECK(S, hipMemcpyAsync(r->hbetas.data(), r->ctx->betas, nk * 8, hipMemcpyDeviceToHost, S->rb_stream));
        SRV_CHECK(S, hipMemcpyAsync(r->hse.data(), r->ctx->se, nk * 8, hipMemcpyDeviceToHost, S->rb_stream));
        SRV_CHECK(S, hipMemcpyAsync(r->hq.data(), r->ctx->qd, (size_t)r->n * 8, hipMemcpyDeviceToHost, S->rb_stream));
    }
    for (int i = 0; i < R.ntimed; ++i)
        SRV_CHECK(S, hipMemcpyAsync(R.timed[i].pairs_host, R.arena.base + R.timed[i].off_pairs,
                                    sizeof(unsigned long long), hipMemcpyDeviceToHost, S->rb_stream));
    SRV_CHECK(S, hipStreamSynchronize(S->rb_stream));
    const double* red = R.red_host;
    for (size_t i = 0; i < checks.size(); ++i) checks[i]->nrefine = R.cnt_host[i];
    // checked evaluations: recompute the listed points next, or go on
    for (Request* r : checks) {
        if (r->nrefine > 0) {
            r->phase = kRefine;
            r->next = 0;
        } else if (r->ev.type == kEvalCrit) {
            r->phase = kDone;
        } else {
            r->phase = kQDiag;
            r->next = 0;
        }
    }
    long long nsolves = 0;
    for (int i = 0; i < nd; ++i) {
        Request* r = finished[i];
        if (r->phase != kDone) continue;   // a criterion whose check listed points
        Eval& ev = r->ev;
        nsolves += r->n;
        r->solves += r->n;
        const double* o = &red[5 * i];
        if (o[3] != 0.0) {
            ev.rc = GWM_ERR_NOT_SPD;
            ev.value = DBL_MAX;
            continue;
        }
        if (ev.type == kEvalCrit) {
            double v = ev.crit == GWM_CRIT_CV ? o[0] : aicc(o[0], (double)r->n, o[1]);
            ev.rc = std::isfinite(v) ? GWM_OK : GWM_ERR_NOT_SPD;
            ev.value = std::isfinite(v) ? v : DBL_MAX;
            ev.cond = o[4];
            ev.bound = ev.precision == GWM_MIXED && std::isfinite(v)
                           ? crit_error_bound(ev.crit, v, o[0], o[1], o[2], r->n, o[4], r->refine_crit,
                                              r->ctx->certify_rel)
                           : 0.0;
        } else {
            ev.rc = GWM_OK;
            ev.shat[0] = o[0];
            ev.shat[1] = o[1];
        }
    }
    for (int i = 0; i < R.ntimed; ++i) {
        const TimedGroup& tg = R.timed[i];
        float gms = 0.0f;
        SRV_CHECK(S, hipEventElapsedTime(&gms, tg.a, tg.b));
        std::lock_guard<std::mutex> lk(S->tune_mu);
        SrvTune& t = S->tune[tg.key];
        const double units = (double)*tg.pairs_host;
        if (t.chosen >= 0 || !(units > 0.0)) continue;
        t.ms[tg.R] += gms;
        t.units[tg.R] += units;
        t.samples[tg.R] += 1;
        if (t.samples[1] >= kTuneRounds && t.samples[2] >= kTuneRounds)
            t.chosen = t.ms[1] / t.units[1] < t.ms[2] / t.units[2] ? 1 : 2;
    }
    R.ntimed = 0;
    float ms = 0.0f;
    SRV_CHECK(S, hipEventElapsedTime(&ms, R.ev0, R.ev1));
    S->rounds += 1;
    S->solves += nsolves;
    S->busy_ms += ms;
    return GWM_OK;
}

int setup(gwm_hip_server* S, Request* R);
void teardown(Request* R);

// Worker: the host algorithm of one request.
constexpr int kSetupSlots = 8;   // concurrent request set-ups

void worker_main(gwm_hip_server* S, Request* R) {
    {
        const double cost = (double)R->n * (double)R->n * (double)R->k;
        {
            std::unique_lock<std::mutex> lk(S->setup_mu);
            S->setup_waiting.emplace_back(cost, R);
            auto first = [&]() {
                auto it = std::max_element(S->setup_waiting.begin(), S->setup_waiting.end(),
                                           [](const std::pair<double, Request*>& a,
                                              const std::pair<double, Request*>& b) { return a.first < b.first; });
                return it->second == R;
            };
            S->setup_cv.wait(lk, [&] { return S->setup_running < kSetupSlots && first(); });
            S->setup_waiting.erase(std::find_if(S->setup_waiting.begin(), S->setup_waiting.end(),
                                                [&](const std::pair<double, Request*>& e) { return e.second == R; }));
            ++S->setup_running;
            S->setup_cv.notify_all();
        }
        const int rc = setup(S, R);
        {
            std::lock_guard<std::mutex> lk(S->setup_mu);
            --S->setup_running;
            S->setup_cv.notify_all();
        }
        if (rc) {
            teardown(R);
            std::unique_lock<std::mutex> lk(S->mu);
            R->status = rc;
            R->finished = true;
            S->cv.notify_all();
            return;
        }
    }
    auto post = [&](int type, double bw, int prec, int crit, double* value, double* bound) -> int {
        std::unique_lock<std::mutex> lk(S->mu);
        R->ev = Eval();
        R->ev.type = type;
        R->ev.bw = bw;
        R->ev.precision = prec;
        R->ev.crit = crit;
        R->phase = kMain;
        R->next = 0;
        R->nrefine = 0;
        R->posted = true;
        S->cv.notify_all();
        R->wcv.wait(lk, [&] { return !R->posted; });
        if (value) *value = R->ev.value;
        if (bound) *bound = R->ev.bound;
        return R->ev.rc;
    };
    int rc = GWM_OK;
    double bw = R->bw;
    double certify_rel;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        certify_rel = S->certify_rel;
    }
    if (R->select_bw) {
        double lo = R->has_lower ? R->lower : (R->adaptive ? 20 : 0.0);
        double hi = R->has_upper ? R->upper : (R->adaptive ? (double)R->n : R->maxdist);
        std::map<double, double> log;
        rc = golden(R->adaptive != 0, lo, hi,
                    [&](double x, bool exact, double* fx, double* fb) {
                        return post(kEvalCrit, x, exact ? GWM_FP64 : S->precision, R->crit, fx, fb);
                    },
                    S->precision == GWM_MIXED, certify_rel, &bw, log);
    }
    if (rc == GWM_OK) rc = post(kEvalFit, bw, S->precision, 0, nullptr, nullptr);
    if (rc == GWM_OK) {
        // device order -> caller's order
        {
            std::vector<double> t(R->hbetas.size());
            to_caller(R->ctx, R->hbetas.data(), t.data(), R->k);
            R->hbetas.swap(t);
            to_caller(R->ctx, R->hse.data(), t.data(), R->k);
            R->hse.swap(t);
            t.resize(R->n);
            to_caller(R->ctx, R->hq.data(), t.data(), 1);
            R->hq.swap(t);
        }
        double shat[2] = {R->ev.shat[0], R->ev.shat[1]};
        double diag[7];
        calc_diagnostic(R->n, R->k, R->X.data(), R->y.data(), R->hbetas.data(), shat, diag);
        const double sigma = diag[0] / (R->n - 2 * shat[0] + shat[1]);
        const size_t nk = (size_t)R->n * R->k;
        if (R->o_betas) std::memcpy(R->o_betas, R->hbetas.data(), nk * 8);
        if (R->o_se)
            for (size_t i = 0; i < nk; ++i) R->o_se[i] = std::sqrt(sigma * R->hse[i]);
        if (R->o_shat) std::memcpy(R->o_shat, shat, sizeof(shat));
        if (R->o_qdiag) std::memcpy(R->o_qdiag, R->hq.data(), (size_t)R->n * 8);
        if (R->o_diag) std::memcpy(R->o_diag, diag, sizeof(diag));
        if (R->o_bw) *R->o_bw = bw;
    }
    teardown(R);
    std::unique_lock<std::mutex> lk(S->mu);
    R->status = rc;
    R->finished = true;
    S->cv.notify_all();
}

void release(gwm_hip_server* S, Request* R) {
    teardown(R);
    S->inflight_bytes -= R->footprint;
}

// Device state of an admitted request: built by its worker thread (so the
// set-ups of many requests run on the host's cores while the scheduler runs
// rounds), from the stream-ordered allocator of the request's own stream, and
// synchronised before the first evaluation is posted.
int setup(gwm_hip_server* S, Request* R) {
    int rc = gwm_hip_create(S->device, &R->ctx);
    if (rc) return rc;
    R->ctx->async_mem = true;
    rc = gwm_hip_set_data(R->ctx, R->n, R->k, R->coords.data(), R->X.data(), R->y.data(), R->dist);
    if (rc) return rc;
    R->ctx->certify_rel = S->certify_rel;
    const size_t bytes = (size_t)R->nseg * (nta_of(R->k) + ntb_of(R->k)) * 16 * R->s_cap * sizeof(double);
    if (hipMallocAsync((void**)&R->part, bytes + 256, R->ctx->stream) != hipSuccess) return GWM_ERR_OOM;
    R->dfail = reinterpret_cast<int*>(reinterpret_cast<char*>(R->part) + bytes);
    if (hipMemsetAsync(R->dfail, 0, sizeof(int), R->ctx->stream) != hipSuccess) return GWM_ERR_HIP;
    if (R->select_bw && !R->adaptive && !R->has_upper) {
        if ((rc = max_distance_impl(R->ctx, &R->maxdist))) return rc;
    }
    if (R->select_bw && R->crit == GWM_CRIT_CV && (rc = ensure_dnn(R->ctx))) return rc;
    if (hipStreamSynchronize(R->ctx->stream) != hipSuccess) return GWM_ERR_HIP;
    return GWM_OK;
}

// Device state of a request whose work has completed: returned to the
// stream-ordered allocator, without a device-wide synchronisation.
void teardown(Request* R) {
    if (R->part) hipFreeAsync(R->part, R->ctx->stream);
    R->part = nullptr;
    R->dfail = nullptr;
    if (R->ctx) gwm_hip_destroy(R->ctx);
    R->ctx = nullptr;
}

void scheduler_main(gwm_hip_server* S) {
    hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        // retire finished requests
        for (size_t i = 0; i < S->active.size();) {
            Request* R = S->active[i];
            if (!R->finished) {
                ++i;
                continue;
            }
            lk.unlock();
            if (R->worker.joinable()) R->worker.join();
            lk.lock();
            release(S, R);
            R->t_done = now_ms();
            R->done = true;
            S->active.erase(S->active.begin() + i);
            S->cv.notify_all();
        }
        // admit waiting requests while slots and memory allow (first come, first served)
        while (!S->pending.empty() && (int)S->active.size() < S->max_inflight) {
            Request* R = S->pending.front();
            if (!S->active.empty() && S->inflight_bytes + R->footprint > S->budget) break;
            S->pending.pop_front();
            R->t_admit = now_ms();
            S->inflight_bytes += R->footprint;
            R->admitted = true;
            S->active.push_back(R);
            S->peak_bytes = std::max(S->peak_bytes, S->inflight_bytes);
            S->peak_inflight = std::max<int>(S->peak_inflight, (int)S->active.size());
            R->worker = std::thread(worker_main, S, R);
        }
        RoundSlot& prev = S->slots[S->slot_next ^ 1];   // the round in flight, if any
        if (S->active.empty() && !prev.busy) {
            if (S->stopping && S->pending.empty()) break;
            S->cv.wait(lk, [&] { return S->stopping || !S->pending.empty(); });
            continue;
        }
        // Wait until every in-flight request outside the round in flight has
        // posted an evaluation or finished, for at most kGatherMs (a worker
        // still setting up or finishing on the host joins a later round), or
        // kGatherBusyMs while the device works on a round in flight.
        const double gather = prev.busy ? kGatherBusyMs : kGatherMs;
        S->cv.wait_for(lk, std::chrono::microseconds((long long)(gather * 1000)), [&] {
            for (Request* R : S->active)
                if (!R->in_round && !R->posted && !R->finished) return false;
            return true;
        });
        std::vector<Request*> batch;
        for (Request* R : S->active)
            if (R->posted && !R->in_round) batch.push_back(R);
        if (batch.empty() && !prev.busy) {
            S->cv.wait(lk, [&] {
                for (Request* R : S->active)
                    if (R->posted || R->finished) return true;
                return S->active.empty();
            });
            continue;
        }
        RoundSlot& cur = S->slots[S->slot_next];
        for (Request* R : batch) R->in_round = true;
        lk.unlock();
        int rc = batch.empty() ? GWM_OK : launch_round(S, batch, cur);
        int rc_prev = prev.busy ? finish_round(S, prev) : GWM_OK;
        lk.lock();
        // requests of the finished round: completed evaluations wake their
        // workers, the others post their next slice
        auto settle = [&](const std::vector<Request*>& rs, int code) {
            for (Request* R : rs) {
                R->in_round = false;
                if (code) {
                    R->ev.rc = code;
                    R->ev.value = DBL_MAX;
                } else if (R->phase != kDone) {
                    continue;   // in progress: more slices in a later round
                }
                R->posted = false;
                R->wcv.notify_one();   // wake only this request's worker
            }
        };
        if (!prev.batch.empty()) {
            settle(prev.batch, rc_prev);
            prev.batch.clear();
        }
        if (rc) {
            // a launch that failed part-way: wait for what it queued, then fail its requests
            hipStreamSynchronize(S->stream);
            cur.busy = false;
            settle(batch, rc);
            cur.batch.clear();
        }
        if (!batch.empty()) S->slot_next ^= 1;
    }
}

}  // namespace

extern "C" {

int gwm_hip_server_create(int device, int precision, double budget_bytes, int max_inflight, gwm_hip_server** out) {
    if (!out || (precision != GWM_FP64 && precision != GWM_MIXED) || max_inflight < 1) return GWM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return GWM_ERR_NO_DEVICE;
    gwm_hip_server* S = new gwm_hip_server();
    S->device = device;
    S->precision = precision;
    S->max_inflight = max_inflight;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&S->rb_stream, hipStreamNonBlocking) != hipSuccess) {
        if (S->stream) hipStreamDestroy(S->stream);
        delete S;
        return GWM_ERR_HIP;
    }
    size_t free_b = 0, total_b = 0;
    hipMemGetInfo(&free_b, &total_b);
    S->budget = budget_bytes > 0 ? (size_t)budget_bytes : (size_t)(0.8 * (double)free_b);
    S->sched = std::thread(scheduler_main, S);
    *out = S;
    return GWM_OK;
}

int gwm_hip_server_set_certify_tolerance(gwm_hip_server* S, double rel) {
    if (!S || !(rel >= 0.0)) return GWM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->certify_rel = rel;
    return GWM_OK;
}

int gwm_hip_server_set_refine_condition(gwm_hip_server* S, double criterion_limit, double fit_limit) {
    if (!S || std::isnan(criterion_limit) || std::isnan(fit_limit)) return GWM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->refine_crit = criterion_limit;
    S->refine_fit = fit_limit;
    return GWM_OK;
}

int gwm_hip_server_set_round_ms(gwm_hip_server* S, double ms) {
    if (!S || std::isnan(ms)) return GWM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->round_ms = ms;
    return GWM_OK;
}

int gwm_hip_server_autotune_report(gwm_hip_server* S, char* buf, int cap) {
    if (!S || !buf || cap <= 0) return GWM_ERR_INVALID;
    static const char* modes[] = {"cv", "hat", "qdiag", "plain"};   // ContractMode order
    std::string out;
    {
        std::lock_guard<std::mutex> lk(S->tune_mu);
        for (const auto& kv : S->tune) {
            const TuneClass& k = kv.first;
            const SrvTune& t = kv.second;
            if (t.samples[1] + t.samples[2] == 0) continue;
            char line[256];
            std::snprintf(line, sizeof(line), "%s %s %s nta=%d ntb=%d chosen_r=%d us_per_Mpair=%.4f,%.4f rounds=%d,%d\n",
                          k.dist == GWM_GREAT_CIRCLE ? "great_circle" : "euclidean",
                          k.mode >= 0 && k.mode < 4 ? modes[k.mode] : "?", k.prec == GWM_MIXED ? "mixed" : "fp64",
                          k.nta, k.ntb, t.chosen,
                          t.units[1] > 0 ? 1e9 * t.ms[1] / t.units[1] : -1.0,
                          t.units[2] > 0 ? 1e9 * t.ms[2] / t.units[2] : -1.0, t.samples[1], t.samples[2]);
            out += line;
        }
    }
    std::snprintf(buf, cap, "%s", out.c_str());
    return GWM_OK;
}

int gwm_hip_server_set_partial_budget(gwm_hip_server* S, double bytes) {
    if (!S || !(bytes >= 1.0)) return GWM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->part_budget = bytes >= 1.8e19 ? SIZE_MAX : (size_t)bytes;
    return GWM_OK;
}

int gwm_hip_server_submit(gwm_hip_server* S, int n, int k, const double* coords, const double* X, const double* y,
                          int distance, int kernel, int adaptive, double bw, int select_bw, int criterion,
                          int has_lower, double lower, int has_upper, double upper, double* betas, double* betasSE,
                          double* shat, double* qDiag, double* diag, double* bw_out, long long* request_id) {
    if (!S || n < 2 || k < 1 || k > 16 || !coords || !X || !y || kernel < 0 || kernel > 4 ||
        (distance != GWM_EUCLIDEAN && distance != GWM_GREAT_CIRCLE))
        return GWM_ERR_INVALID;
    if (!select_bw && adaptive && !(bw >= 1.0)) return GWM_ERR_INVALID;
    Request* R = new Request();
    R->n = n;
    R->k = k;
    R->dist = distance;
    R->kernel = kernel;
    R->adaptive = adaptive;
    R->bw = bw;
    R->select_bw = select_bw;
    R->crit = criterion;
    R->has_lower = has_lower;
    R->lower = lower;
    R->has_upper = has_upper;
    R->upper = upper;
    R->coords.assign(coords, coords + 2 * (size_t)n);
    R->X.assign(X, X + (size_t)n * k);
    R->y.assign(y, y + n);
    R->o_betas = betas;
    R->o_se = betasSE;
    R->o_shat = shat;
    R->o_qdiag = qDiag;
    R->o_diag = diag;
    R->o_bw = bw_out;
    data_segments(n, R->nseg, R->seg_len);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        R->s_cap = slice_cap(n, k, R->nseg, S->part_budget);
    }
    R->footprint = footprint_bytes(n, k, distance, R->nseg, R->s_cap);
    std::lock_guard<std::mutex> lk(S->mu);
    R->refine_crit = S->refine_crit;
    R->refine_fit = S->refine_fit;
    R->id = S->next_id++;
    R->t_submit = now_ms();
    S->reqs[R->id] = R;
    S->pending.push_back(R);
    S->cv.notify_all();
    if (request_id) *request_id = R->id;
    return GWM_OK;
}

int gwm_hip_server_wait(gwm_hip_server* S, long long id, int* status, double* latency_ms, long long* solves) {
    if (!S) return GWM_ERR_INVALID;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return GWM_ERR_INVALID;
    Request* R = it->second;
    S->cv.wait(lk, [&] { return R->done; });
    if (status) *status = R->status;
    if (latency_ms) *latency_ms = R->t_done - R->t_submit;
    if (solves) *solves = R->solves;
    S->reqs.erase(it);
    delete R;
    return GWM_OK;
}

int gwm_hip_server_stats(gwm_hip_server* S, long long* rounds, long long* solves, double* busy_ms,
                         double* peak_bytes, int* peak_inflight) {
    if (!S) return GWM_ERR_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    if (rounds) *rounds = S->rounds;
    if (solves) *solves = S->solves;
    if (busy_ms) *busy_ms = S->busy_ms;
    if (peak_bytes) *peak_bytes = (double)S->peak_bytes;
    if (peak_inflight) *peak_inflight = S->peak_inflight;
    return GWM_OK;
}

const char* gwm_hip_server_last_error(gwm_hip_server* S) { return S ? S->err.c_str() : "null server"; }

void gwm_hip_server_destroy(gwm_hip_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
        S->cv.notify_all();
    }
    if (S->sched.joinable()) S->sched.join();
    for (auto& kv : S->reqs) delete kv.second;
    for (RoundSlot& R : S->slots) {
        if (R.arena.base) hipFree(R.arena.base);
        if (R.host) hipHostFree(R.host);
        if (R.red) hipFree(R.red);
        if (R.red_host) hipHostFree(R.red_host);
        if (R.cnt_host) hipHostFree(R.cnt_host);
        if (R.ev0) hipEventDestroy(R.ev0);
        if (R.ev1) hipEventDestroy(R.ev1);
        for (TimedGroup& tg : R.timed) {
            if (tg.a) hipEventDestroy(tg.a);
            if (tg.b) hipEventDestroy(tg.b);
            if (tg.pairs_host) hipHostFree(tg.pairs_host);
        }
    }
    if (S->rb_stream) hipStreamDestroy(S->rb_stream);
    if (S->sel_left) hipFree(S->sel_left);
    if (S->sel_nleft) hipFree(S->sel_nleft);
    if (S->stream) hipStreamDestroy(S->stream);
    delete S;
}

}  // extern "C"
