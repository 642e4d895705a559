// C++ host interface (include/gwmodelpp/gwmodel.h) over the C ABI of
// include/gwmodel_hip.h.  Behaviour follows the reference's GWRBasic::fit
// (src/gwmodelpp/GWRBasic.cpp:36-128) and GWRMultiscale::fit
// (src/gwmodelpp/GWRMultiscale.cpp:68-197): bandwidth selection bounds,
// criterion list, diagnostics, SE scaling, Status::Terminated with zero
// matrices on cancellation, exceptions on non-SPD local designs and on
// invalid golden-section starting values.
#include "gwmodelpp/gwmodel.h"

#include <cfloat>
#include <string>

#include "gwmodel_hip.h"

namespace gwm {

namespace {

int stop_adapter(void* user) { return static_cast<ITelegram*>(user)->stop() ? 1 : 0; }

// The library's messages and progress reports go to the algorithm's telegram
// as GWM_LOG_INFO / GWM_LOG_ERROR / GWM_LOG_PROGRESS(_PERCENT) would send them
// (include/gwmodelpp/Algorithm.h:21-100); the function and file names are the
// facade's.
void print_adapter(void* user, int level, const char* message) {
    static_cast<ITelegram*>(user)->print(message, static_cast<ITelegram::LogLevel>(level), "fit", __FILE__);
}

void progress_adapter(void* user, double current, double total) {
    ITelegram* t = static_cast<ITelegram*>(user);
    if (total > 0.0) t->progress(static_cast<std::size_t>(current), static_cast<std::size_t>(total), "fit", __FILE__);
    else t->progress(current, "fit", __FILE__);
}

void attach(gwm_hip_ctx* ctx, ITelegram* telegram) {
    gwm_hip_set_stop_callback(ctx, &stop_adapter, telegram);
    gwm_hip_set_telegram(ctx, &print_adapter, &progress_adapter, telegram);
}

void log_stage(ITelegram* telegram, const char* stage) {
    telegram->print(std::string("#stage ") + stage, ITelegram::LogLevel::LOG_INFO, "fit", __FILE__);
}

[[noreturn]] void raise(gwm_hip_ctx* ctx, int rc) {
    std::string msg = ctx ? gwm_hip_last_error(ctx) : "";
    switch (rc) {
    case GWM_ERR_NOT_SPD: throw std::runtime_error("inv_sympd(): matrix is singular or not positive definite");
    case GWM_ERR_INVALID_INITIAL: throw std::runtime_error("Invalid initial values.");
    case GWM_ERR_INVALID: throw std::invalid_argument(msg);
    case GWM_ERR_OOM: throw std::bad_alloc();
    case GWM_ERR_NO_DEVICE: throw std::runtime_error("no HIP device available");
    default: throw std::runtime_error("HIP back-end error: " + msg);
    }
}

int distance_id(const SpatialWeight& sw) {
    if (!sw.isValid() || sw.distance()->type() != Distance::CRSDistance)
        throw std::invalid_argument("the HIP back-end supports CRSDistance only");
    return sw.distance<CRSDistance>()->geographic() ? GWM_GREAT_CIRCLE : GWM_EUCLIDEAN;
}

}  // namespace

HipContext::HipContext(int device) {
    int rc = gwm_hip_create(device, &mCtx);
    if (rc != GWM_OK) raise(nullptr, rc);
}

HipContext::~HipContext() { gwm_hip_destroy(mCtx); }

Mat GWRBasic::fit() {
    mStatus = Status::Success;
    const std::size_t n = mCoords.n_rows();
    std::size_t k = mX.n_cols();
    if (mParallelType != ParallelType::HIP) throw std::logic_error("parallel type not available in this build");
    if (mCoords.n_cols() != 2 || mX.n_rows() != n || mY.n_elem() != n)
        throw std::invalid_argument("coords must be n x 2, X n x k, y n");
    const int dist = distance_id(mSpatialWeight);
    BandwidthWeight* bw = mSpatialWeight.weight<BandwidthWeight>();
    log_stage(mTelegram, "Initializing");
    HipContext ctx(mGpuId);
    attach(ctx.get(), mTelegram);
    int rc = gwm_hip_set_data(ctx.get(), (int)n, (int)k, mCoords.memptr(), mX.memptr(), mY.memptr(), dist);
    if (rc) raise(ctx.get(), rc);
    const int prec = mPrecision == Precision::Mixed ? GWM_MIXED : GWM_FP64;
    const int crit = mBandwidthSelectionCriterion == CV ? GWM_CRIT_CV : GWM_CRIT_AIC;

    if (mIsAutoselectIndepVars) {
        // GWRBasic.cpp:50-72: forward selection over the columns after the
        // intercept, then the fit runs on {intercept} + the selected columns
        const std::size_t kall = k, cap = k * (k + 1) / 2;
        std::vector<int> sel(k), mv(cap * k), mn(cap);
        std::vector<double> mc(cap);
        int nsel = 0, nmod = 0;
        rc = gwm_hip_indep_vars_select(ctx.get(), mHasIntercept ? 1 : 0, mIndepVarSelectionThreshold, sel.data(),
                                       &nsel, mv.data(), mn.data(), mc.data(), (int)cap, &nmod);
        if (rc == GWM_ERR_CANCELLED) {
            mStatus = Status::Terminated;
            mBetas = Mat(n, k);
            return mBetas;
        }
        if (rc) raise(ctx.get(), rc);
        mSelectedIndepVars.assign(sel.begin(), sel.begin() + nsel);
        if (nsel > 0) {
            std::vector<std::size_t> cols;
            if (mHasIntercept) cols.push_back(0);
            for (std::size_t v : mSelectedIndepVars) cols.push_back(v);
            Mat xs(n, cols.size());
            for (std::size_t a = 0; a < cols.size(); ++a) xs.set_col(a, mX.col(cols[a]));
            mX = xs;
            k = cols.size();
            mIndepVarsSelectionCriterionList.clear();
            for (int q = 0; q < nmod && q < (int)cap; ++q) {
                const std::size_t row = (std::size_t)q * kall;
                std::vector<std::size_t> vars(mv.begin() + row, mv.begin() + row + mn[q]);
                mIndepVarsSelectionCriterionList.emplace_back(vars, mc[q]);
            }
            rc = gwm_hip_set_data(ctx.get(), (int)n, (int)k, mCoords.memptr(), mX.memptr(), mY.memptr(), dist);
            if (rc) raise(ctx.get(), rc);
        }
    }

    mBetas = Mat(n, k);
    mBetasSE = Mat(n, k);
    mSHat = Vec(2, 1);
    mQDiag = Vec(n, 1);
    mS = isStoreS() ? Mat(n, n) : Mat(1, n);
    double diag[7] = {0};
    double bw_out = bw->bandwidth();
    std::vector<double> tx(512), tf(512);
    int tlen = 0;
    rc = gwm_hip_gwr_run(ctx.get(), (int)bw->kernel(), bw->adaptive() ? 1 : 0, bw->bandwidth(),
                         mIsAutoselectBandwidth ? 1 : 0, crit, prec, mGoldenLowerBounds.has_value(),
                         mGoldenLowerBounds.value_or(0.0), mGoldenUpperBounds.has_value(),
                         mGoldenUpperBounds.value_or(0.0), mBetas.memptr(), mBetasSE.memptr(), mSHat.memptr(),
                         mQDiag.memptr(), isStoreS() ? mS.memptr() : nullptr, diag, &bw_out, tx.data(), tf.data(),
                         (int)tx.size(), &tlen);
    if (rc == GWM_ERR_CANCELLED) {
        mStatus = Status::Terminated;
        mBetas = Mat(n, k);
        return mBetas;
    }
    if (rc) raise(ctx.get(), rc);
    if (mIsAutoselectBandwidth) {
        bw->setBandwidth(bw_out);
        mBandwidthSelectionCriterionList.clear();
        for (int i = 0; i < tlen && i < (int)tx.size(); ++i) mBandwidthSelectionCriterionList.emplace_back(tx[i], tf[i]);
    }
    mDiagnostic = RegressionDiagnostic{diag[0], diag[1], diag[2], diag[3], diag[4], diag[5], diag[6]};
    return mBetas;
}

Mat GWRBasic::predict(const Mat& locations) {
    mStatus = Status::Success;
    const std::size_t n = mCoords.n_rows(), k = mX.n_cols(), m = locations.n_rows();
    if (mParallelType != ParallelType::HIP) throw std::logic_error("parallel type not available in this build");
    if (mCoords.n_cols() != 2 || mX.n_rows() != n || mY.n_elem() != n || locations.n_cols() != 2)
        throw std::invalid_argument("coords and locations must have 2 columns, X n x k, y n");
    const int dist = distance_id(mSpatialWeight);
    BandwidthWeight* bw = mSpatialWeight.weight<BandwidthWeight>();
    log_stage(mTelegram, "Initialization");   // GWRBasic.cpp:132-144
    HipContext ctx(mGpuId);
    attach(ctx.get(), mTelegram);
    int rc = gwm_hip_set_data(ctx.get(), (int)n, (int)k, mCoords.memptr(), mX.memptr(), mY.memptr(), dist);
    if (rc) raise(ctx.get(), rc);
    const int prec = mPrecision == Precision::Mixed ? GWM_MIXED : GWM_FP64;
    log_stage(mTelegram, "Prediction");
    Mat betas(m, k);
    rc = gwm_hip_predict(ctx.get(), (int)m, locations.memptr(), (int)bw->kernel(), bw->adaptive() ? 1 : 0,
                         bw->bandwidth(), prec, betas.memptr());
    if (rc == GWM_ERR_CANCELLED) {
        // the reference returns an n x k zero matrix when stopped (GWRBasic.cpp:146)
        mStatus = Status::Terminated;
        return Mat(n, k);
    }
    if (rc) raise(ctx.get(), rc);
    mBetas = betas;
    return mBetas;
}

void GWRMultiscale::setBandwidthInitilize(const std::vector<BandwidthInitilizeType>& v) {
    if (v.size() != mX.n_cols()) throw std::length_error("BandwidthInitilize size do not match indepvars");
    mBandwidthInitilize = v;
}

void GWRMultiscale::setBandwidthSelectionApproach(const std::vector<BandwidthSelectionCriterionType>& v) {
    if (v.size() != mX.n_cols()) throw std::length_error("bandwidthSelectionApproach size do not match indepvars");
    mBandwidthSelectionApproach = v;
}

Mat GWRMultiscale::fit() {
    mStatus = Status::Success;
    const std::size_t n = mCoords.n_rows(), k = mX.n_cols();
    if (mParallelType != ParallelType::HIP) throw std::logic_error("parallel type not available in this build");
    if (mCoords.n_cols() != 2 || mX.n_rows() != n || mY.n_elem() != n)
        throw std::invalid_argument("coords must be n x 2, X n x k, y n");
    if (mSpatialWeights.size() != k || mBandwidthInitilize.size() != k || mBandwidthSelectionApproach.size() != k ||
        mPreditorCentered.size() != k || mBandwidthSelectThreshold.size() != k)
        throw std::invalid_argument("per-variable settings must have one entry per independent variable");
    // one distance per variable; the initial search and fit use variable 0's
    // (GWRMultiscale::setSpatialWeights, GWRMultiscale.cpp:1476-1483)
    const int dist = distance_id(mSpatialWeights[0]);
    std::vector<int> kernels(k), adaptive(k), init(k), approach(k), centered(k), dists(k);
    std::vector<double> bws(k);
    for (std::size_t i = 0; i < k; ++i) {
        dists[i] = distance_id(mSpatialWeights[i]);
        const BandwidthWeight* w = mSpatialWeights[i].weight<BandwidthWeight>();
        kernels[i] = (int)w->kernel();
        adaptive[i] = w->adaptive() ? 1 : 0;
        bws[i] = w->bandwidth();
        init[i] = (int)mBandwidthInitilize[i];
        approach[i] = mBandwidthSelectionApproach[i] == CV ? GWM_CRIT_CV : GWM_CRIT_AIC;
        centered[i] = mPreditorCentered[i] ? 1 : 0;
    }
    log_stage(mTelegram, "Initializing");
    HipContext ctx(mGpuId);
    attach(ctx.get(), mTelegram);
    int rc = gwm_hip_set_data(ctx.get(), (int)n, (int)k, mCoords.memptr(), mX.memptr(), mY.memptr(), dist);
    if (rc) raise(ctx.get(), rc);
    if ((rc = gwm_hip_mgwr_set_distances(ctx.get(), (int)k, dists.data()))) raise(ctx.get(), rc);
    gwm_hip_mgwr_set_hat_matrix(ctx.get(), mHasHatMatrix ? 1 : 0);
    mBetas = Mat(n, k);
    std::vector<double> bws_out(k);
    double diag[7] = {0}, rss = 0.0;
    int iters = 0;
    rc = gwm_hip_mgwr_run(ctx.get(), kernels.data(), adaptive.data(), bws.data(), init.data(), approach.data(),
                          centered.data(), mBandwidthSelectThreshold.data(), mHasIntercept ? 1 : 0,
                          mCriterionType == CVR ? 0 : 1, mCriterionThreshold, (int)mMaxIteration,
                          (int)mBandwidthSelectRetryTimes, mAdaptiveLower, mGoldenLowerBounds.has_value(),
                          mGoldenLowerBounds.value_or(0.0), mGoldenUpperBounds.has_value(),
                          mGoldenUpperBounds.value_or(0.0), mPrecision == Precision::Mixed ? GWM_MIXED : GWM_FP64,
                          mBetas.memptr(), bws_out.data(), diag, &rss, &iters);
    if (rc == GWM_ERR_CANCELLED) {
        mStatus = Status::Terminated;
        mBetas = Mat(n, k);
        return mBetas;
    }
    if (rc) raise(ctx.get(), rc);
    for (std::size_t i = 0; i < k; ++i) mSpatialWeights[i].weight<BandwidthWeight>()->setBandwidth(bws_out[i]);
    mDiagnostic = RegressionDiagnostic{diag[0], diag[1], diag[2], diag[3], diag[4], diag[5], diag[6]};
    return mBetas;
}

}  // namespace gwm
