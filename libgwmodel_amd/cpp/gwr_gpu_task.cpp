// IGWRBasicGpuTask over the HIP back-end: the reference's GWRBasicGpuTask
// (include/gwmodelcuda/GWRBasicGpuTask.h, src/gwmodelcuda/GWRBasicGpuTask.cpp:10-66)
// holds the inputs element-wise, runs GWRBasic::fit / predict with
// hasHatMatrix = true and copies the results out; failures return false.
#include <cmath>
#include <exception>
#include <limits>
#include <vector>

#include "gwmodelpp/IGWRBasicGpuTask.h"
#include "gwmodelpp/gwmodel.h"

namespace {

using gwm::Mat;
using gwm::Vec;

class GWRBasicHipTask final : public IGWRBasicGpuTask {
public:
    GWRBasicHipTask(int nDp, int nVar, int distanceType, int nPredict)
        : mX(nDp, nVar), mY(nDp, 1), mCoords(nDp, 2), mLocations(nPredict, 2), mDistanceType(distanceType) {}
    GWRBasicHipTask(const GWRBasicHipTask&) = delete;
    GWRBasicHipTask& operator=(const GWRBasicHipTask&) = delete;

    void setX(int i, int k, double value) override {
        if (in(i, mX.n_rows()) && in(k, mX.n_cols())) mX(i, k) = value;
    }
    void setY(int i, double value) override {
        if (in(i, mY.n_rows())) mY(i, 0) = value;
    }
    void setCoords(int i, double u, double v) override {
        if (!in(i, mCoords.n_rows())) return;
        mCoords(i, 0) = u;
        mCoords(i, 1) = v;
    }
    void setPredictLocations(int i, double u, double v) override {
        if (!in(i, mLocations.n_rows())) return;
        mLocations(i, 0) = u;
        mLocations(i, 1) = v;
    }

    void setDistanceType(int type) override { mDistanceType = type; }
    void setCRSDistanceGergraphic(bool isGeographic) override { mGeographic = isGeographic; }
    void setMinkwoskiDistancePoly(int poly) override { mPoly = poly; }
    void setMinkwoskiDistanceTheta(double theta) override { mTheta = theta; }

    void setBandwidthSize(double bw) override { mBandwidth = bw; }
    void setBandwidthAdaptive(bool adaptive) override { mAdaptive = adaptive; }
    void setBandwidthKernel(int kernel) override { mKernel = kernel; }

    void enableBandwidthOptimization(int criterion) override {
        mOptimizeBandwidth = true;
        mCriterion = criterion;
    }
    void enableVariablesOptimization(double threshold) override {
        mOptimizeVariables = true;
        mThreshold = threshold;
    }

    double betas(int i, int k) override { return at(mBetas, i, k); }
    double betasSE(int i, int k) override { return at(mBetasSE, i, k); }
    double shat1() override { return at(mSHat, 0, 0); }
    double shat2() override { return at(mSHat, 1, 0); }
    double qDiag(int i) override { return at(mQDiag, i, 0); }
    unsigned long long sRows() override { return mS.n_rows(); }
    double s(int i, int k) override { return at(mS, i, k); }

    double diagnosticRSS() override { return mDiagnostic.RSS; }
    double diagnosticAIC() override { return mDiagnostic.AIC; }
    double diagnosticAICc() override { return mDiagnostic.AICc; }
    double diagnosticENP() override { return mDiagnostic.ENP; }
    double diagnosticEDF() override { return mDiagnostic.EDF; }
    double diagnosticRSquare() override { return mDiagnostic.RSquare; }
    double diagnosticRSquareAdjust() override { return mDiagnostic.RSquareAdjust; }

    double optimizedBandwidth() override { return mOptimizedBandwidth; }
    unsigned long long selectedVarSize() override { return mSelected.size(); }
    unsigned long long selectedVar(unsigned long long i) override { return i < mSelected.size() ? mSelected[i] : 0; }
    unsigned long long variableSelectionCriterionSize() override { return mVarCriteria.size(); }
    unsigned long long variableSelectionCriterionItemVarSize(unsigned long long i) override {
        return i < mVarCriteria.size() ? mVarCriteria[i].first.size() : 0;
    }
    unsigned long long variableSelectionCriterionItemVar(unsigned long long i, unsigned long long j) override {
        return i < mVarCriteria.size() && j < mVarCriteria[i].first.size() ? mVarCriteria[i].first[j] : 0;
    }
    double variableSelectionCriterionItemValue(unsigned long long i) override {
        return i < mVarCriteria.size() ? mVarCriteria[i].second : std::numeric_limits<double>::quiet_NaN();
    }

    bool fit(bool hasIntercept) override {
        if (mDistanceType != gwm::Distance::CRSDistance) return false;
        try {
            gwm::CRSDistance distance(mGeographic);
            gwm::BandwidthWeight weight(mBandwidth, mAdaptive, (gwm::BandwidthWeight::KernelFunctionType)mKernel);
            gwm::SpatialWeight sw(&weight, &distance);
            gwm::GWRBasic algorithm(mX, mY, mCoords, sw, true, hasIntercept);
            algorithm.setIsAutoselectBandwidth(mOptimizeBandwidth);
            algorithm.setBandwidthSelectionCriterion((gwm::GWRBasic::BandwidthSelectionCriterionType)mCriterion);
            algorithm.setIsAutoselectIndepVars(mOptimizeVariables);
            algorithm.setIndepVarSelectionThreshold(mThreshold);
            algorithm.setParallelType(gwm::ParallelType::HIP);
            algorithm.fit();
            if (algorithm.status() != gwm::Status::Success) return false;
            mBetas = algorithm.betas();
            mBetasSE = algorithm.betasSE();
            mSHat = algorithm.sHat();
            mQDiag = algorithm.qDiag();
            mS = algorithm.s();
            mDiagnostic = algorithm.diagnostic();
            if (mOptimizeBandwidth) mOptimizedBandwidth = algorithm.spatialWeight().weight<gwm::BandwidthWeight>()->bandwidth();
            if (mOptimizeVariables) {
                mSelected = algorithm.selectedVariables();
                mVarCriteria = algorithm.indepVarsSelectionCriterionList();
            }
        } catch (const std::exception&) {
            return false;
        }
        return true;
    }

    bool predict(bool hasIntercept) override {
        if (mDistanceType != gwm::Distance::CRSDistance) return false;
        try {
            gwm::CRSDistance distance(mGeographic);
            gwm::BandwidthWeight weight(mBandwidth, mAdaptive, (gwm::BandwidthWeight::KernelFunctionType)mKernel);
            gwm::SpatialWeight sw(&weight, &distance);
            gwm::GWRBasic algorithm(mX, mY, mCoords, sw, true, hasIntercept);
            algorithm.setParallelType(gwm::ParallelType::HIP);
            algorithm.predict(mLocations);
            if (algorithm.status() != gwm::Status::Success) return false;
            mBetas = algorithm.betas();
        } catch (const std::exception&) {
            return false;
        }
        return true;
    }

private:
    static bool in(long long i, std::size_t n) { return i >= 0 && (std::size_t)i < n; }
    static double at(const Mat& m, long long i, long long k) {
        return in(i, m.n_rows()) && in(k, m.n_cols()) ? m(i, k) : std::numeric_limits<double>::quiet_NaN();
    }

    Mat mX, mY, mCoords, mLocations;
    Mat mBetas, mBetasSE, mSHat, mQDiag, mS;
    int mDistanceType;
    bool mGeographic = false;
    int mPoly = 2;
    double mTheta = 0.0;
    double mBandwidth = 0.0;
    bool mAdaptive = true;
    int mKernel = gwm::BandwidthWeight::Gaussian;
    bool mOptimizeBandwidth = false;
    int mCriterion = gwm::GWRBasic::CV;   // the reference task's default criterion (GWRBasicGpuTask.h:24)
    bool mOptimizeVariables = false;
    double mThreshold = 3.0;
    gwm::RegressionDiagnostic mDiagnostic{};
    double mOptimizedBandwidth = 0.0;
    std::vector<std::size_t> mSelected;
    gwm::VariablesCriterionList mVarCriteria;
};

IGWRBasicGpuTask* create(int nDp, int nVar, int distanceType, int nPredict) {
    if (nDp < 0 || nVar < 0 || nPredict < 0) return nullptr;
    try {
        return new GWRBasicHipTask(nDp, nVar, distanceType, nPredict);
    } catch (const std::exception&) {
        return nullptr;
    }
}

}  // namespace

extern "C" IGWRBasicGpuTask* GWRBasicGpuTaskFit_Create(int nDp, int nVar, int distanceType) {
    return create(nDp, nVar, distanceType, 0);
}

extern "C" IGWRBasicGpuTask* GWRBasicGpuTaskPredict_Create(int nDp, int nVar, int distanceType, int nPredictPoints) {
    return create(nDp, nVar, distanceType, nPredictPoints);
}

extern "C" void GWRBasicGpuTask_Del(IGWRBasicGpuTask* pInstance) { delete static_cast<GWRBasicHipTask*>(pInstance); }
