// gwmodel.h -- C++ host interface of the MI355X GWR back-end.
//
// Mirrors the parts of libgwmodel's C++ API that the GWR hot path uses, with
// the same names, argument meaning and error behaviour, so that code written
// against gwm::GWRBasic / gwm::GWRMultiscale runs unchanged with
// setParallelType(ParallelType::HIP):
//   - ParallelType            include/gwmodelpp/IParallelizable.h:15-20 (+ HIP = 1 << 3)
//   - Status, RegressionDiagnostic, ITelegram/Logger  Status.h, RegressionDiagnostic.h, Logger.h:31-132
//   - CRSDistance             include/gwmodelpp/spatialweight/CRSDistance.h:17-230
//   - BandwidthWeight         include/gwmodelpp/spatialweight/BandwidthWeight.h:39-261
//   - SpatialWeight           include/gwmodelpp/spatialweight/SpatialWeight.h:43-546
//   - GWRBasic                include/gwmodelpp/GWRBasic.h:29-757
//   - GWRMultiscale           include/gwmodelpp/GWRMultiscale.h:35-1026
// Every computation runs on the device through the C ABI of gwmodel_hip.h;
// there is no CPU back-end here.  Armadillo is replaced by gwm::Mat, a
// column-major FP64 matrix with the same memory layout as arma::mat.
#ifndef GWMODELPP_GWMODEL_H
#define GWMODELPP_GWMODEL_H

#include <cstddef>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

struct gwm_hip_ctx;

namespace gwm {

// ---------------------------------------------------------------- matrices

class Mat {
public:
    Mat() = default;
    Mat(std::size_t rows, std::size_t cols, double fill = 0.0) : r_(rows), c_(cols), d_(rows * cols, fill) {}
    Mat(std::size_t rows, std::size_t cols, const double* colmajor)
        : r_(rows), c_(cols), d_(colmajor, colmajor + rows * cols) {}
    std::size_t n_rows() const { return r_; }
    std::size_t n_cols() const { return c_; }
    std::size_t n_elem() const { return d_.size(); }
    double& operator()(std::size_t i, std::size_t j) { return d_[i + j * r_]; }
    double operator()(std::size_t i, std::size_t j) const { return d_[i + j * r_]; }
    double& operator()(std::size_t i) { return d_[i]; }
    double operator()(std::size_t i) const { return d_[i]; }
    double* memptr() { return d_.data(); }
    const double* memptr() const { return d_.data(); }
    Mat col(std::size_t j) const { return Mat(r_, 1, d_.data() + j * r_); }
    void set_col(std::size_t j, const Mat& v) {
        for (std::size_t i = 0; i < r_; ++i) d_[i + j * r_] = v(i);
    }

private:
    std::size_t r_ = 0, c_ = 0;
    std::vector<double> d_;
};
typedef Mat Vec;

// ---------------------------------------------------------------- framework

enum ParallelType {
    SerialOnly = 1 << 0,
    OpenMP = 1 << 1,
    CUDA = 1 << 2,
    HIP = 1 << 3   //!< MI355X (gfx950) back-end
};

enum class Status { Success = 0, Terminated };

struct RegressionDiagnostic {
    double RSS;
    double AIC;
    double AICc;
    double ENP;
    double EDF;
    double RSquare;
    double RSquareAdjust;
};

struct ITelegram {
    virtual ~ITelegram() {}
    enum class LogLevel { LOG_EMERG = 0, LOG_ALERT, LOG_CRIT, LOG_ERR, LOG_WARNING, LOG_NOTICE, LOG_INFO, LOG_DEBUG };
    virtual void print(std::string message, LogLevel level, std::string fun_name, std::string file_name) = 0;
    virtual void progress(std::size_t current, std::size_t total, std::string fun_name, std::string file_name) = 0;
    virtual void progress(double percent, std::string fun_name, std::string file_name) = 0;
    virtual bool stop() = 0;
};

class Logger : public ITelegram {
public:
    void print(std::string, LogLevel, std::string, std::string) override {}
    void progress(std::size_t, std::size_t, std::string, std::string) override {}
    void progress(double, std::string, std::string) override {}
    bool stop() override { return false; }
};

// Arithmetic of the device sweeps (extension: the reference is FP64 only).
enum class Precision { FP64 = 0, Mixed = 1 };

// ---------------------------------------------------------------- spatial weights

class Distance {
public:
    enum DistanceType { CRSDistance, MinkwoskiDistance, DMatDistance, OneDimDistance, CRSSTDistance };
    virtual ~Distance() {}
    virtual Distance* clone() const = 0;
    virtual DistanceType type() = 0;
};

class CRSDistance : public Distance {
public:
    CRSDistance() : mGeographic(false) {}
    explicit CRSDistance(bool isGeographic) : mGeographic(isGeographic) {}
    Distance* clone() const override { return new CRSDistance(*this); }
    DistanceType type() override { return DistanceType::CRSDistance; }
    bool geographic() const { return mGeographic; }
    void setGeographic(bool geographic) { mGeographic = geographic; }

private:
    bool mGeographic;
};

class Weight {
public:
    enum WeightType { BandwidthWeight };
    virtual ~Weight() {}
    virtual Weight* clone() const = 0;
};

class BandwidthWeight : public Weight {
public:
    enum KernelFunctionType { Gaussian, Exponential, Bisquare, Tricube, Boxcar };
    BandwidthWeight() {}
    BandwidthWeight(double size, bool adaptive, KernelFunctionType kernel)
        : mBandwidth(size), mAdaptive(adaptive), mKernel(kernel) {}
    Weight* clone() const override { return new BandwidthWeight(*this); }
    double bandwidth() const { return mBandwidth; }
    void setBandwidth(double bandwidth) { mBandwidth = bandwidth; }
    bool adaptive() const { return mAdaptive; }
    void setAdaptive(bool adaptive) { mAdaptive = adaptive; }
    KernelFunctionType kernel() const { return mKernel; }
    void setKernel(const KernelFunctionType& kernel) { mKernel = kernel; }

private:
    double mBandwidth = 0.0;
    bool mAdaptive = true;
    KernelFunctionType mKernel = Gaussian;
};

class SpatialWeight {
public:
    SpatialWeight() {}
    SpatialWeight(const Weight* weight, const Distance* distance)
        : mWeight(weight->clone()), mDistance(distance->clone()) {}
    SpatialWeight(const SpatialWeight& o)
        : mWeight(o.mWeight ? o.mWeight->clone() : nullptr), mDistance(o.mDistance ? o.mDistance->clone() : nullptr) {}
    SpatialWeight& operator=(const SpatialWeight& o) {
        if (this != &o) {
            mWeight.reset(o.mWeight ? o.mWeight->clone() : nullptr);
            mDistance.reset(o.mDistance ? o.mDistance->clone() : nullptr);
        }
        return *this;
    }
    template <class T>
    T* weight() const { return static_cast<T*>(mWeight.get()); }
    template <class T>
    T* distance() const { return static_cast<T*>(mDistance.get()); }
    Weight* weight() const { return mWeight.get(); }
    Distance* distance() const { return mDistance.get(); }
    void setWeight(const Weight& w) { mWeight.reset(w.clone()); }
    void setDistance(const Distance& d) { mDistance.reset(d.clone()); }
    bool isValid() const { return mWeight && mDistance; }

private:
    std::unique_ptr<Weight> mWeight;
    std::unique_ptr<Distance> mDistance;
};

typedef std::vector<std::pair<double, double>> BandwidthCriterionList;
// IVarialbeSelectable.h:82
typedef std::vector<std::pair<std::vector<std::size_t>, double>> VariablesCriterionList;

// Owns one device context; shared by the algorithm objects below.
class HipContext {
public:
    explicit HipContext(int device);
    ~HipContext();
    HipContext(const HipContext&) = delete;
    HipContext& operator=(const HipContext&) = delete;
    gwm_hip_ctx* get() const { return mCtx; }

private:
    gwm_hip_ctx* mCtx = nullptr;
};

// ---------------------------------------------------------------- GWRBasic

class GWRBasic {
public:
    enum BandwidthSelectionCriterionType { AIC, CV };

    GWRBasic() {}
    GWRBasic(const Mat& x, const Vec& y, const Mat& coords, const SpatialWeight& spatialWeight,
             bool hasHatMatrix = true, bool hasIntercept = true)
        : mX(x), mY(y), mCoords(coords), mSpatialWeight(spatialWeight), mHasHatMatrix(hasHatMatrix),
          mHasIntercept(hasIntercept) {}

    void setCoords(const Mat& coords) { mCoords = coords; }
    const Mat& coords() const { return mCoords; }
    void setDependentVariable(const Vec& y) { mY = y; }
    void setIndependentVariables(const Mat& x) { mX = x; }
    void setSpatialWeight(const SpatialWeight& w) { mSpatialWeight = w; }
    const SpatialWeight& spatialWeight() const { return mSpatialWeight; }
    void setHasIntercept(bool has) { mHasIntercept = has; }
    bool hasIntercept() const { return mHasIntercept; }
    void setHasHatMatrix(bool has) { mHasHatMatrix = has; }
    bool hasHatMatrix() const { return mHasHatMatrix; }
    void setIsAutoselectBandwidth(bool v) { mIsAutoselectBandwidth = v; }
    bool isAutoselectBandwidth() const { return mIsAutoselectBandwidth; }
    void setBandwidthSelectionCriterion(const BandwidthSelectionCriterionType& c) { mBandwidthSelectionCriterion = c; }
    BandwidthSelectionCriterionType bandwidthSelectionCriterion() const { return mBandwidthSelectionCriterion; }
    void setGoldenUpperBounds(double value) { mGoldenUpperBounds = value; }
    void setGoldenLowerBounds(double value) { mGoldenLowerBounds = value; }
    const BandwidthCriterionList& bandwidthSelectionCriterionList() const { return mBandwidthSelectionCriterionList; }
    // IVarialbeSelectable (GWRBasic.h:213-269, 380-383)
    void setIsAutoselectIndepVars(bool v) { mIsAutoselectIndepVars = v; }
    bool isAutoselectIndepVars() const { return mIsAutoselectIndepVars; }
    void setIndepVarSelectionThreshold(double threshold) { mIndepVarSelectionThreshold = threshold; }
    double indepVarSelectionThreshold() const { return mIndepVarSelectionThreshold; }
    const VariablesCriterionList& indepVarsSelectionCriterionList() const { return mIndepVarsSelectionCriterionList; }
    std::vector<std::size_t> selectedVariables() const { return mSelectedIndepVars; }

    // IParallelizable (IParallelizable.h:35-69): only the HIP back-end exists here.
    int parallelAbility() const { return ParallelType::HIP; }
    ParallelType parallelType() const { return mParallelType; }
    void setParallelType(const ParallelType& type) {
        if (type & parallelAbility()) mParallelType = type;
    }
    void setGPUId(const int gpuId) { mGpuId = gpuId; }
    void setGroupSize(const std::size_t size) { mGroupLength = size; }   // tile-size hint; the kernels tile themselves
    void setPrecision(Precision p) { mPrecision = p; }
    void setTelegram(ITelegram* telegram) { mTelegram = telegram; }
    Status status() const { return mStatus; }

    Mat fit();
    // IGwmodelPredictable::predict (GWRBasic.cpp:130-156): betas at the rows
    // of locations (m x 2) with the current spatial weight; also stored in betas().
    Mat predict(const Mat& locations);

    const Mat& betas() const { return mBetas; }
    const Mat& betasSE() const { return mBetasSE; }
    const Vec& sHat() const { return mSHat; }
    const Vec& qDiag() const { return mQDiag; }
    const Mat& s() const { return mS; }
    RegressionDiagnostic diagnostic() const { return mDiagnostic; }
    bool isStoreS() const { return mHasHatMatrix && (mCoords.n_rows() < 8192); }

private:
    Mat mX, mCoords, mBetas, mBetasSE, mS;
    Vec mY, mSHat, mQDiag;
    SpatialWeight mSpatialWeight;
    bool mHasHatMatrix = true;
    bool mHasIntercept = true;
    bool mIsAutoselectBandwidth = false;
    bool mIsAutoselectIndepVars = false;
    double mIndepVarSelectionThreshold = 3.0;
    VariablesCriterionList mIndepVarsSelectionCriterionList;
    std::vector<std::size_t> mSelectedIndepVars;
    BandwidthSelectionCriterionType mBandwidthSelectionCriterion = AIC;
    std::optional<double> mGoldenUpperBounds, mGoldenLowerBounds;
    BandwidthCriterionList mBandwidthSelectionCriterionList;
    RegressionDiagnostic mDiagnostic{};
    ParallelType mParallelType = ParallelType::HIP;
    int mGpuId = 0;
    std::size_t mGroupLength = 64;
    Precision mPrecision = Precision::FP64;
    Status mStatus = Status::Success;
    Logger mDefaultTelegram;
    ITelegram* mTelegram = &mDefaultTelegram;
};

// ---------------------------------------------------------------- GWRMultiscale

class GWRMultiscale {
public:
    enum BandwidthInitilizeType { Null, Initial, Specified };
    enum BandwidthSelectionCriterionType { CV, AIC };
    enum BackFittingCriterionType { CVR, dCVR };

    void setCoords(const Mat& coords) { mCoords = coords; }
    void setDependentVariable(const Vec& y) { mY = y; }
    void setIndependentVariables(const Mat& x) { mX = x; }
    void setHasIntercept(bool has) { mHasIntercept = has; }
    bool hasIntercept() const { return mHasIntercept; }
    void setSpatialWeights(const std::vector<SpatialWeight>& w) { mSpatialWeights = w; }
    const std::vector<SpatialWeight>& spatialWeights() const { return mSpatialWeights; }
    void setHasHatMatrix(bool has) { mHasHatMatrix = has; }
    void setPreditorCentered(const std::vector<bool>& v) { mPreditorCentered = v; }
    void setBandwidthInitilize(const std::vector<BandwidthInitilizeType>& v);
    void setBandwidthSelectionApproach(const std::vector<BandwidthSelectionCriterionType>& v);
    void setBandwidthSelectThreshold(const std::vector<double>& v) { mBandwidthSelectThreshold = v; }
    void setBandwidthSelectRetryTimes(std::size_t v) { mBandwidthSelectRetryTimes = v; }
    void setMaxIteration(std::size_t v) { mMaxIteration = v; }
    void setCriterionType(const BackFittingCriterionType& v) { mCriterionType = v; }
    void setCriterionThreshold(double v) { mCriterionThreshold = v; }
    void setAdaptiveLower(int v) { mAdaptiveLower = v; }
    void setGoldenUpperBounds(double value) { mGoldenUpperBounds = value; }
    void setGoldenLowerBounds(double value) { mGoldenLowerBounds = value; }

    int parallelAbility() const { return ParallelType::HIP; }
    ParallelType parallelType() const { return mParallelType; }
    void setParallelType(const ParallelType& type) {
        if (type & parallelAbility()) mParallelType = type;
    }
    void setGPUId(const int gpuId) { mGpuId = gpuId; }
    void setGroupSize(const std::size_t size) { mGroupLength = size; }
    void setPrecision(Precision p) { mPrecision = p; }
    void setTelegram(ITelegram* telegram) { mTelegram = telegram; }
    Status status() const { return mStatus; }

    Mat fit();

    const Mat& betas() const { return mBetas; }
    RegressionDiagnostic diagnostic() const { return mDiagnostic; }

private:
    Mat mX, mCoords, mBetas;
    Vec mY;
    std::vector<SpatialWeight> mSpatialWeights;
    bool mHasHatMatrix = true;
    bool mHasIntercept = true;
    std::vector<bool> mPreditorCentered;
    std::vector<BandwidthInitilizeType> mBandwidthInitilize;
    std::vector<BandwidthSelectionCriterionType> mBandwidthSelectionApproach;
    std::vector<double> mBandwidthSelectThreshold;
    std::size_t mBandwidthSelectRetryTimes = 5;
    std::size_t mMaxIteration = 500;
    BackFittingCriterionType mCriterionType = dCVR;
    double mCriterionThreshold = 1e-6;
    int mAdaptiveLower = 10;
    std::optional<double> mGoldenUpperBounds, mGoldenLowerBounds;
    RegressionDiagnostic mDiagnostic{};
    ParallelType mParallelType = ParallelType::HIP;
    int mGpuId = 0;
    std::size_t mGroupLength = 64;
    Precision mPrecision = Precision::FP64;
    Status mStatus = Status::Success;
    Logger mDefaultTelegram;
    ITelegram* mTelegram = &mDefaultTelegram;
};

}  // namespace gwm

#endif
