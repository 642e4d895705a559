// Source- and vtable-compatible replacement of the reference's gwmodelcuda
// entry point (include/gwmodelcuda/IGWRBasicGpuTask.h:14-65): the same
// abstract interface, in the same virtual-function order, and the same three
// extern "C" factory names, served by the HIP back-end through the C++ host
// interface of gwmodel.h.  Differences from the reference's GWRBasicGpuTask:
//  - setCoords(i, u, v) and setPredictLocations(i, u, v) store (u, v) as row i
//    (the reference writes v into column u, GWRBasicGpuTask.h:138-146);
//  - only CRSDistance is served: fit() and predict() return false for the
//    Minkowski distance type (the HIP back-end has no Minkowski path);
//  - setters and getters ignore / answer NaN (0 for sizes) for indices out of
//    range instead of Armadillo's bounds-check exception;
//  - GWRBasicGpuTask_Del destroys the concrete object (the interface has no
//    virtual destructor, as in the reference, to keep its vtable layout).
#ifndef GWM_HIP_IGWRBASICGPUTASK_H
#define GWM_HIP_IGWRBASICGPUTASK_H

#define GWMODELCUDA_API __attribute__((visibility("default")))

class GWMODELCUDA_API IGWRBasicGpuTask {
public:
    virtual void setX(int i, int k, double value) = 0;
    virtual void setY(int i, double value) = 0;
    virtual void setCoords(int i, double u, double v) = 0;
    virtual void setPredictLocations(int i, double u, double v) = 0;

    virtual void setDistanceType(int type) = 0;
    virtual void setCRSDistanceGergraphic(bool isGeographic) = 0;
    virtual void setMinkwoskiDistancePoly(int poly) = 0;
    virtual void setMinkwoskiDistanceTheta(double theta) = 0;

    virtual void setBandwidthSize(double bw) = 0;
    virtual void setBandwidthAdaptive(bool adaptive) = 0;
    virtual void setBandwidthKernel(int kernel) = 0;

    virtual void enableBandwidthOptimization(int criterion) = 0;
    virtual void enableVariablesOptimization(double threshold) = 0;

    virtual double betas(int i, int k) = 0;
    virtual double betasSE(int i, int k) = 0;
    virtual double shat1() = 0;
    virtual double shat2() = 0;
    virtual double qDiag(int i) = 0;
    virtual unsigned long long sRows() = 0;
    virtual double s(int i, int k) = 0;

    virtual double diagnosticRSS() = 0;
    virtual double diagnosticAIC() = 0;
    virtual double diagnosticAICc() = 0;
    virtual double diagnosticENP() = 0;
    virtual double diagnosticEDF() = 0;
    virtual double diagnosticRSquare() = 0;
    virtual double diagnosticRSquareAdjust() = 0;

    virtual double optimizedBandwidth() = 0;
    virtual unsigned long long selectedVarSize() = 0;
    virtual unsigned long long selectedVar(unsigned long long i) = 0;
    virtual unsigned long long variableSelectionCriterionSize() = 0;
    virtual unsigned long long variableSelectionCriterionItemVarSize(unsigned long long i) = 0;
    virtual unsigned long long variableSelectionCriterionItemVar(unsigned long long i, unsigned long long j) = 0;
    virtual double variableSelectionCriterionItemValue(unsigned long long i) = 0;

    virtual bool fit(bool hasIntercept) = 0;

    virtual bool predict(bool hasIntercept) = 0;
};

// IGWRBasicGpuTask.h:63-65.  distanceType follows gwm::Distance::DistanceType
// (0 = CRSDistance, 1 = MinkwoskiDistance).  Null on invalid sizes.
extern "C" GWMODELCUDA_API IGWRBasicGpuTask* GWRBasicGpuTaskFit_Create(int nDp, int nVar, int distanceType);
extern "C" GWMODELCUDA_API IGWRBasicGpuTask* GWRBasicGpuTaskPredict_Create(int nDp, int nVar, int distanceType,
                                                                          int nPredictPoints);
extern "C" GWMODELCUDA_API void GWRBasicGpuTask_Del(IGWRBasicGpuTask* pInstance);

#endif  // GWM_HIP_IGWRBASICGPUTASK_H
