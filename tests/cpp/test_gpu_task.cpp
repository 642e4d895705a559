// The gwmodelcuda-compatible entry (IGWRBasicGpuTask, factory functions) on
// the device: londonhp100 loaded element by element, as a desktop host does,
// checked against the reference's KATs (test/testGWRBasic.cpp:40-68, 105-137,
// 186-240).
// Usage: test_gpu_task <dir with londonhp100coords.csv and londonhp100data.csv>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gwmodelpp/IGWRBasicGpuTask.h"

static int failures = 0;
#define REQUIRE(cond)                                                              \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);          \
            ++failures;                                                            \
        }                                                                          \
    } while (0)
#define REQUIRE_NEAR(a, b, tol)                                                    \
    do {                                                                           \
        double a_ = (a), b_ = (b);                                                 \
        if (!(std::fabs(a_ - b_) <= (tol))) {                                      \
            std::printf("FAILED %s:%d: %s = %.15g, expected %.15g\n", __FILE__, __LINE__, #a, a_, b_); \
            ++failures;                                                            \
        }                                                                          \
    } while (0)

static std::vector<std::vector<double>> read_csv(const std::string& path) {
    std::ifstream f(path);
    std::string line;
    std::getline(f, line);
    std::vector<std::vector<double>> rows;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        std::string cell;
        std::vector<double> r;
        while (std::getline(ss, cell, ',')) r.push_back(std::stod(cell));
        rows.push_back(r);
    }
    return rows;
}

struct Data {
    std::vector<std::vector<double>> coords, data;
    int n() const { return (int)coords.size(); }
};

static void load(IGWRBasicGpuTask* t, const Data& d) {
    for (int i = 0; i < d.n(); ++i) {
        t->setCoords(i, d.coords[i][0], d.coords[i][1]);
        t->setY(i, d.data[i][0]);
        t->setX(i, 0, 1.0);
        for (int k = 1; k < 4; ++k) t->setX(i, k, d.data[i][k]);
    }
}

int main(int argc, char** argv) {
    std::string dir = argc > 1 ? argv[1] : "tests/golden";
    Data d{read_csv(dir + "/londonhp100coords.csv"), read_csv(dir + "/londonhp100data.csv")};
    const int n = d.n();
    std::vector<double> fitted(n * 4);
    {   // adaptive 36, Gaussian, no optimization
        IGWRBasicGpuTask* t = GWRBasicGpuTaskFit_Create(n, 4, 0);
        load(t, d);
        t->setBandwidthSize(36);
        t->setBandwidthAdaptive(true);
        t->setBandwidthKernel(0);
        REQUIRE(t->fit(true));
        REQUIRE_NEAR(t->diagnosticAIC(), 2436.60445730413, 1e-8);
        REQUIRE_NEAR(t->diagnosticAICc(), 2448.27206524754, 1e-8);
        REQUIRE_NEAR(t->diagnosticRSquare(), 0.708010632044736, 1e-8);
        REQUIRE_NEAR(t->diagnosticRSquareAdjust(), 0.674975341723766, 1e-8);
        REQUIRE(t->sRows() == (unsigned long long)n);
        REQUIRE(std::isfinite(t->betasSE(0, 0)) && std::isfinite(t->qDiag(n - 1)));
        REQUIRE(std::isnan(t->betas(n, 0)));
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k) fitted[i + k * n] = t->betas(i, k);
        GWRBasicGpuTask_Del(t);
    }
    {   // CV bandwidth optimization: bw 67
        IGWRBasicGpuTask* t = GWRBasicGpuTaskFit_Create(n, 4, 0);
        load(t, d);
        t->setBandwidthSize(0);
        t->setBandwidthAdaptive(true);
        t->setBandwidthKernel(0);
        t->enableBandwidthOptimization(1);
        REQUIRE(t->fit(true));
        REQUIRE(t->optimizedBandwidth() == 67.0);
        GWRBasicGpuTask_Del(t);
    }
    {   // CV bandwidth + AIC variable optimization: {1, 3}, bw 31
        IGWRBasicGpuTask* t = GWRBasicGpuTaskFit_Create(n, 4, 0);
        load(t, d);
        t->setBandwidthSize(36);
        t->setBandwidthAdaptive(true);
        t->setBandwidthKernel(0);
        t->enableBandwidthOptimization(1);
        t->enableVariablesOptimization(3.0);
        REQUIRE(t->fit(true));
        REQUIRE(t->selectedVarSize() == 2 && t->selectedVar(0) == 1 && t->selectedVar(1) == 3);
        REQUIRE(t->variableSelectionCriterionSize() == 6);
        REQUIRE(t->variableSelectionCriterionItemVarSize(3) == 2);
        REQUIRE(t->variableSelectionCriterionItemVar(3, 0) == 1 && t->variableSelectionCriterionItemVar(3, 1) == 3);
        REQUIRE_NEAR(t->variableSelectionCriterionItemValue(3), 2452.86447942033, 1e-8);
        REQUIRE(t->optimizedBandwidth() == 31.0);
        REQUIRE_NEAR(t->diagnosticAIC(), 2435.8161441795, 1e-8);
        REQUIRE_NEAR(t->diagnosticRSquareAdjust(), 0.678982114793865, 1e-8);
        GWRBasicGpuTask_Del(t);
    }
    {   // predict at the data coordinates: the fit's betas
        IGWRBasicGpuTask* t = GWRBasicGpuTaskPredict_Create(n, 4, 0, n);
        load(t, d);
        for (int i = 0; i < n; ++i) t->setPredictLocations(i, d.coords[i][0], d.coords[i][1]);
        t->setBandwidthSize(36);
        t->setBandwidthAdaptive(true);
        t->setBandwidthKernel(0);
        REQUIRE(t->predict(true));
        double worst = 0.0;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 4; ++k)
                worst = std::fmax(worst, std::fabs(t->betas(i, k) - fitted[i + k * n]) / (1.0 + std::fabs(fitted[i + k * n])));
        REQUIRE_NEAR(worst, 0.0, 1e-9);
        GWRBasicGpuTask_Del(t);
    }
    {   // the Minkowski distance is not served: fit reports failure
        IGWRBasicGpuTask* t = GWRBasicGpuTaskFit_Create(n, 4, 1);
        load(t, d);
        t->setBandwidthSize(36);
        REQUIRE(!t->fit(true));
        GWRBasicGpuTask_Del(t);
        REQUIRE(GWRBasicGpuTaskFit_Create(-1, 4, 0) == nullptr);
    }
    std::printf(failures ? "GPU TASK TESTS FAILED: %d\n" : "GPU TASK TESTS PASSED\n", failures);
    return failures ? 1 : 0;
}
