// Known-answer tests of the C++ host interface on the device, restating the
// reference's Catch2 sections for londonhp100 (test/testGWRBasic.cpp:40-240,
// test/testGWRMultiscale.cpp:45-88, 135-173) with ParallelType::HIP.
// Usage: test_facade <dir with londonhp100coords.csv and londonhp100data.csv>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gwmodelpp/gwmodel.h"

using namespace gwm;

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

static Mat read_csv(const std::string& path) {
    std::ifstream f(path);
    std::string line;
    std::getline(f, line);   // header
    std::vector<std::vector<double>> rows;
    while (std::getline(f, line)) {
        if (line.empty()) continue;
        std::stringstream ss(line);
        std::string cell;
        std::vector<double> r;
        while (std::getline(ss, cell, ',')) r.push_back(std::stod(cell));
        rows.push_back(r);
    }
    Mat m(rows.size(), rows.empty() ? 0 : rows[0].size());
    for (std::size_t i = 0; i < rows.size(); ++i)
        for (std::size_t j = 0; j < rows[i].size(); ++j) m(i, j) = rows[i][j];
    return m;
}

class StopAfter : public Logger {
public:
    bool stop() override { return true; }
};

// Records what the algorithm reports through ITelegram.
class Recorder : public Logger {
public:
    std::vector<std::string> messages;
    std::vector<ITelegram::LogLevel> levels;
    int percent_calls = 0, count_calls = 0;
    void print(std::string message, LogLevel level, std::string, std::string) override {
        messages.push_back(message);
        levels.push_back(level);
    }
    void progress(std::size_t, std::size_t, std::string, std::string) override { ++count_calls; }
    void progress(double percent, std::string, std::string) override {
        if (percent >= 0.0 && percent <= 1.0) ++percent_calls;
    }
};

static bool starts_with(const std::string& s, const std::string& p) { return s.compare(0, p.size(), p) == 0; }

int main(int argc, char** argv) {
    std::string dir = argc > 1 ? argv[1] : "tests/golden";
    Mat coords = read_csv(dir + "/londonhp100coords.csv");
    Mat data = read_csv(dir + "/londonhp100data.csv");
    const std::size_t n = coords.n_rows();
    Vec y = data.col(0);
    Mat x(n, 4), xm(n, 3);
    for (std::size_t i = 0; i < n; ++i) {
        x(i, 0) = 1.0; x(i, 1) = data(i, 1); x(i, 2) = data(i, 2); x(i, 3) = data(i, 3);
        xm(i, 0) = 1.0; xm(i, 1) = data(i, 1); xm(i, 2) = data(i, 3);
    }

    {   // adaptive bandwidth | no bandwidth optimization
        CRSDistance distance(false);
        BandwidthWeight bandwidth(36, true, BandwidthWeight::Gaussian);
        SpatialWeight spatial(&bandwidth, &distance);
        GWRBasic algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(x);
        algorithm.setSpatialWeight(spatial);
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.setGPUId(0);
        algorithm.setGroupSize(64);
        algorithm.fit();
        REQUIRE(algorithm.hasIntercept() == true);
        RegressionDiagnostic d = algorithm.diagnostic();
        REQUIRE_NEAR(d.AIC, 2436.60445730413, 1e-8);
        REQUIRE_NEAR(d.AICc, 2448.27206524754, 1e-8);
        REQUIRE_NEAR(d.RSquare, 0.708010632044736, 1e-8);
        REQUIRE_NEAR(d.RSquareAdjust, 0.674975341723766, 1e-8);
        REQUIRE(algorithm.s().n_rows() == n);   // n < 8192: S is stored
    }
    {   // predict at the data coordinates: the fit's betas (same weights and normal equations)
        CRSDistance distance(false);
        BandwidthWeight bandwidth(36, true, BandwidthWeight::Gaussian);
        SpatialWeight spatial(&bandwidth, &distance);
        GWRBasic algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(x);
        algorithm.setSpatialWeight(spatial);
        algorithm.setParallelType(ParallelType::HIP);
        Mat fitted = algorithm.fit();
        Mat predicted = algorithm.predict(coords);
        REQUIRE(predicted.n_rows() == n && predicted.n_cols() == 4);
        double worst = 0.0;
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < 4; ++j)
                worst = std::fmax(worst, std::fabs(predicted(i, j) - fitted(i, j)) / (1.0 + std::fabs(fitted(i, j))));
        REQUIRE_NEAR(worst, 0.0, 1e-9);
        REQUIRE(algorithm.status() == Status::Success);
    }
    {   // fixed bandwidth | no bandwidth optimization
        CRSDistance distance(false);
        BandwidthWeight bandwidth(5000, false, BandwidthWeight::Gaussian);
        SpatialWeight spatial(&bandwidth, &distance);
        GWRBasic algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(x);
        algorithm.setSpatialWeight(spatial);
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.fit();
        RegressionDiagnostic d = algorithm.diagnostic();
        REQUIRE_NEAR(d.AIC, 2437.649574267587, 1e-8);
        REQUIRE_NEAR(d.AICc, 2447.676281164379, 1e-8);
        REQUIRE_NEAR(d.RSquare, 0.701466295457, 1e-8);
        REQUIRE_NEAR(d.RSquareAdjust, 0.673691594464, 1e-8);
    }
    for (int prec = 0; prec < 2; ++prec) {   // CV bandwidth optimization (FP64 and mixed)
        CRSDistance distance(false);
        BandwidthWeight bandwidth(0, true, BandwidthWeight::Gaussian);
        SpatialWeight spatial(&bandwidth, &distance);
        GWRBasic algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(x);
        algorithm.setSpatialWeight(spatial);
        algorithm.setIsAutoselectBandwidth(true);
        algorithm.setBandwidthSelectionCriterion(GWRBasic::BandwidthSelectionCriterionType::CV);
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.setPrecision(prec ? Precision::Mixed : Precision::FP64);
        Recorder telegram;
        algorithm.setTelegram(&telegram);
        algorithm.fit();
        std::size_t bw = (std::size_t)algorithm.spatialWeight().weight<BandwidthWeight>()->bandwidth();
        REQUIRE(bw == 67);
        const BandwidthCriterionList& list = algorithm.bandwidthSelectionCriterionList();
        REQUIRE(!list.empty());
        // GWRBasic.cpp:38-102: the stage tags, the criterion header, one
        // message per evaluation, then fitting and diagnostics
        const std::vector<std::string>& m = telegram.messages;
        REQUIRE(m.size() >= 7);
        if (m.size() >= 7) {
            REQUIRE(m[0] == "#stage Initializing");
            REQUIRE(m[1] == "#stage Bandwidth selection");
            REQUIRE(m[2] == "#bandwidth-criterion adaptive,criterion");
            REQUIRE(m[m.size() - 2] == "#stage Model fitting");
            REQUIRE(m[m.size() - 1] == "#stage Model Diagnostic");
            std::size_t evals = 0;
            bool seen67 = false;
            for (std::size_t q = 3; q + 2 < m.size(); ++q) {
                REQUIRE(starts_with(m[q], "#bandwidth-criterion "));
                const std::string body = m[q].substr(21);
                const std::string bws = body.substr(0, body.find(','));
                bool listed = false;
                for (const auto& e : list) listed = listed || std::to_string((int)e.first) == bws;
                REQUIRE(listed);
                seen67 = seen67 || bws == "67";
                ++evals;
            }
            REQUIRE(seen67);
            REQUIRE(evals >= list.size());
            REQUIRE(telegram.percent_calls == (int)evals);
            REQUIRE(telegram.count_calls == 1);   // the fit
            for (auto lv : telegram.levels) REQUIRE(lv == ITelegram::LogLevel::LOG_INFO);
        }
    }
    for (int optimize_bw = 0; optimize_bw < 2; ++optimize_bw) {
        // adaptive bandwidth | {no, CV} bandwidth optimization | AIC variable optimization
        CRSDistance distance(false);
        BandwidthWeight bandwidth(36, true, BandwidthWeight::Gaussian);
        SpatialWeight spatial(&bandwidth, &distance);
        GWRBasic algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(x);
        algorithm.setSpatialWeight(spatial);
        if (optimize_bw) {
            algorithm.setIsAutoselectBandwidth(true);
            algorithm.setBandwidthSelectionCriterion(GWRBasic::BandwidthSelectionCriterionType::CV);
        }
        algorithm.setIsAutoselectIndepVars(true);
        algorithm.setIndepVarSelectionThreshold(3.0);
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.fit();
        const VariablesCriterionList& c = algorithm.indepVarsSelectionCriterionList();
        const std::vector<std::vector<std::size_t>> vars = {{2}, {3}, {1}, {1, 3}, {1, 2}, {1, 2, 3}};
        const double values[] = {2551.61359020599, 2551.30032201349, 2468.93236280013,
                                 2452.86447942033, 2450.59642666509, 2452.80388934625};
        REQUIRE(c.size() == 6);
        for (std::size_t q = 0; q < c.size() && q < 6; ++q) {
            REQUIRE(c[q].first == vars[q]);
            REQUIRE_NEAR(c[q].second, values[q], 1e-8);
        }
        REQUIRE(algorithm.selectedVariables() == std::vector<std::size_t>({1, 3}));
        if (optimize_bw) {
            REQUIRE(algorithm.spatialWeight().weight<BandwidthWeight>()->bandwidth() == 31.0);
            RegressionDiagnostic d = algorithm.diagnostic();
            REQUIRE_NEAR(d.AIC, 2435.8161441795, 1e-8);
            REQUIRE_NEAR(d.AICc, 2445.49629974057, 1e-8);
            REQUIRE_NEAR(d.RSquare, 0.706143867720706, 1e-8);
            REQUIRE_NEAR(d.RSquareAdjust, 0.678982114793865, 1e-8);
        }
    }
    {   // cancellation: Status::Terminated and zero betas
        CRSDistance distance(false);
        BandwidthWeight bandwidth(0, true, BandwidthWeight::Gaussian);
        SpatialWeight spatial(&bandwidth, &distance);
        GWRBasic algorithm;
        StopAfter telegram;
        algorithm.setTelegram(&telegram);
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(x);
        algorithm.setSpatialWeight(spatial);
        algorithm.setIsAutoselectBandwidth(true);
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.fit();
        REQUIRE(algorithm.status() == Status::Terminated);
    }
    {   // MGWR: optim bw cv | null init bw | adaptive | bisquare | without hat matrix
        std::vector<SpatialWeight> spatials;
        std::vector<bool> centered;
        std::vector<GWRMultiscale::BandwidthInitilizeType> init;
        std::vector<GWRMultiscale::BandwidthSelectionCriterionType> approach;
        for (int i = 0; i < 3; ++i) {
            CRSDistance distance;
            BandwidthWeight bandwidth(0, true, BandwidthWeight::Bisquare);
            spatials.push_back(SpatialWeight(&bandwidth, &distance));
            centered.push_back(i != 0);
            init.push_back(GWRMultiscale::BandwidthInitilizeType::Null);
            approach.push_back(GWRMultiscale::BandwidthSelectionCriterionType::CV);
        }
        GWRMultiscale algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(xm);
        algorithm.setSpatialWeights(spatials);
        algorithm.setHasHatMatrix(false);
        algorithm.setPreditorCentered(centered);
        algorithm.setBandwidthInitilize(init);
        algorithm.setBandwidthSelectionApproach(approach);
        algorithm.setBandwidthSelectThreshold(std::vector<double>(3, 1e-5));
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.fit();
        const std::vector<SpatialWeight>& sw = algorithm.spatialWeights();
        REQUIRE(sw[0].weight<BandwidthWeight>()->bandwidth() == 35);
        REQUIRE(sw[1].weight<BandwidthWeight>()->bandwidth() == 98);
        REQUIRE(sw[2].weight<BandwidthWeight>()->bandwidth() == 98);
        REQUIRE_NEAR(algorithm.diagnostic().RSquare, 0.757377391669, 1e-6);
    }
    {   // MGWR: optim bw cv | null init bw | fixed | bisquare | with hat matrix (testGWRMultiscale.cpp:45-88)
        std::vector<SpatialWeight> spatials;
        std::vector<bool> centered;
        std::vector<GWRMultiscale::BandwidthInitilizeType> init;
        std::vector<GWRMultiscale::BandwidthSelectionCriterionType> approach;
        for (int i = 0; i < 3; ++i) {
            CRSDistance distance;
            BandwidthWeight bandwidth(0, false, BandwidthWeight::Bisquare);
            spatials.push_back(SpatialWeight(&bandwidth, &distance));
            centered.push_back(i != 0);
            init.push_back(GWRMultiscale::BandwidthInitilizeType::Null);
            approach.push_back(GWRMultiscale::BandwidthSelectionCriterionType::CV);
        }
        GWRMultiscale algorithm;
        algorithm.setCoords(coords);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(xm);
        algorithm.setSpatialWeights(spatials);
        algorithm.setHasHatMatrix(true);
        algorithm.setPreditorCentered(centered);
        algorithm.setBandwidthInitilize(init);
        algorithm.setBandwidthSelectionApproach(approach);
        algorithm.setBandwidthSelectThreshold(std::vector<double>(3, 1e-5));
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.fit();
        const std::vector<SpatialWeight>& sw = algorithm.spatialWeights();
        REQUIRE_NEAR(sw[0].weight<BandwidthWeight>()->bandwidth(), 4623.78, 0.1);
        REQUIRE_NEAR(sw[1].weight<BandwidthWeight>()->bandwidth(), 12665.70, 0.1);
        REQUIRE_NEAR(sw[2].weight<BandwidthWeight>()->bandwidth(), 12665.70, 0.1);
        RegressionDiagnostic d = algorithm.diagnostic();
        REQUIRE_NEAR(d.AICc, 2437.09277417389, 1e-6);
        REQUIRE_NEAR(d.RSquare, 0.744649364494, 1e-6);
        REQUIRE_NEAR(d.RSquareAdjust, 0.712344894394, 1e-6);
    }
    {   // MGWR with one distance per variable: a great-circle variable among
        // Euclidean ones (GWRMultiscale.h:578); coordinates read as degrees
        Mat ll(n, 2);
        for (std::size_t i = 0; i < n; ++i) {   // londonhp100 metres to a lon/lat box
            ll(i, 0) = -0.5 + (coords(i, 0) - 500000.0) / 70000.0;
            ll(i, 1) = 51.3 + (coords(i, 1) - 155000.0) / 111000.0;
        }
        std::vector<SpatialWeight> spatials;
        for (int i = 0; i < 3; ++i) {
            CRSDistance distance(i == 1);
            BandwidthWeight bandwidth(0, true, BandwidthWeight::Bisquare);
            spatials.push_back(SpatialWeight(&bandwidth, &distance));
        }
        GWRMultiscale algorithm;
        Recorder telegram;
        algorithm.setTelegram(&telegram);
        algorithm.setCoords(ll);
        algorithm.setDependentVariable(y);
        algorithm.setIndependentVariables(xm);
        algorithm.setSpatialWeights(spatials);
        algorithm.setHasHatMatrix(false);
        algorithm.setPreditorCentered({false, true, true});
        algorithm.setBandwidthInitilize(std::vector<GWRMultiscale::BandwidthInitilizeType>(
            3, GWRMultiscale::BandwidthInitilizeType::Null));
        algorithm.setBandwidthSelectionApproach(std::vector<GWRMultiscale::BandwidthSelectionCriterionType>(
            3, GWRMultiscale::BandwidthSelectionCriterionType::CV));
        algorithm.setBandwidthSelectThreshold(std::vector<double>(3, 1e-5));
        algorithm.setParallelType(ParallelType::HIP);
        algorithm.fit();
        REQUIRE(algorithm.status() == Status::Success);
        for (int i = 0; i < 3; ++i) {
            const double b = algorithm.spatialWeights()[i].weight<BandwidthWeight>()->bandwidth();
            REQUIRE(b >= 10 && b <= (double)n);
        }
        const std::vector<std::string>& m = telegram.messages;
        REQUIRE(!m.empty() && m[0] == "#stage Initializing");
        REQUIRE(!m.empty() && m.back() == "#stage Model Diagnostic");
    }
    std::printf(failures ? "FACADE TESTS FAILED: %d\n" : "FACADE TESTS PASSED\n", failures);
    return failures ? 1 : 0;
}
