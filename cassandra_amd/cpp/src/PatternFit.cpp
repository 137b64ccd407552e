// FitWeights (PatternFit/PatternFit.cpp:7-39) and the per-stage fit on the GPU:
// positions are indexed by the HIP kernels straight into the device matrix.
#include "PatternFit/PatternFit.h"

#include <memory>
#include <stdexcept>

#include "device.h"

using namespace cad_cpp;

namespace {

struct Fit {
    std::unique_ptr<cad_patterns, int (*)(cad_patterns*)> pat{nullptr, cad_patterns_destroy};
    std::unique_ptr<cad_matrix, int (*)(cad_matrix*)> mat{nullptr, cad_matrix_destroy};
    int64_t cols = 0;
};

// b in disc-difference halves (Score), as FitWeights feeds the solver.
Vector fit(const std::vector<uint64_t>& pattern, const ScoredPosition* sp, std::size_t n, int iterations)
{
    cad_context* ctx = Context();
    Fit f;
    cad_patterns* p = nullptr;
    Check(cad_patterns_create(ctx, pattern.data(), int(pattern.size()), &p));
    f.pat.reset(p);
    int k = 0, np = 0;
    Check(cad_patterns_info(p, &k, &f.cols, &np));
    Vector x(std::size_t(f.cols), 0.0f);
    if (n == 0) return x;
    std::vector<uint64_t> po(2 * n);
    std::vector<double> b(n);
    for (std::size_t i = 0; i < n; i++) {
        po[2 * i] = sp[i].pos.Player();
        po[2 * i + 1] = sp[i].pos.Opponent();
        b[i] = double(int8_t(sp[i].score));
    }
    DevBuf dpos(sizeof(uint64_t) * po.size()), db(sizeof(double) * n), dx(sizeof(double) * std::size_t(f.cols));
    Check(cad_memcpy_h2d(ctx, dpos.as<void>(), po.data(), sizeof(uint64_t) * po.size()));
    Check(cad_memcpy_h2d(ctx, db.as<void>(), b.data(), sizeof(double) * n));
    cad_matrix* m = nullptr;
    Check(cad_matrix_from_positions(p, dpos.as<uint64_t>(), int64_t(n), &m));
    f.mat.reset(m);
    cad_solve_options o;
    cad_solve_options_default(&o);
    o.solver = CAD_SOLVER_PCG;
    o.precision = CAD_PREC_FP32;
    o.max_iterations = iterations;
    o.tol = 0.0;
    o.check_every = std::max(1, iterations);
    cad_solve_report rep;
    Check(cad_solve(m, db.as<double>(), dx.as<double>(), &o, &rep));
    std::vector<double> xd(std::size_t(f.cols));
    Check(cad_memcpy_d2h(ctx, xd.data(), dx.as<void>(), sizeof(double) * xd.size()));
    for (std::size_t c = 0; c < xd.size(); c++) x[c] = float(xd[c]);
    return x;
}

}  // namespace

Vector FitWeights(const std::vector<uint64_t>& pattern, std::span<ScoredPosition> scored_pos, int iterations)
{
    return fit(pattern, scored_pos.data(), scored_pos.size(), iterations);
}

void FitStages(PatternBasedEstimator& estimator, std::span<const ScoredPosition> scored_pos, int iterations)
{
    const int s = estimator.StageSize();
    // the input masks of every stage are the same as the first stage's
    const auto pattern = estimator.Pattern();
    for (int stage = 0; stage < estimator.Stages(); stage++) {
        std::vector<ScoredPosition> part;
        for (const ScoredPosition& sp : scored_pos) {
            const int e = sp.EmptyCount();
            if (stage * s <= e && e <= stage * s + s - 1) part.push_back(sp);
        }
        Vector w = fit(pattern, part.data(), part.size(), iterations);
        estimator.score.estimators[stage] = ScoreEstimator(pattern, w);
    }
}
