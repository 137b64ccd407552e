// FitWeights (PatternFit/PatternFit.cpp:7-39) and the per-stage fit of the
// stage loop (PatternFit/PatternFit.cpp:40-84) on the GPU: positions are
// indexed by the HIP kernels straight into the device matrix, and the stages
// are fitted together as one block-diagonal batch.
#include "PatternFit/PatternFit.h"

#include <algorithm>
#include <memory>
#include <stdexcept>

#include "Pattern/Indexer.h"

#include "device.h"

using namespace cad_cpp;

namespace {

struct Fit {
    std::unique_ptr<cad_patterns, int (*)(cad_patterns*)> pat{nullptr, cad_patterns_destroy};
    std::unique_ptr<cad_matrix, int (*)(cad_matrix*)> mat{nullptr, cad_matrix_destroy};
    int64_t cols = 0;
};

// Jacobi-PCG in float with FP64 scalars, `iterations` steps from 0, as
// FitWeights runs it (PatternFit/PatternFit.cpp:29-38).
cad_solve_options pcg_options(int iterations)
{
    cad_solve_options o;
    cad_solve_options_default(&o);
    o.solver = CAD_SOLVER_PCG;
    o.precision = CAD_PREC_FP32;
    o.max_iterations = iterations;
    o.tol = 0.0;
    o.check_every = std::max(1, iterations);
    return o;
}

// Fits the systems given by consecutive runs of positions: run s is
// [row_begin[s], row_begin[s+1]) of sp.  One device matrix holds all runs
// (block-diagonal) and one solve fits them together; returns the weights of
// each run.  b is the Score (disc-difference halves) as FitWeights feeds it.
std::vector<Vector> fit_runs(const std::vector<uint64_t>& pattern, const std::vector<const ScoredPosition*>& sp,
                             const std::vector<int64_t>& row_begin, int iterations)
{
    cad_context* ctx = Context();
    Fit f;
    cad_patterns* p = nullptr;
    Check(cad_patterns_create(ctx, pattern.data(), int(pattern.size()), &p));
    f.pat.reset(p);
    int k = 0, np = 0;
    Check(cad_patterns_info(p, &k, &f.cols, &np));
    const int nsys = int(row_begin.size()) - 1;
    const std::size_t n = sp.size();
    std::vector<uint64_t> po(2 * n);
    std::vector<double> b(n);
    for (std::size_t i = 0; i < n; i++) {
        po[2 * i] = sp[i]->pos.Player();
        po[2 * i + 1] = sp[i]->pos.Opponent();
        b[i] = double(int8_t(sp[i]->score));
    }
    const std::size_t ncols = std::size_t(f.cols) * std::size_t(nsys);
    DevBuf dpos(sizeof(uint64_t) * po.size()), db(sizeof(double) * n), dx(sizeof(double) * ncols);
    Check(cad_memcpy_h2d(ctx, dpos.as<void>(), po.data(), sizeof(uint64_t) * po.size()));
    Check(cad_memcpy_h2d(ctx, db.as<void>(), b.data(), sizeof(double) * n));
    cad_matrix* m = nullptr;
    Check(cad_matrix_from_positions_systems(p, dpos.as<uint64_t>(), row_begin.data(), nsys, &m));
    f.mat.reset(m);
    const cad_solve_options o = pcg_options(iterations);
    cad_solve_report rep;
    Check(cad_solve_systems(m, db.as<double>(), dx.as<double>(), &o, &rep, nullptr, nullptr));
    std::vector<double> xd(ncols);
    Check(cad_memcpy_d2h(ctx, xd.data(), dx.as<void>(), sizeof(double) * ncols));
    std::vector<Vector> out(nsys, Vector(std::size_t(f.cols)));
    for (int s = 0; s < nsys; s++)
        for (int64_t c = 0; c < f.cols; c++) out[s][c] = float(xd[std::size_t(s) * f.cols + c]);
    return out;
}

}  // namespace

Vector FitWeights(const std::vector<uint64_t>& pattern, std::span<ScoredPosition> scored_pos, int iterations)
{
    if (scored_pos.empty()) return Vector(std::size_t(ConfigurationCount(pattern)), 0.0f);
    std::vector<const ScoredPosition*> sp;
    for (const ScoredPosition& p : scored_pos) sp.push_back(&p);
    return fit_runs(pattern, sp, {0, int64_t(sp.size())}, iterations)[0];
}

void FitStages(PatternBasedEstimator& estimator, std::span<const ScoredPosition> scored_pos, int iterations)
{
    const int size = estimator.StageSize();
    const int stages = estimator.Stages();
    // the input masks of every stage are the same as the first stage's
    const auto pattern = estimator.Pattern();
    // positions grouped by stage (stable), one system per non-empty stage
    std::vector<std::vector<const ScoredPosition*>> by_stage(stages);
    for (const ScoredPosition& sp : scored_pos) {
        const int stage = sp.EmptyCount() / size;
        if (stage >= 0 && stage < stages) by_stage[stage].push_back(&sp);
    }
    std::vector<const ScoredPosition*> all;
    std::vector<int64_t> row_begin{0};
    std::vector<int> fitted;
    for (int stage = 0; stage < stages; stage++) {
        if (by_stage[stage].empty()) continue;
        all.insert(all.end(), by_stage[stage].begin(), by_stage[stage].end());
        row_begin.push_back(int64_t(all.size()));
        fitted.push_back(stage);
    }
    const Vector zero(std::size_t(ConfigurationCount(pattern)), 0.0f);
    for (int stage = 0; stage < stages; stage++) estimator.score.estimators[stage] = ScoreEstimator(pattern, zero);
    if (fitted.empty()) return;
    const std::vector<Vector> w = fit_runs(pattern, all, row_begin, iterations);
    for (std::size_t i = 0; i < fitted.size(); i++)
        estimator.score.estimators[fitted[i]] = ScoreEstimator(pattern, w[i]);
}
