// Pattern/Evaluator.cpp:6-140 semantics; EvalBatch on the GPU.
#include "Pattern/Evaluator.h"

#include <cmath>
#include <mutex>
#include <stdexcept>

#include "Pattern/Helpers.h"
#include "Pattern/Indexer.h"
#include "device.h"

using namespace cad_cpp;

std::size_t Stages(int stage_size) { return static_cast<std::size_t>(std::ceil(65.0 / stage_size)); }

ScoreEstimator::ScoreEstimator(std::vector<uint64_t> patterns, std::span<const float> weights)
    : input_pattern(patterns), compact(weights.begin(), weights.end())
{
    for (uint64_t p : patterns) {
        auto ix = CreateIndexer(p);
        const auto vars = ix->Variations();
        const int n3 = pown(3, unsigned(std::popcount(p)));
        pattern.push_back(ix->pattern);
        if (weights.size() < std::size_t(ix->index_space_size)) throw std::runtime_error("Size mismatch.");
        for (std::size_t i = 0; i < vars.size(); i++) {
            masks.push_back(vars[i]);
            w.emplace_back(n3, 0.0f);
            for (Position cfg : Configurations(vars[i]))
                w.back()[FastIndex(cfg, vars[i])] = weights[ix->Indices(cfg)[i]];
        }
        weights = weights.subspan(ix->index_space_size);
    }
}

ScoreEstimator::ScoreEstimator(std::vector<uint64_t> patterns) : input_pattern(patterns)
{
    std::size_t cols = 0;
    for (uint64_t p : patterns) {
        auto ix = CreateIndexer(p);
        const int n3 = pown(3, unsigned(std::popcount(p)));
        pattern.push_back(ix->pattern);
        for (uint64_t v : ix->Variations()) {
            masks.push_back(v);
            w.emplace_back(n3, 0.0f);
        }
        cols += std::size_t(ix->index_space_size);
    }
    compact.assign(cols, 0.0f);
}

float ScoreEstimator::Eval(const Position& pos) const noexcept
{
    float s = 0.0f;
    for (std::size_t i = 0; i < masks.size(); i++) s += w[i][FastIndex(pos, masks[i])];
    return s;
}

std::vector<float> ScoreEstimator::Weights() const
{
    std::vector<float> out;
    std::size_t t = 0;
    for (uint64_t p : pattern) {
        auto ix = CreateIndexer(p);
        std::vector<float> dense(std::size_t(ix->index_space_size), 0.0f);
        for (Position cfg : Configurations(p)) dense[ix->Index(cfg)] = w[t][FastIndex(cfg, p)];
        out.insert(out.end(), dense.begin(), dense.end());
        t += ix->Variations().size();
    }
    return out;
}

// The pattern group and the expanded tables of one ScoreEstimator on the
// device; calls on one estimator take turns (its staging buffers are shared).
struct ScoreEstimatorDevice {
    std::mutex mu;
    cad_patterns* patterns = nullptr;
    cad_evaluator* eval = nullptr;
    ~ScoreEstimatorDevice()
    {
        if (eval) cad_evaluator_destroy(eval);
        if (patterns) cad_patterns_destroy(patterns);
    }
};

std::vector<float> ScoreEstimator::EvalBatch(std::span<const Position> pos) const
{
    std::vector<float> out(pos.size());
    if (pos.empty()) return out;
    std::shared_ptr<ScoreEstimatorDevice> d;
    {
        static std::mutex create;
        std::lock_guard<std::mutex> lock(create);
        if (!device) {
            auto nd = std::make_shared<ScoreEstimatorDevice>();
            cad_context* ctx = Context();
            Check(cad_patterns_create(ctx, input_pattern.data(), int(input_pattern.size()), &nd->patterns));
            int k = 0, np = 0;
            int64_t cols = 0;
            Check(cad_patterns_info(nd->patterns, &k, &cols, &np));
            if (compact.size() < std::size_t(cols)) throw std::runtime_error("Size mismatch.");
            DevBuf dw(sizeof(float) * std::size_t(cols));
            Check(cad_memcpy_h2d(ctx, dw.as<void>(), compact.data(), sizeof(float) * std::size_t(cols)));
            Check(cad_evaluator_create(nd->patterns, dw.as<float>(), &nd->eval));
            Check(cad_context_synchronize(ctx));  // the expansion has read dw
            device = std::move(nd);
        }
        d = device;
    }
    std::lock_guard<std::mutex> lock(d->mu);
    static_assert(sizeof(Position) == 16, "Position is {P, O}");
    Check(cad_evaluator_eval_host(d->eval, reinterpret_cast<const uint64_t*>(pos.data()), int64_t(pos.size()),
                                  out.data()));
    return out;
}

MultiStageScoreEstimator::MultiStageScoreEstimator(int stage_size, std::vector<uint64_t> pattern)
    : stage_size(stage_size)
{
    const int stages = int(Stages(stage_size));
    for (int i = 0; i < stages; i++) estimators.emplace_back(pattern);
}

float MultiStageScoreEstimator::Eval(const Position& pos) const noexcept
{
    return estimators[pos.EmptyCount() / stage_size].Eval(pos);
}

float AccuracyModel::Eval(int D, int d, int E) const noexcept
{
    const float alpha = param_values[0], beta = param_values[1], gamma = param_values[2], delta = param_values[3],
                epsilon = param_values[4];
    return (std::exp(alpha * d) + beta) * std::pow(static_cast<float>(D - d), gamma) * (delta * E + epsilon);
}

PatternBasedEstimator::PatternBasedEstimator(MultiStageScoreEstimator s, AccuracyModel a)
    : score(std::move(s)), accuracy(std::move(a))
{
}

PatternBasedEstimator::PatternBasedEstimator(int stage_size, std::vector<uint64_t> pattern)
    : score(stage_size, std::move(pattern))
{
}
