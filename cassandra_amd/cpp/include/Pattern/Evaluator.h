// Drop-in for Pattern/Evaluator.h:9-68.  ScoreEstimator expands the compact
// weights into per-variation tables and sums them per position, as the
// reference does; EvalBatch scores many positions on the GPU from the same
// tables, expanded once on the device and kept there (same terms, same order:
// the same floats).
#pragma once
#include <cstdint>
#include <memory>
#include <span>
#include <vector>

#include "Board/Position.h"
#include "Search/Estimator.h"

std::size_t Stages(int stage_size);

class ScoreEstimator
{
    std::vector<uint64_t> masks;              // one per variation
    std::vector<std::vector<float>> w;        // expanded weights
    std::vector<uint64_t> pattern;            // transformed masks
    std::vector<uint64_t> input_pattern;      // masks as given
    std::vector<float> compact;               // compact weights (for EvalBatch)
    // the device evaluator of EvalBatch, created on its first call and shared by copies
    mutable std::shared_ptr<struct ScoreEstimatorDevice> device;

public:
    ScoreEstimator(std::vector<uint64_t> pattern, std::span<const float> weights);
    explicit ScoreEstimator(std::vector<uint64_t> pattern);

    float Eval(const Position&) const noexcept;
    std::vector<float> EvalBatch(std::span<const Position>) const;  // GPU

    std::vector<uint64_t> Pattern() const noexcept { return pattern; }
    std::vector<float> Weights() const;
};

class MultiStageScoreEstimator
{
    int stage_size;

public:
    std::vector<ScoreEstimator> estimators;

    MultiStageScoreEstimator(int stage_size, std::vector<uint64_t> pattern);

    int StageSize() const noexcept { return stage_size; }
    std::vector<uint64_t> Pattern() const noexcept { return estimators.front().Pattern(); }

    float Eval(const Position&) const noexcept;
};

class AccuracyModel
{
    std::vector<float> param_values;

public:
    AccuracyModel(std::vector<float> param_values = {-0.16594754f, 0.95554334f, 0.27973529f, -0.01616647f,
                                                     1.28078789f}) noexcept
        : param_values(std::move(param_values)) {}
    const std::vector<float>& ParameterValues() const { return param_values; }
    float Eval(int D, int d, int E) const noexcept;
};

class PatternBasedEstimator final : public Estimator
{
public:
    MultiStageScoreEstimator score;
    AccuracyModel accuracy;

    PatternBasedEstimator(MultiStageScoreEstimator, AccuracyModel);
    PatternBasedEstimator(int stage_size, std::vector<uint64_t> pattern);

    int Stages() const noexcept { return static_cast<int>(score.estimators.size()); }
    int StageSize() const noexcept { return score.StageSize(); }
    std::vector<uint64_t> Pattern() const noexcept { return score.Pattern(); }

    float Score(const Position& pos) const noexcept override { return score.Eval(pos); }
    float Accuracy(int empty_count, int small_depth, int big_depth) const noexcept override
    {
        return accuracy.Eval(big_depth, small_depth, empty_count);
    }
};
