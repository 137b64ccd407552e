// The weight file format (IO/Binary.cpp:3-68).
#include "IO/Binary.h"

void Serialize(const ScoreEstimator& e, std::ostream& stream)
{
    Serialize(e.Pattern(), stream);
    Serialize(e.Weights(), stream);
}

template <>
ScoreEstimator Deserialize<ScoreEstimator>(std::istream& stream)
{
    auto pattern = Deserialize<std::vector<uint64_t>>(stream);
    auto weights = Deserialize<std::vector<float>>(stream);
    return ScoreEstimator(std::move(pattern), weights);
}

void Serialize(const MultiStageScoreEstimator& m, std::ostream& stream)
{
    Serialize(m.StageSize(), stream);
    Serialize(m.Pattern(), stream);
    Serialize(m.estimators, stream);
}

template <>
MultiStageScoreEstimator Deserialize<MultiStageScoreEstimator>(std::istream& stream)
{
    const int stage_size = Deserialize<int>(stream);
    auto pattern = Deserialize<std::vector<uint64_t>>(stream);
    MultiStageScoreEstimator m(stage_size, std::move(pattern));
    m.estimators = Deserialize<std::vector<ScoreEstimator>>(stream);
    return m;
}

void Serialize(const AccuracyModel& a, std::ostream& stream) { Serialize(a.ParameterValues(), stream); }

template <>
AccuracyModel Deserialize<AccuracyModel>(std::istream& stream)
{
    return AccuracyModel(Deserialize<std::vector<float>>(stream));
}

void Serialize(const PatternBasedEstimator& e, std::ostream& stream)
{
    Serialize(e.score, stream);
    Serialize(e.accuracy, stream);
}

template <>
PatternBasedEstimator Deserialize<PatternBasedEstimator>(std::istream& stream)
{
    auto score = Deserialize<MultiStageScoreEstimator>(stream);
    auto accuracy = Deserialize<AccuracyModel>(stream);
    return PatternBasedEstimator(std::move(score), std::move(accuracy));
}

void Save(const PatternBasedEstimator& e, const std::filesystem::path& file) { Serialize(e, file); }

PatternBasedEstimator LoadPatternBasedEstimator(const std::filesystem::path& file)
{
    return Deserialize<PatternBasedEstimator>(file);
}
