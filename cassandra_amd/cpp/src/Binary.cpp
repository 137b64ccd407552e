// Weight files (the on-disk format of IO/Binary.cpp:3-68, laid out in
// IO/Binary.h): written and read field by field by one writer and one reader,
// which the reference's Serialize / Deserialize entry points call.
#include "IO/Binary.h"

namespace {

// Every field is a little-endian scalar or a counted run of scalars (u64
// count first); an estimator is its masks then its compact weights.
class WeightWriter {
public:
    explicit WeightWriter(std::ostream& s) : s_(s) {}

    template <class T>
    void scalar(T v)
    {
        Serialize(v, s_);
    }
    template <class T>
    void run(const std::vector<T>& v)
    {
        scalar(static_cast<std::uint64_t>(v.size()));
        if (!v.empty()) {
            s_.write(reinterpret_cast<const char*>(v.data()), std::streamsize(sizeof(T) * v.size()));
            cad_io_detail::check(s_);
        }
    }
    void estimator(const ScoreEstimator& e)
    {
        run(e.Pattern());
        run(e.Weights());
    }
    void stages(const MultiStageScoreEstimator& m)
    {
        scalar(static_cast<std::int32_t>(m.StageSize()));
        run(m.Pattern());
        scalar(static_cast<std::uint64_t>(m.estimators.size()));
        for (const ScoreEstimator& e : m.estimators) estimator(e);
    }
    void accuracy(const AccuracyModel& a) { run(a.ParameterValues()); }

private:
    std::ostream& s_;
};

class WeightReader {
public:
    explicit WeightReader(std::istream& s) : s_(s) {}

    template <class T>
    T scalar()
    {
        return Deserialize<T>(s_);
    }
    template <class T>
    std::vector<T> run()
    {
        std::vector<T> v(scalar<std::uint64_t>());
        if (!v.empty()) {
            s_.read(reinterpret_cast<char*>(v.data()), std::streamsize(sizeof(T) * v.size()));
            cad_io_detail::check(s_);
        }
        return v;
    }
    ScoreEstimator estimator()
    {
        std::vector<std::uint64_t> masks = run<std::uint64_t>();
        const std::vector<float> weights = run<float>();
        return ScoreEstimator(std::move(masks), weights);
    }
    MultiStageScoreEstimator stages()
    {
        const int stage_size = scalar<std::int32_t>();
        MultiStageScoreEstimator m(stage_size, run<std::uint64_t>());
        const auto n = scalar<std::uint64_t>();
        m.estimators.clear();
        m.estimators.reserve(n);
        for (std::uint64_t i = 0; i < n; i++) m.estimators.push_back(estimator());
        return m;
    }
    AccuracyModel accuracy() { return AccuracyModel(run<float>()); }

private:
    std::istream& s_;
};

}  // namespace

void Serialize(const ScoreEstimator& e, std::ostream& stream) { WeightWriter(stream).estimator(e); }
void Serialize(const MultiStageScoreEstimator& m, std::ostream& stream) { WeightWriter(stream).stages(m); }
void Serialize(const AccuracyModel& a, std::ostream& stream) { WeightWriter(stream).accuracy(a); }

void Serialize(const PatternBasedEstimator& e, std::ostream& stream)
{
    WeightWriter w(stream);
    w.stages(e.score);
    w.accuracy(e.accuracy);
}

template <>
ScoreEstimator Deserialize<ScoreEstimator>(std::istream& stream)
{
    return WeightReader(stream).estimator();
}

template <>
MultiStageScoreEstimator Deserialize<MultiStageScoreEstimator>(std::istream& stream)
{
    return WeightReader(stream).stages();
}

template <>
AccuracyModel Deserialize<AccuracyModel>(std::istream& stream)
{
    return WeightReader(stream).accuracy();
}

template <>
PatternBasedEstimator Deserialize<PatternBasedEstimator>(std::istream& stream)
{
    WeightReader r(stream);
    MultiStageScoreEstimator score = r.stages();
    AccuracyModel accuracy = r.accuracy();
    return PatternBasedEstimator(std::move(score), std::move(accuracy));
}

void Save(const PatternBasedEstimator& e, const std::filesystem::path& file) { Serialize(e, file); }

PatternBasedEstimator LoadPatternBasedEstimator(const std::filesystem::path& file)
{
    return Deserialize<PatternBasedEstimator>(file);
}
