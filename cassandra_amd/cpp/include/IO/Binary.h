// Drop-in for IO/Binary.h:1-69 and IO/Binary.cpp:3-68: the on-disk format of
// fitted weights.  Little-endian, native sizes (size_t = 8 bytes):
//   vector<T>              u64 n, then n elements
//   ScoreEstimator         vector<u64> pattern (the transformed masks), vector<f32> compact weights
//   MultiStageScoreEstimator  i32 stage_size, vector<u64> pattern, u64 n, n ScoreEstimators
//   AccuracyModel          vector<f32> parameters
//   PatternBasedEstimator  MultiStageScoreEstimator, AccuracyModel
#pragma once
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <type_traits>
#include <vector>

#include "Pattern/Evaluator.h"

namespace cad_io_detail {
template <class, template <class...> class>
inline constexpr bool is_instance = false;
template <template <class...> class T, class... A>
inline constexpr bool is_instance<T<A...>, T> = true;

inline void check(std::ios& s)
{
    if (!s) throw std::runtime_error("Binary stream failed (truncated or unwritable).");
}
}  // namespace cad_io_detail

template <typename T>
T Deserialize(std::istream&);

template <typename T>
void Serialize(const T& t, std::ostream& stream)
    requires std::is_arithmetic_v<T> || std::is_enum_v<T>
{
    stream.write(reinterpret_cast<const char*>(&t), sizeof(T));
    cad_io_detail::check(stream);
}

template <typename T>
T Deserialize(std::istream& stream)
    requires std::is_arithmetic_v<T> || std::is_enum_v<T>
{
    T t{};
    stream.read(reinterpret_cast<char*>(&t), sizeof(T));
    cad_io_detail::check(stream);
    return t;
}

void Serialize(const ScoreEstimator&, std::ostream&);
void Serialize(const MultiStageScoreEstimator&, std::ostream&);
void Serialize(const AccuracyModel&, std::ostream&);
void Serialize(const PatternBasedEstimator&, std::ostream&);

template <typename T>
void Serialize(const std::vector<T>& v, std::ostream& stream)
{
    Serialize(static_cast<std::size_t>(v.size()), stream);
    if constexpr (std::is_arithmetic_v<T>) {
        stream.write(reinterpret_cast<const char*>(v.data()), std::streamsize(sizeof(T) * v.size()));
        cad_io_detail::check(stream);
    } else {
        for (const T& t : v) Serialize(t, stream);
    }
}

template <typename T>
T Deserialize(std::istream& stream)
    requires cad_io_detail::is_instance<T, std::vector>
{
    using E = typename T::value_type;
    const auto n = Deserialize<std::size_t>(stream);
    T v;
    if constexpr (std::is_arithmetic_v<E>) {
        v.resize(n);
        stream.read(reinterpret_cast<char*>(v.data()), std::streamsize(sizeof(E) * n));
        cad_io_detail::check(stream);
    } else {
        v.reserve(n);
        for (std::size_t i = 0; i < n; i++) v.push_back(Deserialize<E>(stream));
    }
    return v;
}

template <>
ScoreEstimator Deserialize<ScoreEstimator>(std::istream&);
template <>
MultiStageScoreEstimator Deserialize<MultiStageScoreEstimator>(std::istream&);
template <>
AccuracyModel Deserialize<AccuracyModel>(std::istream&);
template <>
PatternBasedEstimator Deserialize<PatternBasedEstimator>(std::istream&);

template <typename T>
void Serialize(const T& t, const std::filesystem::path& file)
{
    std::ofstream stream(file, std::ios::binary | std::ios::out | std::ios::trunc);
    if (!stream) throw std::runtime_error("Can't open " + file.string());
    Serialize(t, stream);
}

template <typename T>
T Deserialize(const std::filesystem::path& file)
{
    std::ifstream stream(file, std::ios::binary | std::ios::in);
    if (!stream) throw std::runtime_error("Can't open " + file.string());
    return Deserialize<T>(stream);
}

void Save(const PatternBasedEstimator&, const std::filesystem::path&);
PatternBasedEstimator LoadPatternBasedEstimator(const std::filesystem::path&);
