// Drop-in for Game/Score.h and Game/ScoredPosition.h: Score is half the disc
// difference in an int8 (text formats hold the disc difference).
#pragma once
#include <cstdint>

#include "Board/Position.h"

class Score
{
    int8_t value{};

public:
    constexpr Score() noexcept = default;
    constexpr Score(int v) noexcept : value(int8_t(v)) {}
    constexpr operator int8_t() const { return value; }
    Score operator-() const noexcept { return -value; }
};

inline constexpr Score min_score{-32};
inline constexpr Score max_score{+32};
inline constexpr Score inf_score{+33};
inline constexpr Score undefined_score{+35};

struct ScoredPosition
{
    Position pos;
    Score score = undefined_score;

    int EmptyCount() const noexcept { return pos.EmptyCount(); }
    bool HasScore() const noexcept { return score != undefined_score; }
};
