// Drop-in for Search/Estimator.h:7-12: the interface the search uses.
#pragma once
#include "Board/Position.h"

class Estimator
{
public:
    virtual ~Estimator() = default;
    virtual float Score(const Position&) const noexcept = 0;
    virtual float Accuracy(int empty_count, int small_depth, int big_depth) const noexcept = 0;
};
