// Drop-in for PatternFit/PatternFit.cpp:7-39 (FitWeights, which the reference
// does not declare in its header) plus the per-stage fit of its stage loop
// EvaluateIteratively (PatternFit/PatternFit.cpp:40-84) without the search
// that rescores the positions (with every position already scored and
// reevaluate = false the loop is exactly FitStages).  Both run on the GPU.
#pragma once
#include <span>
#include <vector>

#include "Game/ScoredPosition.h"
#include "Math/Vector.h"
#include "Pattern/Evaluator.h"

// Jacobi-PCG on A^T A from 0 for `iterations` iterations in float, as the
// reference; returns one weight per column of GroupIndexer(pattern).
Vector FitWeights(const std::vector<uint64_t>& pattern, std::span<ScoredPosition> scored_pos, int iterations);

// Fits every empty-count stage of a MultiStageScoreEstimator (positions with
// EmptyCount in [stage*stage_size, stage*stage_size + stage_size - 1]) with
// FitWeights' solver, all stages in one batched solve (per-stage scalars, so
// each stage gets the fit FitWeights gives it alone).  A stage without
// positions gets zero weights.
void FitStages(PatternBasedEstimator& estimator, std::span<const ScoredPosition> scored_pos, int iterations);
