// Expanded-table evaluator on the device (Pattern/Evaluator.cpp:11-73).
#pragma once
#include "../../include/cassandra_amd.h"
#include "runtime.hpp"

namespace cad {

// The table layout of a pattern group (uploaded on first use).
const ExpandedLayout& expanded_layout(cad_patterns* p);
// ScoreEstimator(pattern, weights): compact weights -> expanded tables.
void weights_expand(cad_patterns* p, const float* d_w, float* d_tables);
// ScoreEstimator::Weights(): expanded tables -> compact weights (the last
// configuration of each column in Configurations order wins, as in the
// reference; synchronises the stream).
void weights_compact(cad_patterns* p, const float* d_tables, float* d_w);
// ScoreEstimator::Eval for n device positions from expanded tables.
void eval_tables(cad_patterns* p, const float* d_tables, const uint64_t* d_pos, int64_t n, float* d_out);

}  // namespace cad
