// Drop-in for Pattern/Names.h: the pattern masks (bit b = x + 8y) and sets.
#pragma once
#include <cstdint>
#include <vector>

namespace pattern
{
inline constexpr uint64_t L0 = 0x00000000000000FFULL;
inline constexpr uint64_t L1 = 0x000000000000FF00ULL;
inline constexpr uint64_t L2 = 0x0000000000FF0000ULL;
inline constexpr uint64_t L3 = 0x00000000FF000000ULL;
inline constexpr uint64_t D4 = 0x0000000001020408ULL;
inline constexpr uint64_t D5 = 0x0000000102040810ULL;
inline constexpr uint64_t D6 = 0x0000010204081020ULL;
inline constexpr uint64_t D7 = 0x0001020408102040ULL;
inline constexpr uint64_t D8 = 0x0102040810204080ULL;
inline constexpr uint64_t L02X = 0x00000000000042FFULL;
inline constexpr uint64_t B4 = 0x0000000000000F0FULL;
inline constexpr uint64_t B5 = 0x0000000000001F1FULL;
inline constexpr uint64_t B6 = 0x0000000000003F3FULL;
inline constexpr uint64_t Q3 = 0x0000000000070707ULL;
inline constexpr uint64_t C3 = 0x0000000000010307ULL;
inline constexpr uint64_t C3p1 = 0x000000000101030FULL;
inline constexpr uint64_t C3p2 = 0x000000010101031FULL;
inline constexpr uint64_t Ep = 0x0000000000003CBDULL;
inline constexpr uint64_t Comet = 0x8040201008040303ULL;

inline const std::vector<uint64_t> logistello{D4, D5, D6, D7, D8, L02X, L1, L2, L3, Q3, B5};
inline const std::vector<uint64_t> edax{D4, D5, D6, D7, D8, L02X, L1, L2, L3, Q3, C3p2, Ep};
inline const std::vector<uint64_t> cassandra{D4, D5, D6, D7, Comet, L02X, L1, L2, L3, Q3, C3p2, B5};
}  // namespace pattern
