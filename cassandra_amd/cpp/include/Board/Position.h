// Drop-in for the parts of Board/Position.h and Board/BitBoard.h the fit uses.
// Bit b is the square (x, y) with b = x + 8y; P holds the discs of the side to
// move (Board/Position.h:11-33).
#pragma once
#include <bit>
#include <cstdint>

class Position
{
    uint64_t P{}, O{};

public:
    Position() noexcept = default;
    constexpr Position(uint64_t P, uint64_t O) noexcept : P(P), O(O) {}

    static Position Start() noexcept { return {0x0000000810000000ULL, 0x0000001008000000ULL}; }

    bool operator==(const Position& o) const noexcept { return P == o.P && O == o.O; }
    bool operator!=(const Position& o) const noexcept { return !(*this == o); }
    bool operator<(const Position& o) const noexcept { return P < o.P || (P == o.P && O < o.O); }

    uint64_t Player() const noexcept { return P; }
    uint64_t Opponent() const noexcept { return O; }
    uint64_t Discs() const noexcept { return P | O; }
    uint64_t Empties() const noexcept { return ~Discs(); }
    int EmptyCount() const noexcept { return std::popcount(Empties()); }
};
static_assert(sizeof(Position) == 16, "positions are 16 bytes on the device");

uint64_t FlippedCodiagonal(uint64_t) noexcept;
uint64_t FlippedDiagonal(uint64_t) noexcept;
uint64_t FlippedHorizontal(uint64_t) noexcept;
uint64_t FlippedVertical(uint64_t) noexcept;
Position FlippedCodiagonal(const Position&) noexcept;
Position FlippedDiagonal(const Position&) noexcept;
Position FlippedHorizontal(const Position&) noexcept;
Position FlippedVertical(const Position&) noexcept;
bool IsCodiagonallySymmetric(uint64_t) noexcept;
bool IsDiagonallySymmetric(uint64_t) noexcept;
bool IsHorizontallySymmetric(uint64_t) noexcept;
bool IsVerticallySymmetric(uint64_t) noexcept;

uint64_t PossibleMoves(const Position&) noexcept;
uint64_t Flips(const Position&, int move) noexcept;
