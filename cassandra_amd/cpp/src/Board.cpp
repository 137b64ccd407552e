// Board/BitBoard.cpp:4-104 and the move generator through the library's host
// board code (the code the GPU kernels compile).
#include "Board/Position.h"

#include "device.h"

uint64_t FlippedCodiagonal(uint64_t b) noexcept { return cad_flip(4, b); }
uint64_t FlippedDiagonal(uint64_t b) noexcept { return cad_flip(3, b); }
uint64_t FlippedHorizontal(uint64_t b) noexcept { return cad_flip(2, b); }
uint64_t FlippedVertical(uint64_t b) noexcept { return cad_flip(1, b); }

Position FlippedCodiagonal(const Position& p) noexcept { return {FlippedCodiagonal(p.Player()), FlippedCodiagonal(p.Opponent())}; }
Position FlippedDiagonal(const Position& p) noexcept { return {FlippedDiagonal(p.Player()), FlippedDiagonal(p.Opponent())}; }
Position FlippedHorizontal(const Position& p) noexcept { return {FlippedHorizontal(p.Player()), FlippedHorizontal(p.Opponent())}; }
Position FlippedVertical(const Position& p) noexcept { return {FlippedVertical(p.Player()), FlippedVertical(p.Opponent())}; }

bool IsCodiagonallySymmetric(uint64_t b) noexcept { return FlippedCodiagonal(b) == b; }
bool IsDiagonallySymmetric(uint64_t b) noexcept { return FlippedDiagonal(b) == b; }
bool IsHorizontallySymmetric(uint64_t b) noexcept { return FlippedHorizontal(b) == b; }
bool IsVerticallySymmetric(uint64_t b) noexcept { return FlippedVertical(b) == b; }

uint64_t PossibleMoves(const Position& p) noexcept { return cad_possible_moves(p.Player(), p.Opponent()); }
uint64_t Flips(const Position& p, int move) noexcept { return cad_flips(p.Player(), p.Opponent(), move); }
