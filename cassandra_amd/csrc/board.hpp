// Bitboard primitives shared by host and device code.
//
// A position is two 64-bit boards {P, O}: P = discs of the side to move,
// O = discs of the opponent (Board/Position.h:11-33).  Bit b is the square
// (x, y) with b = x + 8y (Board/Field.h:7-18).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CAD_HD __host__ __device__ inline
#else
#define CAD_HD inline
#endif

namespace cad {

struct Pos {
    uint64_t P, O;
};

CAD_HD int popcount64(uint64_t b) { return __builtin_popcountll(b); }
CAD_HD int ctz64(uint64_t b) { return __builtin_ctzll(b); }

// ---- symmetries (Board/BitBoard.cpp:4-84) ----------------------------------
// vertical   (x,y) -> (x,7-y)   : byte swap
// horizontal (x,y) -> (7-x,y)   : reverse the bits inside every byte
// diagonal   (x,y) -> (y,x)     : 8x8 bit-matrix transpose (three block swaps)
// codiagonal (x,y) -> (7-y,7-x) : transpose followed by a 180 degree turn
CAD_HD uint64_t flip_vertical(uint64_t b) { return __builtin_bswap64(b); }
CAD_HD uint64_t flip_horizontal(uint64_t b)
{
    // full 64-bit reversal turns (x,y) into (7-x,7-y); the byte swap then
    // restores the rank.
    return __builtin_bswap64(__builtin_bitreverse64(b));
}
CAD_HD uint64_t flip_diagonal(uint64_t b)
{
    // Swap the off-diagonal 4x4, 2x2 and 1x1 blocks of the bit matrix.
    // Square (x,y) and (y,x) differ by 7*(x-y) bits.
    uint64_t t;
    t = (b ^ (b >> 28)) & 0x00000000F0F0F0F0ULL;  // 4x4 blocks
    b ^= t ^ (t << 28);
    t = (b ^ (b >> 14)) & 0x0000CCCC0000CCCCULL;  // 2x2 blocks
    b ^= t ^ (t << 14);
    t = (b ^ (b >> 7)) & 0x00AA00AA00AA00AAULL;   // single squares
    b ^= t ^ (t << 7);
    return b;
}
CAD_HD uint64_t flip_codiagonal(uint64_t b) { return __builtin_bitreverse64(flip_diagonal(b)); }

// Symmetry ids used by the index programs.  A composite "F2 then F1" is
// written F1∘F2 and acts on positions as F1(F2(pos)).
enum Sym : int { kId = 0, kV = 1, kH = 2, kD = 3, kC = 4 };

CAD_HD uint64_t apply_sym(int s, uint64_t b)
{
    switch (s) {
    case kV: return flip_vertical(b);
    case kH: return flip_horizontal(b);
    case kD: return flip_diagonal(b);
    case kC: return flip_codiagonal(b);
    default: return b;
    }
}

// The square map of each symmetry (an involution: it is its own inverse).
CAD_HD int sym_square(int s, int q)
{
    int x = q & 7, y = q >> 3;
    switch (s) {
    case kV: return x + 8 * (7 - y);
    case kH: return (7 - x) + 8 * y;
    case kD: return y + 8 * x;
    case kC: return (7 - y) + 8 * (7 - x);
    default: return q;
    }
}

// ---- move generation (Board/PossibleMoves.cpp:86-156, Board/Flips.cpp) -----
// Shift of a whole board one step in direction d (0..7), dropping squares
// that would wrap around a file edge.
CAD_HD uint64_t step(uint64_t b, int d)
{
    constexpr uint64_t notA = 0xFEFEFEFEFEFEFEFEULL;  // x != 0
    constexpr uint64_t notH = 0x7F7F7F7F7F7F7F7FULL;  // x != 7
    switch (d) {
    case 0: return (b << 1) & notA;  // +x
    case 1: return (b >> 1) & notH;  // -x
    case 2: return b << 8;           // +y
    case 3: return b >> 8;           // -y
    case 4: return (b << 9) & notA;  // +x +y
    case 5: return (b << 7) & notH;  // -x +y
    case 6: return (b >> 7) & notA;  // +x -y
    default: return (b >> 9) & notH; // -x -y
    }
}

CAD_HD uint64_t possible_moves(uint64_t P, uint64_t O)
{
    const uint64_t empty = ~(P | O);
    uint64_t moves = 0;
    for (int d = 0; d < 8; d++) {
        uint64_t run = step(P, d) & O;
        for (int i = 0; i < 5; i++) run |= step(run, d) & O;
        moves |= step(run, d) & empty;
    }
    return moves;
}

// Discs flipped by playing square `move` (which must be empty).
CAD_HD uint64_t flips(uint64_t P, uint64_t O, int move)
{
    const uint64_t m = 1ULL << move;
    uint64_t out = 0;
    for (int d = 0; d < 8; d++) {
        uint64_t run = 0, cur = step(m, d);
        while (cur & O) {
            run |= cur;
            cur = step(cur, d);
        }
        if (cur & P) out |= run;
    }
    return out;
}

// Board/Position.cpp:84-89: the mover's discs become the opponent's.
CAD_HD Pos play(Pos p, int move)
{
    const uint64_t f = flips(p.P, p.O, move);
    return Pos{p.O ^ f, p.P ^ f ^ (1ULL << move)};
}

// Game/GameClass.cpp:4-13: pass when the mover cannot move but the opponent can.
CAD_HD Pos implicit_pass(Pos p)
{
    if (!possible_moves(p.P, p.O) && possible_moves(p.O, p.P)) return Pos{p.O, p.P};
    return p;
}

CAD_HD Pos start_position() { return Pos{0x0000000810000000ULL, 0x0000001008000000ULL}; }

// ---- synthetic data streams ------------------------------------------------
// Every game / column / row owns an independent SplitMix64 stream seeded with
// seed ^ (0xD1B54A32D192ED03 * (i + 1)), so results do not depend on how the
// work is split across threads or devices.
struct Stream {
    uint64_t s;
    CAD_HD explicit Stream(uint64_t seed, uint64_t i) : s(seed ^ (0xD1B54A32D192ED03ULL * (i + 1))) {}
    CAD_HD uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // floor(u * n / 2^64): an index in [0, n) from one 64-bit draw.
    CAD_HD uint32_t below(uint32_t n)
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return uint32_t(__umul64hi(next(), uint64_t(n)));
#else
        return uint32_t((static_cast<unsigned __int128>(next()) * n) >> 64);
#endif
    }
    // Approximately standard normal: Irwin-Hall sum of 12 uniforms of 32-bit
    // resolution minus 6.  Every term is a multiple of 2^-32, so the sum is
    // exact in binary64 and identical on every platform.
    CAD_HD double ih12()
    {
        double acc = 0.0;
        for (int i = 0; i < 12; i++) acc += double(uint32_t(next() >> 32)) * (1.0 / 4294967296.0);
        return acc - 6.0;
    }
};

// Uniformly random legal play from the start position (Game/Player.cpp:3-16,
// Game/GameGenerator.cpp:3-28, Game/GameClass.cpp:32-44): the empty count E is
// drawn from [e_lo, e_hi] first (when the range is not a single value), then
// moves are drawn as the r-th set bit (LSB first) of the move set, r = below(n).
// Positions are kept after the implicit pass.  A game that ends before reaching
// E empties restarts from the start position with the stream continuing.
CAD_HD Pos random_play(uint64_t seed, uint64_t game, int e_lo, int e_hi)
{
    Stream rng(seed, game);
    const int E = (e_lo == e_hi) ? e_lo : e_lo + int(rng.below(uint32_t(e_hi - e_lo + 1)));
    Pos p = start_position();
    for (int guard = 0; guard < 1 << 20; guard++) {
        if (popcount64(~(p.P | p.O)) == E) break;
        uint64_t m = possible_moves(p.P, p.O);
        if (!m) {
            p = start_position();
            continue;
        }
        uint32_t r = rng.below(uint32_t(popcount64(m)));
        for (uint32_t i = 0; i < r; i++) m &= m - 1;  // drop r lowest moves
        p = implicit_pass(play(p, ctz64(m)));
    }
    return p;
}

}  // namespace cad
