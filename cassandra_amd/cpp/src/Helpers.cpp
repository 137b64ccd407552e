// Pattern/Helpers.cpp:5-102.
#include "Pattern/Helpers.h"

#include <bit>

int pown(int base, unsigned int exponent)
{
    int r = 1;
    while (exponent--) r *= base;
    return r;
}

uint64_t PExt(uint64_t src, uint64_t mask) noexcept
{
    uint64_t out = 0;
    int j = 0;
    for (uint64_t b = mask; b; b &= b - 1, j++) out |= ((src >> std::countr_zero(b)) & 1ULL) << j;
    return out;
}

uint64_t PDep(uint64_t src, uint64_t mask) noexcept
{
    uint64_t out = 0;
    int j = 0;
    for (uint64_t b = mask; b; b &= b - 1, j++) out |= ((src >> j) & 1ULL) << std::countr_zero(b);
    return out;
}

uint32_t FastIndex(Position pos, uint64_t pattern) noexcept
{
    uint32_t v = 0, w = 1;
    for (uint64_t b = pattern; b; b &= b - 1, w *= 3) {
        const int q = std::countr_zero(b);
        v += w * (uint32_t((pos.Player() >> q) & 1) + 2 * uint32_t((pos.Opponent() >> q) & 1));
    }
    return v;
}

Configurations::Iterator::Iterator(uint64_t mask) noexcept : mask(mask), size(1ULL << std::popcount(mask)) {}

Configurations::Iterator& Configurations::Iterator::operator++()
{
    o++;
    for (; p < size; p++) {
        for (; o < size; o++)
            if ((p & o) == 0) return *this;
        o = 0;
    }
    *this = Iterator();
    return *this;
}

Position Configurations::Iterator::operator*() const { return {PDep(p, mask), PDep(o, mask)}; }
