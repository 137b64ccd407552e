// Drop-in for Pattern/Helpers.h:6-41.
#pragma once
#include <cstddef>
#include <cstdint>
#include <iterator>

#include "Board/Position.h"

int pown(int base, unsigned int exponent);

// sum_i 3^i * (P[s_i] + 2 O[s_i]) over the set bits s_0 < s_1 < ... of the mask
uint32_t FastIndex(Position, uint64_t pattern) noexcept;

uint64_t PDep(uint64_t src, uint64_t mask) noexcept;
uint64_t PExt(uint64_t src, uint64_t mask) noexcept;

// All (p, o) with p & o == 0 over the squares of a mask, p outer and o inner,
// deposited onto the mask (Pattern/Helpers.cpp:81-102).
class Configurations
{
    uint64_t mask;

public:
    class Iterator
    {
        uint64_t mask{}, size{0}, p{0}, o{0};

    public:
        using difference_type = std::ptrdiff_t;
        using value_type = Position;
        using iterator_category = std::forward_iterator_tag;
        Iterator() noexcept = default;
        explicit Iterator(uint64_t mask) noexcept;
        Iterator& operator++();
        Position operator*() const;
        bool operator==(const Iterator&) const noexcept = default;
        bool operator!=(const Iterator&) const noexcept = default;
    };

    explicit Configurations(uint64_t mask) noexcept : mask(mask) {}
    Iterator begin() const noexcept { return Iterator(mask); }
    static Iterator end() noexcept { return {}; }
};
