// Host-side pattern catalogue: symmetry classification of a pattern mask,
// index-space layout of a pattern group, and the "index programs" the device
// kernels run to compute a row of column indices from a 16-byte position.
//
// Semantics follow Pattern/Indexer.cpp:18-422 (CreateIndexer, the six indexer
// classes, GroupIndexer) and Pattern/Helpers.cpp:5-102 (FastIndex,
// Configurations).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cad {

enum PatternClass : int { kA = 0, kVcls = 1, kDcls = 2, kVH = 3, kDC = 4, kVHDC = 5 };

// One list of squares whose digits form a base-3 number:
//   value = sum_j 3^j * (P[src_j] + 2 O[src_j]).
// src_j are squares of the *untransformed* position: the symmetry of the slot
// is folded into the square numbers.
struct SquareList {
    std::vector<uint8_t> src;
};

// How one column index of a row is computed (one "slot" = one element of the
// row, Pattern/Indexer.cpp InsertIndices).
struct SlotProgram {
    int kind = 0;              // 0: a, 1: pair(a,b), 2: pair(a,b)*ds + d, 3: dense[a]
    SquareList list[3];        // a, b, d
    uint32_t hs = 0, ds = 0;   // pairing sizes (V / D classes)
    uint32_t col_offset = 0;   // first column of the pattern's block
    uint32_t dense_offset = 0; // into PatternGroup::dense (kind 3)
    uint32_t block_size = 0;   // index_space_size of the pattern
};

struct PatternInfo {
    uint64_t input_mask = 0;
    uint64_t mask = 0;         // transformed mask the indexer keeps (Indexer::pattern)
    PatternClass cls = kA;
    int order = 0;             // symmetry_group_order
    uint32_t size = 0;         // index_space_size
    uint32_t offset = 0;       // first column in the group
    std::vector<uint32_t> dense;      // VH / DC / VHDC orbit numbering (3^n entries)
    std::vector<uint64_t> variations; // Indexer::Variations()
};

class PatternGroup {
public:
    explicit PatternGroup(const std::vector<uint64_t>& masks);

    int k() const { return int(slots.size()); }
    uint64_t cols() const { return cols_; }

    std::vector<PatternInfo> patterns;
    std::vector<SlotProgram> slots;
    std::vector<uint32_t> dense;  // concatenated dense tables

    // Host evaluation of one row (used to validate the device programs and by
    // the C++ API's Indices()).
    void row(uint64_t P, uint64_t O, uint32_t* out) const;

private:
    uint64_t cols_ = 0;
};

// Pattern/Helpers.cpp building blocks (host).
uint32_t fast_index(uint64_t P, uint64_t O, uint64_t mask);
uint64_t pdep(uint64_t src, uint64_t mask);
uint64_t pext(uint64_t src, uint64_t mask);
uint32_t pow3(int n);

// Classification and per-pattern info (throws std::runtime_error on failure).
PatternInfo make_pattern_info(uint64_t mask);

}  // namespace cad
