// Drop-in for Pattern/Indexer.h:10-40.  The symmetry classification, index
// formulas and orbit numbering are libcassandra_amd's (the same tables the GPU
// index kernels run); the classes keep the reference interface.
#pragma once
#include <cstdint>
#include <memory>
#include <span>
#include <vector>

#include "Board/Position.h"

struct cad_patterns;

struct Indexer
{
    uint64_t pattern;          // the transformed mask the indexer works on
    int symmetry_group_order;  // indices per position
    int index_space_size = 0;

    Indexer(uint64_t pattern, int symmetry_group_order) noexcept;
    virtual ~Indexer() = default;

    virtual uint32_t Index(Position) const = 0;
    std::vector<uint32_t> Indices(Position) const;
    virtual void InsertIndices(Position, std::span<uint32_t> location, uint32_t offset = 0) const = 0;
    virtual std::vector<uint64_t> Variations() const = 0;
};

// Throws std::runtime_error for masks that cannot be indexed (more than 16 squares).
std::unique_ptr<Indexer> CreateIndexer(uint64_t pattern);

std::size_t ConfigurationCount(uint64_t pattern);
std::size_t ConfigurationCount(std::vector<uint64_t> pattern);

class GroupIndexer
{
    std::shared_ptr<cad_patterns> group;
    int k = 0;

public:
    int index_space_size = 0;

    explicit GroupIndexer(std::vector<uint64_t> pattern);

    std::vector<uint32_t> Indices(Position) const;
    void InsertIndices(Position, std::span<uint32_t> location) const;
    std::vector<uint64_t> Variations() const;

    // every row of a batch at once (row-major n x Variations().size())
    std::vector<uint32_t> Indices(std::span<const Position>) const;
};
