// Pattern/Indexer.h interface over the library's pattern groups (host-only
// groups: no device is needed to index positions on the host).
#include "Pattern/Indexer.h"

#include <stdexcept>

#include "device.h"

using namespace cad_cpp;

namespace {

std::shared_ptr<cad_patterns> make_group(const std::vector<uint64_t>& masks)
{
    cad_patterns* p = nullptr;
    const int rc = cad_patterns_create(nullptr, masks.data(), int(masks.size()), &p);
    if (rc != CAD_OK) throw std::runtime_error(cad_last_error());
    return std::shared_ptr<cad_patterns>(p, [](cad_patterns* q) { cad_patterns_destroy(q); });
}

// One pattern: class, transformed mask and slot programs come from the library.
class PatternIndexer final : public Indexer
{
    std::shared_ptr<cad_patterns> group;
    std::vector<uint64_t> variations;

public:
    PatternIndexer(std::shared_ptr<cad_patterns> g, uint64_t mask, int order, int size)
        : Indexer(mask, order), group(std::move(g)), variations(order)
    {
        index_space_size = size;
        Check(cad_patterns_variations(group.get(), variations.data()));
    }

    uint32_t Index(Position pos) const override
    {
        // slot 0 applies no symmetry: it is Index(pos) of the reference classes
        std::vector<uint32_t> row(symmetry_group_order);
        uint64_t po[2] = {pos.Player(), pos.Opponent()};
        Check(cad_patterns_indices_host(group.get(), po, 1, row.data()));
        return row[0];
    }

    void InsertIndices(Position pos, std::span<uint32_t> location, uint32_t offset) const override
    {
        uint64_t po[2] = {pos.Player(), pos.Opponent()};
        Check(cad_patterns_indices_host(group.get(), po, 1, location.data()));
        for (int i = 0; i < symmetry_group_order; i++) location[i] += offset;
    }

    std::vector<uint64_t> Variations() const override { return variations; }
};

}  // namespace

Indexer::Indexer(uint64_t pattern, int symmetry_group_order) noexcept
    : pattern(pattern), symmetry_group_order(symmetry_group_order)
{
}

std::vector<uint32_t> Indexer::Indices(Position pos) const
{
    std::vector<uint32_t> out(symmetry_group_order, 0);
    InsertIndices(pos, out);
    return out;
}

std::unique_ptr<Indexer> CreateIndexer(uint64_t pattern)
{
    auto g = make_group({pattern});
    int cls = 0, order = 0;
    uint64_t mask = 0;
    int64_t size = 0, off = 0;
    Check(cad_patterns_pattern_info(g.get(), 0, &cls, &mask, &size, &order, &off));
    return std::make_unique<PatternIndexer>(g, mask, order, int(size));
}

std::size_t ConfigurationCount(uint64_t pattern) { return std::size_t(CreateIndexer(pattern)->index_space_size); }

std::size_t ConfigurationCount(std::vector<uint64_t> pattern)
{
    std::size_t n = 0;
    for (uint64_t p : pattern) n += ConfigurationCount(p);
    return n;
}

GroupIndexer::GroupIndexer(std::vector<uint64_t> pattern) : group(make_group(pattern))
{
    int64_t cols = 0;
    int npat = 0;
    Check(cad_patterns_info(group.get(), &k, &cols, &npat));
    index_space_size = int(cols);
}

std::vector<uint32_t> GroupIndexer::Indices(Position pos) const
{
    std::vector<uint32_t> out(k);
    InsertIndices(pos, out);
    return out;
}

void GroupIndexer::InsertIndices(Position pos, std::span<uint32_t> location) const
{
    uint64_t po[2] = {pos.Player(), pos.Opponent()};
    Check(cad_patterns_indices_host(group.get(), po, 1, location.data()));
}

std::vector<uint64_t> GroupIndexer::Variations() const
{
    std::vector<uint64_t> out(k);
    Check(cad_patterns_variations(group.get(), out.data()));
    return out;
}

std::vector<uint32_t> GroupIndexer::Indices(std::span<const Position> pos) const
{
    std::vector<uint32_t> out(pos.size() * std::size_t(k));
    Check(cad_patterns_indices_host(group.get(), reinterpret_cast<const uint64_t*>(pos.data()), int64_t(pos.size()),
                                    out.data()));
    return out;
}
