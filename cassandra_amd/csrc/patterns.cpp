#include "patterns.hpp"

#include <algorithm>
#include <stdexcept>

#include "board.hpp"

namespace cad {

namespace {
constexpr uint64_t kHalfV = 0x00000000FFFFFFFFULL;  // rows 0-3 (Pattern/Indexer.cpp:64)
constexpr uint64_t kHalfD = 0x0080C0E0F0F8FCFEULL;  // x > y  (Pattern/Indexer.cpp:107)
constexpr uint64_t kDiagL = 0x8040201008040201ULL;  // x == y (Pattern/Indexer.cpp:108)

bool symmetric(int s, uint64_t m) { return apply_sym(s, m) == m; }

Pos apply(int s, Pos p) { return Pos{apply_sym(s, p.P), apply_sym(s, p.O)}; }

// Square list of FastIndex(U(pos), m) where U applies ops in order.
SquareList list_for(const std::vector<int>& ops, uint64_t m)
{
    SquareList l;
    for (uint64_t b = m; b; b &= b - 1) {
        int q = ctz64(b);
        for (auto it = ops.rbegin(); it != ops.rend(); ++it) q = sym_square(*it, q);
        l.src.push_back(uint8_t(q));
    }
    return l;
}

uint32_t eval_list(const SquareList& l, uint64_t P, uint64_t O)
{
    uint32_t v = 0;
    for (size_t j = l.src.size(); j-- > 0;) {
        int q = l.src[j];
        v = v * 3 + uint32_t((P >> q) & 1) + 2 * uint32_t((O >> q) & 1);
    }
    return v;
}

// Slot symmetries in InsertIndices order, each as a list of flips in the
// order they are applied to the position.
std::vector<std::vector<int>> slot_transforms(PatternClass c)
{
    switch (c) {
    case kA:  // Pattern/Indexer.cpp:31-41
        return {{}, {kC}, {kD}, {kH}, {kV}, {kC, kH}, {kD, kH}, {kV, kH}};
    case kVcls:  // :82-88
        return {{}, {kC}, {kD}, {kH}};
    case kDcls:  // :128-134
        return {{}, {kC}, {kH}, {kV}};
    case kVH:  // :186-190
        return {{}, {kD}};
    case kDC:  // :240-244
        return {{}, {kV}};
    default:  // :306-309
        return {{}};
    }
}

// Orbit numbering of VH / DC / VHDC patterns: ids in first-seen order of the
// Configurations enumeration, shared by every image of a configuration under
// the pattern's symmetry group (Pattern/Indexer.cpp:159-180, 213-234, 271-300).
std::vector<uint32_t> dense_numbering(PatternClass c, uint64_t m, uint32_t& count)
{
    const int n = popcount64(m);
    std::vector<uint32_t> dense(pow3(n), 0xFFFFFFFFu);
    std::vector<std::vector<int>> group;
    if (c == kVH) group = {{}, {kV}, {kH}, {kV, kH}};
    else if (c == kDC) group = {{}, {kD}, {kC}, {kD, kC}};
    else group = {{}, {kC}, {kD}, {kH}, {kV}, {kH, kC}, {kH, kD}, {kH, kV}};
    uint32_t next = 0;
    const uint64_t span = 1ULL << n;
    for (uint64_t p = 0; p < span; p++)
        for (uint64_t o = 0; o < span; o++) {
            if (p & o) continue;
            Pos cfg{pdep(p, m), pdep(o, m)};
            if (dense[fast_index(cfg.P, cfg.O, m)] != 0xFFFFFFFFu) continue;
            for (auto& ops : group) {
                Pos q = cfg;
                for (int s : ops) q = apply(s, q);
                dense[fast_index(q.P, q.O, m)] = next;
            }
            next++;
        }
    count = next;
    return dense;
}
}  // namespace

uint32_t pow3(int n)
{
    uint32_t r = 1;
    while (n-- > 0) r *= 3;
    return r;
}

uint64_t pext(uint64_t src, uint64_t mask)
{
    uint64_t out = 0;
    int j = 0;
    for (uint64_t b = mask; b; b &= b - 1, j++) out |= ((src >> ctz64(b)) & 1ULL) << j;
    return out;
}

uint64_t pdep(uint64_t src, uint64_t mask)
{
    uint64_t out = 0;
    int j = 0;
    for (uint64_t b = mask; b; b &= b - 1, j++) out |= ((src >> j) & 1ULL) << ctz64(b);
    return out;
}

uint32_t fast_index(uint64_t P, uint64_t O, uint64_t mask)
{
    uint32_t v = 0, w = 1;
    for (uint64_t b = mask; b; b &= b - 1, w *= 3) {
        int q = ctz64(b);
        v += w * (uint32_t((P >> q) & 1) + 2 * uint32_t((O >> q) & 1));
    }
    return v;
}

PatternInfo make_pattern_info(uint64_t p)
{
    PatternInfo info;
    info.input_mask = p;
    if (popcount64(p) > 16) throw std::runtime_error("Pattern has more than 16 squares.");
    const bool v = symmetric(kV, p), h = symmetric(kH, p), d = symmetric(kD, p), c = symmetric(kC, p);
    // Pattern/Indexer.cpp:345-364
    if (v && h && d && c) { info.cls = kVHDC; info.mask = p; }
    else if (v && h) { info.cls = kVH; info.mask = p; }
    else if (d && c) { info.cls = kDC; info.mask = p; }
    else if (v) { info.cls = kVcls; info.mask = p; }
    else if (h) { info.cls = kVcls; info.mask = flip_diagonal(p); }
    else if (d) { info.cls = kDcls; info.mask = p; }
    else if (c) { info.cls = kDcls; info.mask = flip_vertical(p); }
    else { info.cls = kA; info.mask = p; }

    const uint64_t m = info.mask;
    switch (info.cls) {
    case kA:
        info.order = 8;
        info.size = pow3(popcount64(m));
        info.variations = {m, flip_codiagonal(m), flip_diagonal(m), flip_horizontal(m), flip_vertical(m),
                           flip_codiagonal(flip_horizontal(m)), flip_diagonal(flip_horizontal(m)),
                           flip_vertical(flip_horizontal(m))};
        break;
    case kVcls: {
        info.order = 4;
        uint64_t hs = pow3(popcount64(m & kHalfV));
        info.size = uint32_t(hs * (hs + 1) / 2);
        info.variations = {m, flip_codiagonal(m), flip_diagonal(m), flip_horizontal(m)};
        break;
    }
    case kDcls: {
        info.order = 4;
        uint64_t hs = pow3(popcount64(m & kHalfD)), ds = pow3(popcount64(m & kDiagL));
        info.size = uint32_t(ds * hs * (hs + 1) / 2);
        info.variations = {m, flip_codiagonal(m), flip_horizontal(m), flip_vertical(m)};
        break;
    }
    case kVH:
        info.order = 2;
        info.dense = dense_numbering(info.cls, m, info.size);
        info.variations = {m, flip_diagonal(m)};
        break;
    case kDC:
        info.order = 2;
        info.dense = dense_numbering(info.cls, m, info.size);
        info.variations = {m, flip_vertical(m)};
        break;
    case kVHDC:
        info.order = 1;
        info.dense = dense_numbering(info.cls, m, info.size);
        info.variations = {m};
        break;
    }
    return info;
}

PatternGroup::PatternGroup(const std::vector<uint64_t>& masks)
{
    uint64_t offset = 0;
    for (uint64_t in : masks) {
        PatternInfo info = make_pattern_info(in);
        info.offset = uint32_t(offset);
        const uint64_t m = info.mask;
        uint32_t dense_off = uint32_t(dense.size());
        dense.insert(dense.end(), info.dense.begin(), info.dense.end());
        for (auto& T : slot_transforms(info.cls)) {
            SlotProgram s;
            s.col_offset = info.offset;
            s.block_size = info.size;
            switch (info.cls) {
            case kA:
                s.kind = 0;
                s.list[0] = list_for(T, m);
                break;
            case kVcls: {
                s.kind = 1;
                auto Tv = T;
                Tv.push_back(kV);
                s.list[0] = list_for(T, m & kHalfV);
                s.list[1] = list_for(Tv, m & kHalfV);
                s.hs = pow3(popcount64(m & kHalfV));
                break;
            }
            case kDcls: {
                s.kind = 2;
                auto Td = T;
                Td.push_back(kD);
                s.list[0] = list_for(T, m & kHalfD);
                s.list[1] = list_for(Td, m & kHalfD);
                s.list[2] = list_for(T, m & kDiagL);
                s.hs = pow3(popcount64(m & kHalfD));
                s.ds = pow3(popcount64(m & kDiagL));
                break;
            }
            default:
                s.kind = 3;
                s.list[0] = list_for(T, m);
                s.dense_offset = dense_off;
                break;
            }
            slots.push_back(s);
        }
        offset += info.size;
        if (offset > 0xFFFFFFFFull) throw std::runtime_error("Index space exceeds 32 bits.");
        patterns.push_back(std::move(info));
    }
    cols_ = offset;
}

void PatternGroup::row(uint64_t P, uint64_t O, uint32_t* out) const
{
    for (size_t i = 0; i < slots.size(); i++) {
        const SlotProgram& s = slots[i];
        uint32_t a = eval_list(s.list[0], P, O), idx;
        if (s.kind == 0) {
            idx = a;
        } else if (s.kind == 3) {
            idx = dense[s.dense_offset + a];
        } else {
            uint32_t b = eval_list(s.list[1], P, O);
            uint32_t lo = std::min(a, b), hi = std::max(a, b);
            idx = lo * s.hs + hi - lo * (lo + 1) / 2;
            if (s.kind == 2) idx = idx * s.ds + eval_list(s.list[2], P, O);
        }
        out[i] = s.col_offset + idx;
    }
}

}  // namespace cad
