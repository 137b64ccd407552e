// The expanded-table evaluator on the device: expansion of compact weights
// into one table of 3^n floats per variation (ScoreEstimator's constructor,
// Pattern/Evaluator.cpp:11-29), compaction back (ScoreEstimator::Weights,
// Pattern/Evaluator.cpp:56-73), and a resident evaluator that keeps the
// tables in HBM and scores batches of positions (ScoreEstimator::Eval,
// Pattern/Evaluator.cpp:48-54, for the leaves of a search).
#include "evaluator.hpp"

#include <algorithm>
#include <cstring>

#include "kernels.hpp"

namespace cad {

namespace {

// w_i[t] = x[column of variation i for the configuration whose FastIndex over
// the variation's mask is t]: t's base-3 digits (least significant first) are
// the squares of the mask in ascending order, digit 1 = P, 2 = O.  One thread
// per table entry; blockIdx.y is the slot.
__global__ __launch_bounds__(256) void k_expand_tables(const EvalSlot* __restrict__ es,
                                                       const DevSlot* __restrict__ slots,
                                                       const uint8_t* __restrict__ sq,
                                                       const uint32_t* __restrict__ dense,
                                                       const float* __restrict__ w, float* __restrict__ tables)
{
    const EvalSlot s = es[blockIdx.y];
    const DevSlot sl = slots[blockIdx.y];
    const uint32_t n3 = s.n3;
    for (uint32_t t = blockIdx.x * 256u + threadIdx.x; t < n3; t += gridDim.x * 256u) {
        uint64_t P = 0, O = 0;
        uint32_t rest = t;
        for (int j = 0; j < int(s.n); j++) {
            const uint32_t d = rest % 3u;
            rest /= 3u;
            const uint64_t bit = uint64_t(1) << s.sq[j];
            if (d == 1u) P |= bit;
            else if (d == 2u) O |= bit;
        }
        tables[s.table_offset + t] = w[sl.col_offset + slot_index(sl, sq, dense, P, O)];
    }
}

// Compaction, pass 1: for every configuration (p, o) of a pattern's mask
// (p & o == 0 over the mask's n bits, enumerated p outer, o inner as
// Configurations does, Pattern/Helpers.cpp:81-102), the key p * 2^n + o + 1
// orders the configurations as the enumeration does; each compact column
// keeps the largest key of the configurations whose index it is (the last
// one the reference writes).  blockIdx.y is the pattern.
__device__ __forceinline__ void config_position(const EvalSlot& s, uint32_t pb, uint32_t ob, uint64_t& P,
                                                uint64_t& O)
{
    P = O = 0;
    for (int j = 0; j < int(s.n); j++) {
        const uint64_t bit = uint64_t(1) << s.sq[j];
        if ((pb >> j) & 1u) P |= bit;
        if ((ob >> j) & 1u) O |= bit;
    }
}

__global__ __launch_bounds__(256) void k_compact_keys(const EvalSlot* __restrict__ es,
                                                      const int32_t* __restrict__ first_slot,
                                                      const DevSlot* __restrict__ slots,
                                                      const uint8_t* __restrict__ sq,
                                                      const uint32_t* __restrict__ dense, uint32_t* __restrict__ key)
{
    const int s0 = first_slot[blockIdx.y];
    const EvalSlot s = es[s0];
    const DevSlot sl = slots[s0];
    const uint32_t n = s.n, nn = 1u << (2 * n);
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < nn; c += gridDim.x * 256u) {
        const uint32_t pb = c >> n, ob = c & ((1u << n) - 1u);
        if (pb & ob) continue;
        uint64_t P, O;
        config_position(s, pb, ob, P, O);
        atomicMax(&key[sl.col_offset + slot_index(sl, sq, dense, P, O)], c + 1u);
    }
}

// Compaction, pass 2: the configuration holding a column's key writes
// w[column] = table 0 of the pattern at its FastIndex.
__global__ __launch_bounds__(256) void k_compact_write(const EvalSlot* __restrict__ es,
                                                       const int32_t* __restrict__ first_slot,
                                                       const DevSlot* __restrict__ slots,
                                                       const uint8_t* __restrict__ sq,
                                                       const uint32_t* __restrict__ dense,
                                                       const uint32_t* __restrict__ key,
                                                       const float* __restrict__ tables, float* __restrict__ w)
{
    const int s0 = first_slot[blockIdx.y];
    const EvalSlot s = es[s0];
    const DevSlot sl = slots[s0];
    const uint32_t n = s.n, nn = 1u << (2 * n);
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < nn; c += gridDim.x * 256u) {
        const uint32_t pb = c >> n, ob = c & ((1u << n) - 1u);
        if (pb & ob) continue;
        uint64_t P, O;
        config_position(s, pb, ob, P, O);
        const uint32_t col = sl.col_offset + slot_index(sl, sq, dense, P, O);
        if (key[col] != c + 1u) continue;
        uint32_t t = 0;
        for (int j = int(n) - 1; j >= 0; --j) t = t * 3u + ((pb >> j) & 1u) + 2u * ((ob >> j) & 1u);
        w[col] = tables[s.table_offset + t];
    }
}

// Eval for every position: the sum over the variations of table_i[FastIndex(
// pos, mask_i)], in variation order from 0.0f (bit-identical to
// ScoreEstimator::Eval and to (A x)[row]).  A thread scores kEvalRows
// positions (r, r + 256, ...) from their ternary words; each variation's
// FastIndex is its ternary-word program (the variation's squares in
// ascending order), read once for all of them.  The descriptors are the same
// for the whole wave.
constexpr int kEvalRows = 4;

__global__ __launch_bounds__(256) void k_eval_tables(const uint64_t* __restrict__ pos, int64_t n,
                                                     const EvalSlot* __restrict__ es,
                                                     const DevTern* __restrict__ ts,
                                                     const uint32_t* __restrict__ tern, int k,
                                                     const float* __restrict__ tables, float* __restrict__ out)
{
    const int64_t r0 = blockIdx.x * int64_t(256 * kEvalRows) + threadIdx.x;
    if (r0 >= n) return;
    uint32_t T[kEvalRows][4];
    bool ok[kEvalRows];
    float sum[kEvalRows];
#pragma unroll
    for (int i = 0; i < kEvalRows; i++) {
        const int64_t r = r0 + int64_t(i) * 256;
        ok[i] = r < n;
        const int64_t rr = ok[i] ? r : r0;  // a valid row's words for the lanes past the end
        ternary_words(pos[2 * rr], pos[2 * rr + 1], T[i]);
        sum[i] = 0.0f;
    }
    for (int s = 0; s < k; s++) {
        const EvalSlot e = uniform_load(es + s);
        const DevTern t = uniform_load(ts + s);
        uint32_t v[kEvalRows];
        tern_values<kEvalRows>(tern, t.off[0], t.nent[0], T, v);
#pragma unroll
        for (int i = 0; i < kEvalRows; i++) sum[i] += tables[e.table_offset + v[i]];
    }
#pragma unroll
    for (int i = 0; i < kEvalRows; i++)
        if (ok[i]) out[r0 + int64_t(i) * 256] = sum[i];
}

int blocks_for(uint64_t n, int cap = 4096)
{
    return int(std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, uint64_t(cap))));
}

}  // namespace

// Per-slot table layout of a pattern group (variation after variation, in
// slot order), uploaded once per pattern group.
const ExpandedLayout& expanded_layout(cad_patterns* p)
{
    if (p->layout) return *p->layout;
    std::unique_ptr<ExpandedLayout> L(new ExpandedLayout);
    const PatternGroup& g = *p->group;
    std::vector<EvalSlot> es;
    int slot = 0;
    uint64_t off = 0;
    for (const PatternInfo& pi : g.patterns) {
        L->first_slot.push_back(slot);
        for (uint64_t v : pi.variations) {
            EvalSlot e;
            std::memset(&e, 0, sizeof(e));
            int j = 0;
            for (int q = 0; q < 64; q++)
                if ((v >> q) & 1) e.sq[j++] = uint8_t(q);
            if (j > 16) throw Error(CAD_ERR_PATTERN, "expanded tables need patterns of at most 16 squares");
            e.n = uint8_t(j);
            e.n3 = pow3(j);
            e.table_offset = off;
            off += e.n3;
            es.push_back(e);
            slot++;
        }
    }
    if (slot != g.k()) throw Error(CAD_ERR_PATTERN, "variation count differs from the slot count");
    // each variation's FastIndex as a ternary-word program (list 0)
    std::vector<DevTern> ts;
    std::vector<uint32_t> tern;
    for (const EvalSlot& e : es) {
        DevTern t;
        std::memset(&t, 0, sizeof(t));
        append_ternary_list(e.sq, int(e.n), tern, t, 0);
        ts.push_back(t);
    }
    tern.insert(tern.end(), kTernPadding, 0u);
    L->slots = es;
    L->size = int64_t(off);
    hipStream_t s = p->ctx->stream;
    L->d_tslots = to_device(ts, s);
    L->d_tern = to_device(tern, s);
    L->d_slots = to_device(es, s);
    L->d_first_slot = to_device(L->first_slot, s);
    CAD_HIP(hipStreamSynchronize(s));  // the host vectors are uploaded
    p->layout = std::move(L);
    return *p->layout;
}

void weights_expand(cad_patterns* p, const float* w, float* tables)
{
    const ExpandedLayout& L = expanded_layout(p);
    uint32_t widest = 1;
    for (const EvalSlot& e : L.slots) widest = std::max(widest, e.n3);
    hipLaunchKernelGGL(k_expand_tables, dim3(unsigned(blocks_for(widest)), unsigned(L.slots.size())), dim3(256), 0,
                       p->ctx->stream, L.d_slots.as<EvalSlot>(), p->d_slots.as<DevSlot>(), p->d_sq.as<uint8_t>(),
                       p->d_dense.as<uint32_t>(), w, tables);
    CAD_HIP(hipGetLastError());
}

void weights_compact(cad_patterns* p, const float* tables, float* w)
{
    const ExpandedLayout& L = expanded_layout(p);
    const int64_t cols = int64_t(p->group->cols());
    hipStream_t s = p->ctx->stream;
    DBuf key(sizeof(uint32_t) * size_t(std::max<int64_t>(1, cols)));
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(unsigned(blocks_for(uint64_t(cols)))), dim3(256), 0, s,
                       key.as<uint32_t>(), cols, 0u);
    hipLaunchKernelGGL(k_fill<float>, dim3(unsigned(blocks_for(uint64_t(cols)))), dim3(256), 0, s, w, cols, 0.0f);
    uint32_t widest = 1;
    for (int f : L.first_slot) widest = std::max(widest, 1u << (2 * L.slots[size_t(f)].n));
    const dim3 grid(unsigned(blocks_for(widest)), unsigned(L.first_slot.size()));
    hipLaunchKernelGGL(k_compact_keys, grid, dim3(256), 0, s, L.d_slots.as<EvalSlot>(), L.d_first_slot.as<int32_t>(),
                       p->d_slots.as<DevSlot>(), p->d_sq.as<uint8_t>(), p->d_dense.as<uint32_t>(),
                       key.as<uint32_t>());
    hipLaunchKernelGGL(k_compact_write, grid, dim3(256), 0, s, L.d_slots.as<EvalSlot>(), L.d_first_slot.as<int32_t>(),
                       p->d_slots.as<DevSlot>(), p->d_sq.as<uint8_t>(), p->d_dense.as<uint32_t>(),
                       (const uint32_t*)key.as<uint32_t>(), tables, w);
    CAD_HIP(hipGetLastError());
    CAD_HIP(hipStreamSynchronize(s));  // the key buffer is freed on return
}

void eval_tables(cad_patterns* p, const float* tables, const uint64_t* pos, int64_t n, float* out)
{
    if (n <= 0) return;
    const ExpandedLayout& L = expanded_layout(p);
    hipLaunchKernelGGL(k_eval_tables, dim3(unsigned((n + 256 * kEvalRows - 1) / (256 * kEvalRows))), dim3(256), 0,
                       p->ctx->stream, pos, n, L.d_slots.as<EvalSlot>(), L.d_tslots.as<DevTern>(),
                       L.d_tern.as<uint32_t>(), int(L.slots.size()), tables, out);
    CAD_HIP(hipGetLastError());
}

}  // namespace cad
