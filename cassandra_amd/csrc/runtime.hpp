// Host runtime objects behind the C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "device.hpp"
#include "patterns.hpp"

namespace cad {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what);
#define CAD_HIP(x) ::cad::hip_check((x), #x)

struct Context;

// Owning device buffer.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    explicit DBuf(size_t n);
    ~DBuf();
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DBuf& operator=(DBuf&& o) noexcept;
    template <class T> T* as() const { return static_cast<T*>(p); }
    void upload(const void* h, size_t n, hipStream_t s);
};

template <class T>
DBuf to_device(const std::vector<T>& v, hipStream_t s)
{
    DBuf b(std::max<size_t>(1, v.size()) * sizeof(T));
    if (!v.empty()) b.upload(v.data(), v.size() * sizeof(T), s);
    return b;
}

// Appends list l of a ternary-word program (DevTern): the squares src[j] of a
// base-3 number sum_j 3^j (P[src_j] + 2 O[src_j]) as entries {2 (q mod 16),
// 3^j} grouped by the word q / 16; sets t.off[l] and t.nent[l].  Throws for
// more than 16 squares (the kernels multiply by 3^j in 24 bits).  A program
// array must end with kTernPadding zero words.
void append_ternary_list(const uint8_t* src, int n, std::vector<uint32_t>& tern, DevTern& t, int l);
constexpr int kTernPadding = 6;

// The two-stage scatter of one value size: tiles (window x row chunk) that
// write partial slices, and the column reduction over the slices.
struct ScatterLists {
    // the slot groups of these lists: the pattern groups, with consecutive
    // groups narrower than a window merged while they fit one window together
    // (at most kMaxSegSlots slots), so one pass over a row chunk serves them
    std::vector<SlotGroup> groups;
    std::vector<uint16_t> slot_list;
    DBuf d_groups, d_slot_list;
    std::vector<ScatterWork> work;
    std::vector<ReduceWork> reduce;
    DBuf d_work, d_reduce;
    uint32_t widest = 0;     // widest window (LDS of the tile kernel)
    int64_t part_elems = 0;  // partial buffer length
    DBuf part;               // partial buffer (allocated on first use)
    int cbits = 0;           // a window column gets fewer than 2^cbits terms per work item
    // reduce items of each system (the list is ordered by system), for a
    // reduce that also sums over its columns (block partials per item)
    std::vector<int32_t> red_first, red_count;
    DBuf d_red_first, d_red_count, d_red_counters, d_red_partials;
    GridSys reduce_sys() const;
};

// Slot groups + per-slot column ranges -> scatter work lists.
struct ScatterPlan {
    std::vector<SlotGroup> groups;
    std::vector<uint16_t> slot_list;
    DBuf d_groups, d_slot_list;
    ScatterLists f4, f8;  // float and double (and uint32 counts) windows
    DBuf d_ymax;          // per system: bits of max |y| of the last FP32 scatter (fixed-point scale)
    DBuf d_ymax_part;     // the same per row block of the grid plan
};

// LDS-staged gather of one value size: row tiles and the slot segments.
struct GatherPlan {
    bool ok = false;  // the matrix has the pattern structure the tile kernel needs
    int rpt = 0;      // rows per thread
    int bs = 1024;    // threads per block (1024, or 512: two blocks and half the LDS each per CU)
    int sub = 1;      // sub-tiles of bs * rpt rows that share each staging
    uint32_t wmax = 0;
    uint32_t lds_elems = 0;
    std::vector<GatherSeg> segs;     // matrix-free kernel: one staging per segment
    std::vector<GatherStep> steps;   // stored-index kernel: packed phases, prefetched steps
    std::vector<PhaseSeg> phases;
    std::vector<RowWork> tiles;
    std::vector<int32_t> first, count;
    DBuf d_segs, d_steps, d_phases, d_tiles, d_first, d_count, d_counters, d_partials;
    GridSys sys() const;
};

// Blocks of rows / columns of the systems, with the per-system grid
// reduction bookkeeping.
struct GridPlan {
    std::vector<RowWork> rows;
    std::vector<ColWork> cols;
    DBuf d_rows, d_cols;
    std::vector<int32_t> row_first, row_count, col_first, col_count;
    DBuf d_row_first, d_row_count, d_col_first, d_col_count;
    DBuf d_row_counters, d_col_counters, d_partials;
    GridSys row_sys() const;
    GridSys col_sys() const;
};

// Schedule knobs of a matrix's plans, chosen by the autotuner (or defaults).
struct TuneParams {
    int rpt4 = 4;               // rows per thread of the FP32 tile gather (0: row kernel)
    int rpt8 = 2;               // rows per thread of the FP64 tile gather (0: row kernel)
    int bs4 = 1024, bs8 = 1024; // threads per block of the FP32 / FP64 tile gather (1024 or 512)
    int sub4 = 1, sub8 = 1;     // sub-tiles per staging of the FP32 / FP64 tile gather (1..3; 1 with 512)
    int srpt = 4;               // consecutive rows per step of the stored-index scatter (4, or 8 for u16)
    double blocks_per_cu = 1.0; // scatter tiles per CU the row chunks aim at
    double partial_cap = 1.0;   // partial-slice traffic cap as a fraction of index traffic
    int scatter_order = 0;      // 0: chunks per window by work share; 1: shared chunks, XCD-major
    double scatter_balance = 1.0; // order 0: weight of slot loads against atomics in a window's work (1: loads)
    double small_floor_bytes = 8.0 * 1024 * 1024;  // partial slices a window may take to spread a small system
};

// A matrix (or, in a batch, the systems still iterating) of at most this many
// rows gathers with the row kernel: a staged tile costs about the same (≈22
// µs) whatever its rows, because it stages every x window of the pattern set,
// while the row kernel (one row a thread, x read through L2) takes ≈10 µs at
// 100k rows and ≈15 µs at 250k (FIT4).
constexpr int64_t kRowKernelRows = 262144;

// Default product schedule of a new stored-index matrix of one system.
inline void default_schedule(TuneParams& t, int ibytes, int64_t rows)
{
    t.rpt4 = ibytes == 2 ? 8 : 4;
    if (rows <= kRowKernelRows) t.rpt4 = t.rpt8 = 0;
}

// Slot columns of a stored index matrix start at multiples of this many rows,
// so a thread can load the entries of up to 8 consecutive rows of one slot in
// one aligned vector load.
constexpr int64_t kIndexAlign = 64;
inline int64_t padded_rows(int64_t rows) { return (rows + kIndexAlign - 1) / kIndexAlign * kIndexAlign; }

struct Workspace;

// Expanded tables of a pattern group: one table of 3^n floats per variation,
// variation after variation in slot order (ScoreEstimator's w[i]).
struct ExpandedLayout {
    std::vector<EvalSlot> slots;
    std::vector<int32_t> first_slot;  // first slot (variation 0) of each pattern
    int64_t size = 0;                 // floats of all tables
    DBuf d_slots, d_first_slot;
    DBuf d_tslots, d_tern;            // FastIndex of every variation as a ternary-word program
};

struct Context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    int cus = 0;
    size_t max_lds = 0;
    ~Context();
};

}  // namespace cad

struct cad_context : cad::Context {};

struct cad_patterns {
    cad_context* ctx = nullptr;
    std::unique_ptr<cad::PatternGroup> group;
    cad::DBuf d_slots, d_sq, d_dense;
    cad::DBuf d_tslots, d_tern;             // the ternary-word form of the slot programs
    std::vector<uint32_t> slot_lo, slot_hi;  // global column range of every slot
    // scatter plans of the evaluator backward, one per minibatch size
    std::map<int64_t, std::unique_ptr<cad::ScatterPlan>> eval_plans;
    cad::DBuf d_base0, d_syscol0;            // zeros
    std::unique_ptr<cad::ExpandedLayout> layout;  // expanded tables (built on first use)
    // the training step's buffers: per-row gradient of the loss, the weight
    // gradient, the forward's block sums of the loss, the minibatch's index
    // rows made contiguous (grown on demand)
    cad::DBuf d_train_g, d_train_grad, d_train_part, d_train_rows;
};

// A resident evaluator: the expanded tables of one weight vector in HBM and
// staging for host batches (ScoreEstimator on the device).
struct cad_evaluator {
    cad_context* ctx = nullptr;        // the pattern group's context
    cad_patterns* patterns = nullptr;
    cad::DBuf tables;
    cad::DBuf d_pos, d_out;     // staging of cad_evaluator_eval_host
    int64_t staged = 0;         // rows the staging holds
};

struct cad_matrix {
    cad_context* ctx = nullptr;
    int64_t rows = 0, cols = 0;
    int k = 0;
    int ibytes = 4;                 // 2: u16 local + slot base, 4: u32 with base 0, 0: matrix-free
    cad::DBuf d_idx, d_base;        // slot-major: entry (r, s) at d_idx[s * ld + r]
    int64_t ld = 0;                 // rows rounded up to kIndexAlign (rows r >= rows are zero)
    // matrix-free: the positions and a copy of the pattern group's slot programs
    bool mf = false;
    cad::DBuf d_pos, d_slots, d_sq, d_dense, d_tslots, d_tern, d_tw;
    std::vector<uint32_t> slot_lo, slot_hi;
    int nsys = 1;
    std::vector<int64_t> sys_row_begin;  // nsys + 1
    std::vector<uint32_t> sys_col;       // column base of each system
    int64_t cols_per_sys = 0;
    cad::DBuf d_sys_col;
    cad::ScatterPlan splan;
    cad::GridPlan grid;
    cad::GatherPlan gplan4, gplan8;  // float / double tiles
    cad::TuneParams tune;           // schedule of the plans below
    bool tuned = false;
    cad::DBuf d_jacobi_cnt;         // cached exact multiplicity counts (u32)
    bool have_cnt = false;
    std::map<int, std::unique_ptr<cad::Workspace>> ws;
    ~cad_matrix();
};

namespace cad {
void build_plans(cad_matrix* m);
void build_gather_plan(cad_matrix* m, GatherPlan& gp, int vbytes, int rpt, int bs, int sub);
void build_scatter_plan(ScatterPlan& sp, const std::vector<uint32_t>& lo, const std::vector<uint32_t>& hi,
                        const std::vector<int64_t>& sys_rows, const std::vector<uint32_t>& sys_col,
                        int64_t cols_per_sys, int ibytes, const Context& ctx, const TuneParams& tune);
// The 8-byte window lists of the systems in `on` alone (host vectors only:
// work, reduce items ordered by system with each system's range, part_elems,
// cbits), their row chunks set for the windows those systems run side by
// side.  The groups are the plan's (the same packing), so the plan's device
// groups and slot lists serve the work items.
void plan_scatter_subset(const ScatterPlan& sp, const std::vector<int64_t>& sys_rows,
                         const std::vector<uint32_t>& sys_col, int64_t cols_per_sys, int ibytes, const Context& ctx,
                         const TuneParams& tune, const std::vector<uint8_t>& on, ScatterLists& L);
void set_error(const std::string& s);
}  // namespace cad
