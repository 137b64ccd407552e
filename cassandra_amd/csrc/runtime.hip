// Device buffers, contexts, pattern groups, matrices and their launch plans.
#include "runtime.hpp"

#include "../../include/cassandra_amd.h"

#include <algorithm>
#include <cstring>
#include <numeric>

#include "kernels.hpp"

namespace cad {

void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw Error(-5, std::string("out of device memory: ") + what);
        throw Error(-4, std::string(hipGetErrorString(e)) + " in " + what);
    }
}

DBuf::DBuf(size_t n) : bytes(n)
{
    if (n) {
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(-5, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
    }
}

DBuf::~DBuf()
{
    if (p) (void)hipFree(p);
}

DBuf& DBuf::operator=(DBuf&& o) noexcept
{
    if (this != &o) {
        if (p) (void)hipFree(p);
        p = o.p;
        bytes = o.bytes;
        o.p = nullptr;
        o.bytes = 0;
    }
    return *this;
}

void DBuf::upload(const void* h, size_t n, hipStream_t s)
{
    CAD_HIP(hipMemcpyAsync(p, h, n, hipMemcpyHostToDevice, s));
    CAD_HIP(hipStreamSynchronize(s));
}

Context::~Context()
{
    if (own) (void)hipStreamDestroy(own);
}

GridSys GridPlan::row_sys() const
{
    return GridSys{d_row_first.as<int32_t>(), d_row_count.as<int32_t>(), d_row_counters.as<unsigned>(),
                   d_partials.as<double>()};
}

GridSys GridPlan::col_sys() const
{
    return GridSys{d_col_first.as<int32_t>(), d_col_count.as<int32_t>(), d_col_counters.as<unsigned>(),
                   d_partials.as<double>()};
}

namespace {

constexpr uint32_t kReduceCols = kRowBlock;  // columns of one scatter-reduce work item

// Row chunks of one window of a system: the share of the target tile count,
// capped so that the partial slices stay near partial_cap of the window's
// index traffic -- and, where that cap would leave CUs idle, raised towards
// one tile per CU while its slices stay under the floor bytes a window
// (their write and read then cost less than the parallelism gains):
// small_floor (8 MB by default) for a small system (at most kRowKernelRows
// rows, whose tiles are latency-bound: a 100k-row PCG iteration 47 -> 42 µs),
// 2 MB for a large one (8 MB took the 5M-row matrix's XCD-major plan from 240
// to 360 tiles, 165 -> 245 µs).
// `windows` counts the windows of every system planned together: those of
// many systems already fill the CUs side by side.  At least kScatterBlock
// rows a chunk.
constexpr double kPartialFloorBytes = 2.0 * 1024 * 1024;
constexpr double kMultiSystemTiles = 8.0;  // target tiles a CU, over blocks_per_cu, for several systems

int64_t window_chunks(int64_t R, double share_chunks, double cap_chunks, int cus, int windows, uint32_t W,
                      int vbytes, double small_floor)
{
    int64_t C = std::max<int64_t>(1, std::min<int64_t>(std::llround(share_chunks), int64_t(cap_chunks)));
    const int64_t par = std::max<int64_t>(1, cus / std::max(1, windows));
    const double floor_bytes = R <= kRowKernelRows ? small_floor : kPartialFloorBytes;
    const int64_t floor_c = std::max<int64_t>(1, int64_t(floor_bytes / (2.0 * double(W) * vbytes)));
    C = std::max(C, std::min(par, floor_c));
    return std::max<int64_t>(1, std::min<int64_t>(C, (R + kScatterBlock - 1) / kScatterBlock));
}

// Group slots whose column ranges overlap (transitively).
void make_groups(const std::vector<uint32_t>& lo, const std::vector<uint32_t>& hi, ScatterPlan& sp)
{
    const int k = int(lo.size());
    std::vector<int> order(k);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return lo[a] < lo[b]; });
    sp.groups.clear();
    sp.slot_list.clear();
    int i = 0;
    while (i < k) {
        uint32_t glo = lo[order[i]], ghi = hi[order[i]];
        int j = i + 1;
        while (j < k && lo[order[j]] < ghi) {
            ghi = std::max(ghi, hi[order[j]]);
            j++;
        }
        std::vector<uint16_t> members;
        for (int t = i; t < j; t++) members.push_back(uint16_t(order[t]));
        std::sort(members.begin(), members.end());  // keep slot order inside a group
        SlotGroup g;
        g.col_lo = glo;
        g.col_hi = ghi;
        g.nslots = uint16_t(members.size());
        g.first = uint16_t(sp.slot_list.size());
        sp.slot_list.insert(sp.slot_list.end(), members.begin(), members.end());
        sp.groups.push_back(g);
        i = j;
    }
}

// The scatter groups of a window size (ScatterLists::groups).  Pattern
// groups come in column order and never overlap.
void pack_groups(const ScatterPlan& sp, uint32_t wmax, ScatterLists& L)
{
    L.groups.clear();
    L.slot_list.clear();
    size_t i = 0;
    while (i < sp.groups.size()) {
        SlotGroup g = sp.groups[i];
        std::vector<uint16_t> slots(sp.slot_list.begin() + g.first, sp.slot_list.begin() + g.first + g.nslots);
        size_t j = i + 1;
        if (g.col_hi - g.col_lo <= wmax)
            for (; j < sp.groups.size(); j++) {
                const SlotGroup& h = sp.groups[j];
                if (h.col_lo < g.col_hi || h.col_hi - g.col_lo > wmax || g.nslots + h.nslots > kMaxSegSlots) break;
                slots.insert(slots.end(), sp.slot_list.begin() + h.first, sp.slot_list.begin() + h.first + h.nslots);
                g.col_hi = h.col_hi;
                g.nslots = uint16_t(g.nslots + h.nslots);
            }
        g.first = uint16_t(L.slot_list.size());
        L.slot_list.insert(L.slot_list.end(), slots.begin(), slots.end());
        L.groups.push_back(g);
        i = j;
    }
}

// Reduce items of one window ([wl, wh) of system s, chunk slices at off with
// stride W) and of the columns outside every window (written as zeros).
void push_reduce(ScatterLists& L, int s, uint32_t sys_col, uint32_t wl, uint32_t wh, int64_t off, int64_t C)
{
    // one column per thread of a reduce block (kReduceCols), so the reduce
    // launches enough blocks to cover the CUs
    for (uint32_t c0 = wl; c0 < wh; c0 += kReduceCols) {
        ReduceWork r;
        r.col_begin = int64_t(sys_col) + c0;
        r.col_end = int64_t(sys_col) + std::min(wh, c0 + kReduceCols);
        r.part_off = C > 0 ? off + (c0 - wl) : 0;
        r.nchunk = int32_t(C);
        r.stride = C > 0 ? wh - wl : 0;
        r.sys = s;
        r.pad = 0;
        L.reduce.push_back(r);
    }
}

// Chunk-major scatter lists: every window of a system uses the same row
// chunks, and the tiles of one chunk (all its windows, the heavier ones split
// into row parts of equal work) are placed on one XCD one after the other
// (workgroup i runs on XCD i mod 8), so the y values and the index entries
// that the windows of a chunk share are read from HBM once and then hit that
// XCD's L2.  The chunk count is a multiple of the XCD count near the target
// tile count.  Slots left over at the end of an XCD's lane hold empty tiles
// (no rows, no columns).
constexpr int kXcds = 8;

void make_scatter_lists_xcd(const ScatterPlan& sp, const std::vector<int64_t>& sys_rows,
                            const std::vector<uint32_t>& sys_col, int64_t cols_per_sys, uint32_t wmax, int vbytes,
                            int ibytes, double target, int cus, const TuneParams& tune, ScatterLists& L,
                            const std::vector<uint8_t>* on)
{
    (void)sp;
    const int nsys = int(sys_col.size());
    auto in = [&](int s) { return !on || (*on)[size_t(s)] != 0; };
    // Windows of the slot groups; a window with more slots than the lightest
    // is split into parts (row sub-ranges of the chunk), so that every tile
    // of a chunk does about the same work and the chunk's tiles, running
    // side by side on one XCD, stream its rows at the same pace.
    struct Win { uint32_t lo, hi; int group; int nslots; int parts; };
    std::vector<Win> wins;
    double work_per_row = 0, width = 0;
    int minslots = 1 << 30;
    for (size_t gi = 0; gi < L.groups.size(); gi++) {
        const SlotGroup& g = L.groups[gi];
        for (uint32_t wl = g.col_lo; wl < g.col_hi; wl += wmax) {
            const uint32_t wh = std::min(g.col_hi, wl + wmax);
            wins.push_back(Win{wl, wh, int(gi), g.nslots, 1});
            L.widest = std::max(L.widest, wh - wl);
            work_per_row += g.nslots;
            width += wh - wl;
            minslots = std::min(minslots, int(g.nslots));
        }
    }
    const int nw = int(wins.size());
    int tiles_per_chunk = 0;
    for (Win& w : wins) {
        w.parts = std::max(1, int(std::lround(double(w.nslots) / std::max(1, minslots))));
        tiles_per_chunk += w.parts;
    }
    int64_t total_rows = 0;
    int nrun = 0;  // systems planned here that have rows: their windows run side by side
    for (int s = 0; s < nsys; s++)
        if (in(s)) {
            total_rows += sys_rows[s + 1] - sys_rows[s];
            nrun += sys_rows[s + 1] > sys_rows[s];
        }
    // chunks: about target tiles in all, a multiple of the XCD count
    int64_t chunks = std::max<int64_t>(1, std::llround(target / std::max(1, tiles_per_chunk)));
    if (chunks > kXcds) chunks = chunks / kXcds * kXcds;
    std::vector<std::vector<ScatterWork>> lanes(kXcds);
    int64_t off = 0;
    int chunk_id = 0;
    for (int s = 0; s < nsys; s++) {
        if (!in(s)) continue;
        const int64_t R = sys_rows[s + 1] - sys_rows[s];
        int64_t C = 0;
        std::vector<int64_t> woff(size_t(std::max(1, nw)));
        if (R > 0 && nw > 0) {
            const double share = total_rows > 0 ? double(R) / double(total_rows) : 1.0;
            const double cap = tune.partial_cap * double(R) * work_per_row * ibytes /
                               (2.0 * width * vbytes * double(tiles_per_chunk) / std::max(1, nw));
            C = window_chunks(R, double(chunks) * share, cap, cus, tiles_per_chunk * std::max(1, nrun),
                              uint32_t(width / std::max(1, nw)), vbytes, tune.small_floor_bytes);
            const int64_t chunk = (R + C - 1) / C;
            C = (R + chunk - 1) / chunk;
            for (int j = 0; j < nw; j++) {
                woff[size_t(j)] = off;
                off += C * wins[size_t(j)].parts * int64_t(wins[size_t(j)].hi - wins[size_t(j)].lo);
            }
            for (int64_t c = 0; c < C; c++, chunk_id++) {
                auto& lane = lanes[size_t(chunk_id % kXcds)];
                const int64_t c0 = c * chunk, c1 = std::min(R, (c + 1) * chunk);
                for (int j = 0; j < nw; j++) {
                    const Win& wj = wins[size_t(j)];
                    const int64_t W = int64_t(wj.hi - wj.lo);
                    for (int q = 0; q < wj.parts; q++) {
                        ScatterWork w;
                        w.row_begin = sys_rows[s] + c0 + (c1 - c0) * q / wj.parts;
                        w.row_end = sys_rows[s] + c0 + (c1 - c0) * (q + 1) / wj.parts;
                        w.part_off = woff[size_t(j)] + (c * wj.parts + q) * W;
                        w.win_lo = sys_col[s] + wj.lo;
                        w.win_hi = sys_col[s] + wj.hi;
                        w.group = wj.group;
                        w.sys = s;
                        lane.push_back(w);
                    }
                }
            }
            for (int j = 0; j < nw; j++)
                push_reduce(L, s, sys_col[s], wins[size_t(j)].lo, wins[size_t(j)].hi, woff[size_t(j)],
                            C * wins[size_t(j)].parts);
        }
        // columns outside every window are written as zeros
        std::vector<std::pair<uint32_t, uint32_t>> covered;
        if (R > 0)
            for (const Win& w : wins) covered.push_back({w.lo, w.hi});
        std::sort(covered.begin(), covered.end());
        covered.push_back({uint32_t(cols_per_sys), uint32_t(cols_per_sys)});
        uint32_t at = 0;
        for (auto& iv : covered) {
            if (at < iv.first) push_reduce(L, s, sys_col[s], at, iv.first, 0, 0);
            at = std::max(at, iv.second);
        }
    }
    size_t depth = 0;
    for (auto& lane : lanes) depth = std::max(depth, lane.size());
    ScatterWork empty{};
    empty.row_begin = empty.row_end = 0;
    empty.part_off = 0;
    empty.win_lo = empty.win_hi = 0;
    empty.group = 0;
    empty.sys = 0;
    L.work.assign(depth * kXcds, empty);
    for (int x = 0; x < kXcds; x++)
        for (size_t p = 0; p < lanes[size_t(x)].size(); p++) L.work[p * kXcds + size_t(x)] = lanes[size_t(x)][p];
    L.part_elems = off;
}

// Scatter tiles of (system rows) x (group window).  The number of row chunks
// of a window is set by its share of the work (about two tiles per CU in
// total), capped so that the partial slices (written and read once) stay near
// 10% of the window's index traffic.
void make_scatter_lists(const ScatterPlan& sp, const std::vector<int64_t>& sys_rows,
                        const std::vector<uint32_t>& sys_col, int64_t cols_per_sys, uint32_t wmax, int vbytes,
                        int ibytes, int cus, const TuneParams& tune, ScatterLists& L,
                        const std::vector<uint8_t>* on = nullptr)
{
    L.work.clear();
    L.reduce.clear();
    L.widest = 0;
    pack_groups(sp, wmax, L);
    const int nsys = int(sys_col.size());
    auto in = [&](int s) { return !on || (*on)[size_t(s)] != 0; };
    // Work of a window per row: its slot loads, and its atomics -- the share
    // of the group's entries that fall into the window, taken as its share of
    // the group's columns.  scatter_balance weighs the two (1: loads only).
    const double lam = std::min(1.0, std::max(0.0, tune.scatter_balance));
    auto win_cost = [&](const SlotGroup& g, uint32_t W) {
        const double frac = double(W) / double(std::max<uint32_t>(1, g.col_hi - g.col_lo));
        return g.nslots * (lam + (1.0 - lam) * frac);
    };
    double total = 0;
    int nrun = 0;  // systems planned here that have rows: their windows run side by side
    for (int s = 0; s < nsys; s++) {
        if (!in(s)) continue;
        const int64_t R = sys_rows[s + 1] - sys_rows[s];
        nrun += R > 0;
        for (auto& g : L.groups)
            for (uint32_t wl = g.col_lo; wl < g.col_hi; wl += wmax)
                total += double(R) * win_cost(g, std::min(g.col_hi, wl + wmax) - wl);
    }
    // several systems side by side have tiles of very different sizes (their
    // row counts differ): more tiles a CU let the launch balance them
    const double target = tune.blocks_per_cu * std::max(cus, 1) * (nrun > 1 ? kMultiSystemTiles : 1.0);
    int windows = 0;
    for (auto& g : L.groups) windows += int((g.col_hi - g.col_lo + wmax - 1) / wmax);
    int64_t off = 0;
    if (tune.scatter_order == 1) {
        make_scatter_lists_xcd(sp, sys_rows, sys_col, cols_per_sys, wmax, vbytes, ibytes, target, cus, tune, L, on);
        return;
    }
    for (int s = 0; s < nsys; s++) {
        if (!in(s)) continue;
        const int64_t R = sys_rows[s + 1] - sys_rows[s];
        std::vector<std::pair<uint32_t, uint32_t>> covered;
        for (size_t gi = 0; gi < L.groups.size(); gi++) {
            const SlotGroup& g = L.groups[gi];
            for (uint32_t wl = g.col_lo; wl < g.col_hi; wl += wmax) {
                const uint32_t wh = std::min(g.col_hi, wl + wmax);
                const uint32_t W = wh - wl;
                covered.push_back({wl, wh});
                L.widest = std::max(L.widest, W);
                int64_t C = 0;
                if (R > 0) {
                    const double share = total > 0 ? double(R) * win_cost(g, W) / total : 1.0;
                    const double cap = tune.partial_cap * double(R) * g.nslots * ibytes / (2.0 * W * vbytes);
                    C = window_chunks(R, target * share, cap, cus, windows * std::max(1, nrun), W, vbytes,
                                      tune.small_floor_bytes);
                    const int64_t chunk = (R + C - 1) / C;
                    C = (R + chunk - 1) / chunk;
                    for (int64_t c = 0; c < C; c++) {
                        ScatterWork w;
                        w.row_begin = sys_rows[s] + c * chunk;
                        w.row_end = sys_rows[s] + std::min(R, (c + 1) * chunk);
                        w.part_off = off + c * int64_t(W);
                        w.win_lo = sys_col[s] + wl;
                        w.win_hi = sys_col[s] + wh;
                        w.group = int32_t(gi);
                        w.sys = s;
                        L.work.push_back(w);
                    }
                }
                // one column per thread of a reduce block (kReduceCols), so the
                // reduce launches enough blocks to cover the CUs
                for (uint32_t c0 = wl; c0 < wh; c0 += kReduceCols) {
                    ReduceWork r;
                    r.col_begin = int64_t(sys_col[s]) + c0;
                    r.col_end = int64_t(sys_col[s]) + std::min(wh, c0 + kReduceCols);
                    r.part_off = off + (c0 - wl);
                    r.nchunk = int32_t(C);
                    r.stride = W;
                    r.sys = s;
                    r.pad = 0;
                    L.reduce.push_back(r);
                }
                off += C * int64_t(W);
            }
        }
        // columns outside every window are written as zeros
        std::sort(covered.begin(), covered.end());
        covered.push_back({uint32_t(cols_per_sys), uint32_t(cols_per_sys)});
        uint32_t at = 0;
        for (auto& iv : covered) {
            for (uint32_t c0 = at; c0 < iv.first; c0 += kReduceCols) {
                ReduceWork r;
                r.col_begin = int64_t(sys_col[s]) + c0;
                r.col_end = int64_t(sys_col[s]) + std::min(iv.first, c0 + kReduceCols);
                r.part_off = 0;
                r.nchunk = 0;
                r.stride = 0;
                r.sys = s;
                r.pad = 0;
                L.reduce.push_back(r);
            }
            at = std::max(at, iv.second);
        }
    }
    L.part_elems = off;
    // longest first: a launch of more tiles than CUs then starts its big
    // chunks in the first wave and fills in with the small ones (each tile
    // writes its own slice, so the order changes no sum)
    auto cost = [&](const ScatterWork& w) {
        return double(w.row_end - w.row_begin) * win_cost(L.groups[size_t(w.group)], w.win_hi - w.win_lo);
    };
    std::stable_sort(L.work.begin(), L.work.end(),
                     [&](const ScatterWork& a, const ScatterWork& b) { return cost(a) > cost(b); });
}

// The host side of a list after planning: the bound on the terms a window
// column receives in one work item (rows x slots of the group), for the
// fixed-point scale of the FP32 scatter, and the reduce items ordered by
// system with each system's range of them.
void finish_lists(ScatterLists& L, int nsys)
{
    L.cbits = 0;
    for (const ScatterWork& w : L.work) {
        uint64_t n = uint64_t(std::max<int64_t>(0, w.row_end - w.row_begin)) * L.groups[size_t(w.group)].nslots;
        int b = 0;
        while (n) b++, n >>= 1;
        L.cbits = std::max(L.cbits, b);
    }
    std::stable_sort(L.reduce.begin(), L.reduce.end(),
                     [](const ReduceWork& a, const ReduceWork& b) { return a.sys < b.sys; });
    L.red_first.assign(size_t(nsys), 0);
    L.red_count.assign(size_t(nsys), 0);
    for (size_t i = L.reduce.size(); i-- > 0;) {
        L.red_first[size_t(L.reduce[i].sys)] = int32_t(i);
        L.red_count[size_t(L.reduce[i].sys)]++;
    }
}

}  // namespace

void append_ternary_list(const uint8_t* src, int n, std::vector<uint32_t>& tern, DevTern& t, int l)
{
    // the kernels multiply by 3^j in 24-bit arithmetic: 3^15 < 2^24
    if (n > 16) throw Error(CAD_ERR_PATTERN, "a pattern list of more than 16 squares");
    t.off[l] = uint32_t(tern.size() / 2);
    t.nent[l] = 0;
    for (int w = 0; w < 4; w++) {
        uint32_t cnt = 0;
        for (int j = 0; j < n; j++)
            if (src[j] / 16 == w) {
                tern.push_back(uint32_t(2 * (src[j] % 16)));
                tern.push_back(pow3(j));
                cnt++;
            }
        t.nent[l] |= cnt << (8 * w);
    }
}

// Dynamic LDS a window may use: the device limit minus 1 KB kept for the
// kernels' static LDS (block-reduction scratch).
size_t lds_budget(const Context& ctx)
{
    return std::min<size_t>(ctx.max_lds, 160 * 1024) - 1024;
}

GridSys GatherPlan::sys() const
{
    return GridSys{d_first.as<int32_t>(), d_count.as<int32_t>(), d_counters.as<unsigned>(), d_partials.as<double>()};
}

void build_scatter_plan(ScatterPlan& sp, const std::vector<uint32_t>& lo, const std::vector<uint32_t>& hi,
                        const std::vector<int64_t>& sys_rows, const std::vector<uint32_t>& sys_col,
                        int64_t cols_per_sys, int ibytes, const Context& ctx, const TuneParams& tune)
{
    make_groups(lo, hi, sp);
    const size_t lds = lds_budget(ctx);
    // the 4-byte windows serve the Jacobi counts only (once per matrix): per-
    // window chunks, several tiles per CU, the u32 partial slices uncapped
    TuneParams t4 = tune;
    t4.scatter_order = 0;
    t4.blocks_per_cu = std::max(4.0, tune.blocks_per_cu);
    t4.partial_cap = 1.0;
    t4.scatter_balance = 1.0;
    make_scatter_lists(sp, sys_rows, sys_col, cols_per_sys, uint32_t(lds / 4), 4, ibytes, ctx.cus, t4, sp.f4);
    make_scatter_lists(sp, sys_rows, sys_col, cols_per_sys, uint32_t(lds / 8), 8, ibytes, ctx.cus, tune, sp.f8);
    const int nsys = int(sys_rows.size()) - 1;
    finish_lists(sp.f4, nsys);
    finish_lists(sp.f8, nsys);
    sp.d_groups = to_device(sp.groups, ctx.stream);
    // slot lists padded to whole 32-bit words (the kernels read them as words)
    auto padded = [](std::vector<uint16_t> v) {
        if (v.size() % 2) v.push_back(0);
        return v;
    };
    sp.d_slot_list = to_device(padded(sp.slot_list), ctx.stream);
    for (ScatterLists* L : {&sp.f4, &sp.f8}) {
        L->d_groups = to_device(L->groups, ctx.stream);
        L->d_slot_list = to_device(padded(L->slot_list), ctx.stream);
        L->d_work = to_device(L->work, ctx.stream);
        L->d_reduce = to_device(L->reduce, ctx.stream);
        L->d_red_first = to_device(L->red_first, ctx.stream);
        L->d_red_count = to_device(L->red_count, ctx.stream);
        L->d_red_counters = to_device(std::vector<unsigned>(size_t(std::max(1, nsys)), 0u), ctx.stream);
        L->d_red_partials = DBuf(2 * sizeof(double) * std::max<size_t>(1, L->reduce.size()));
        L->part = DBuf();
    }
}

void plan_scatter_subset(const ScatterPlan& sp, const std::vector<int64_t>& sys_rows,
                         const std::vector<uint32_t>& sys_col, int64_t cols_per_sys, int ibytes, const Context& ctx,
                         const TuneParams& tune, const std::vector<uint8_t>& on, ScatterLists& L)
{
    make_scatter_lists(sp, sys_rows, sys_col, cols_per_sys, uint32_t(lds_budget(ctx) / 8), 8, ibytes, ctx.cus, tune, L,
                       &on);
    finish_lists(L, int(sys_rows.size()) - 1);
}

GridSys ScatterLists::reduce_sys() const
{
    return GridSys{d_red_first.as<int32_t>(), d_red_count.as<int32_t>(), d_red_counters.as<unsigned>(),
                   d_red_partials.as<double>()};
}

// Segments of the tiled gather: the slot groups must be runs of consecutive
// slots that follow each other in slot order (true for pattern matrices, whose
// groups are the patterns); groups wider than the LDS may hold at most
// kMaxSegSlots slots.
void build_gather_plan(cad_matrix* m, GatherPlan& gp, int vbytes, int rpt, int bs, int sub)
{
    const Context& ctx = *m->ctx;
    const ScatterPlan& sp = m->splan;
    gp.ok = false;
    gp.rpt = rpt;
    gp.bs = bs;
    gp.sub = sub;
    if (rpt <= 0) return;  // the row kernel
    if (bs != kTileBlock && (bs != kTileBlock / 2 || m->mf)) return;
    if (sub < 1 || sub > 3 || (sub > 1 && (bs != kTileBlock || m->mf))) return;
    const int per_cu = kTileBlock / bs;  // blocks resident per CU
    // A staged run of columns starts at its element offset inside a 16-byte
    // line (up to Q - 1 elements) so that whole lines map to whole LDS lines:
    // windows leave that room, and phase segments start on line boundaries.
    const uint32_t total = uint32_t((lds_budget(ctx) + 1024) / per_cu - 1024) / uint32_t(vbytes);
    const uint32_t Q = uint32_t(16 / vbytes);
    gp.wmax = total - (Q - 1);
    std::vector<const SlotGroup*> order;
    for (auto& g : sp.groups) order.push_back(&g);
    std::sort(order.begin(), order.end(),
              [&](const SlotGroup* a, const SlotGroup* b) { return sp.slot_list[a->first] < sp.slot_list[b->first]; });
    int next = 0;
    gp.segs.clear();
    gp.lds_elems = 0;
    for (const SlotGroup* g : order) {
        const int first = sp.slot_list[g->first];
        if (first != next) return;
        for (int j = 0; j < g->nslots; j++)
            if (sp.slot_list[g->first + j] != first + j) return;
        GatherSeg seg;
        seg.col_lo = g->col_lo;
        seg.width = g->col_hi - g->col_lo;
        seg.slot_first = uint16_t(first);
        seg.nslots = g->nslots;
        seg.nwin = uint16_t((seg.width + gp.wmax - 1) / gp.wmax);
        seg.pad = 0;
        if (seg.nwin > 1 && seg.nslots > kMaxSegSlots) return;
        gp.lds_elems = std::max(gp.lds_elems, std::min(seg.width, gp.wmax) + Q - 1);
        gp.segs.push_back(seg);
        next = first + g->nslots;
    }
    if (next != m->k || gp.segs.empty()) return;
    // Steps of the stored-index kernel: consecutive single-window segments are
    // packed into one staging phase while they fit the LDS together; a wide
    // segment is a step of its own.
    gp.steps.clear();
    gp.phases.clear();
    uint32_t phase_w = 0;
    int open_step = -1;  // the step that stages the current phase
    for (const GatherSeg& sg : gp.segs) {
        const bool wide = sg.nwin > 1;
        if (wide || open_step < 0 || phase_w + sg.width + Q - 1 > total) {
            open_step = -1;
            phase_w = 0;
        }
        const int per_step = wide ? wide_step_slots(rpt * m->ibytes) : kMaxSegSlots;
        for (int j0 = 0; j0 < sg.nslots; j0 += per_step) {
            GatherStep st{};
            st.slot_first = uint16_t(sg.slot_first + j0);
            st.nslots = uint16_t(std::min<int>(per_step, sg.nslots - j0));
            st.col_lo = sg.col_lo;
            st.width = sg.width;
            st.nwin = sg.nwin;
            st.phase_first = -1;
            st.phase_count = 0;
            for (int j = 0; j < st.nslots; j++) {
                const int slot = st.slot_first + j;
                st.off[j] = (m->ibytes == 2 ? m->slot_lo[size_t(slot)] : 0u) - sg.col_lo;
            }
            if (wide && j0 > 0) {
                // the windows of this step in the other direction, from the
                // one the step before ended on
                st.wflags = uint16_t(kWideStaged | ((j0 / per_step) % 2 ? kWideReverse : 0));
            }
            if (!wide) {
                if (j0 == 0) {
                    if (open_step < 0) {
                        st.phase_first = int16_t(gp.phases.size());
                        open_step = int(gp.steps.size());
                    }
                    gp.phases.push_back(PhaseSeg{sg.col_lo, sg.width, phase_w, 0u});
                }
                st.lds_off = phase_w;
                for (int j = 0; j < st.nslots; j++) st.off[j] += phase_w;
            }
            gp.steps.push_back(st);
        }
        if (!wide) {
            GatherStep& opener = gp.steps[size_t(open_step)];
            opener.phase_count = int16_t(int(gp.phases.size()) - opener.phase_first);
            gp.lds_elems = std::max(gp.lds_elems, phase_w + sg.width + Q - 1);
            phase_w = (phase_w + sg.width + Q - 1 + Q - 1) / Q * Q;
        }
    }
    // Row tiles: a thread owns rpt consecutive rows counted from the tile's
    // first row rounded down to rpt, so a tile spans at most sub * bs * rpt
    // rows from there; tiles of a system are a multiple of rpt rows long (their
    // misalignment is the system's).  A single system is cut into a multiple of
    // the CU count so that every CU gets the same number of tiles.  In a batch,
    // a small system is cut into tiles of about kSmallTileRows rows (up to
    // kSmallTiles of them): once the large systems have converged, its
    // iterations run alone, and full tiles would leave all but a few CUs idle.
    constexpr int64_t kSmallTileRows = 2048, kSmallTiles = 32;
    const int64_t cap = int64_t(sub) * bs * rpt - (rpt - 1);
    gp.tiles.clear();
    gp.first.assign(m->nsys, 0);
    gp.count.assign(m->nsys, 0);
    for (int s = 0; s < m->nsys; s++) {
        const int64_t r0 = m->sys_row_begin[s], R = m->sys_row_begin[s + 1] - r0;
        gp.first[s] = int32_t(gp.tiles.size());
        int64_t n = (R + cap - 1) / cap;
        if (m->nsys == 1 && n > 0) n = (n + ctx.cus * per_cu - 1) / (ctx.cus * per_cu) * (ctx.cus * per_cu);
        if (m->nsys > 1) n = std::max(n, std::min((R + kSmallTileRows - 1) / kSmallTileRows, kSmallTiles));
        int64_t per = n > 0 ? (R + n - 1) / n : 0;
        per = std::min(cap / rpt * rpt, (per + rpt - 1) / rpt * rpt);
        for (int64_t t = 0; per > 0 && t * per < R; t++)
            gp.tiles.push_back(RowWork{r0 + t * per, r0 + std::min(R, (t + 1) * per), s, 0});
        gp.count[s] = int32_t(gp.tiles.size()) - gp.first[s];
    }
    gp.d_segs = to_device(gp.segs, ctx.stream);
    gp.d_steps = to_device(gp.steps, ctx.stream);
    gp.d_phases = to_device(gp.phases, ctx.stream);
    gp.d_tiles = to_device(gp.tiles, ctx.stream);
    gp.d_first = to_device(gp.first, ctx.stream);
    gp.d_count = to_device(gp.count, ctx.stream);
    gp.d_counters = to_device(std::vector<unsigned>(std::max(1, m->nsys), 0u), ctx.stream);
    gp.d_partials = DBuf(2 * sizeof(double) * std::max<size_t>(1, gp.tiles.size()));
    gp.ok = true;
}

void build_plans(cad_matrix* m)
{
    const Context& ctx = *m->ctx;
    // per-system column bases (block-diagonal batches); one system by default
    if (m->sys_row_begin.empty()) {
        m->nsys = 1;
        m->sys_row_begin = {0, m->rows};
        m->sys_col = {0};
        m->cols_per_sys = m->cols;
    }
    m->d_sys_col = to_device(m->sys_col, ctx.stream);
    // (a matrix-free matrix sizes its scatter partials as a u16 matrix would)
    build_scatter_plan(m->splan, m->slot_lo, m->slot_hi, m->sys_row_begin, m->sys_col, m->cols_per_sys,
                       m->mf ? 2 : m->ibytes, ctx, m->tune);
    build_gather_plan(m, m->gplan4, 4, m->tune.rpt4, m->tune.bs4, m->tune.sub4);
    build_gather_plan(m, m->gplan8, 8, m->tune.rpt8, m->tune.bs8, m->tune.sub8);

    GridPlan& g = m->grid;
    g.rows.clear();
    g.cols.clear();
    g.row_first.assign(m->nsys, 0);
    g.row_count.assign(m->nsys, 0);
    g.col_first.assign(m->nsys, 0);
    g.col_count.assign(m->nsys, 0);
    // row blocks of the row kernels: one row a thread for a small matrix (its
    // gathers use the row kernel, whose rows' loads are then all in flight)
    const int64_t rchunk = int64_t(kRowBlock) * (m->rows <= kRowKernelRows ? 1 : kRowsPerThread);
    const int64_t cchunk = int64_t(kRowBlock) * 2;  // enough column blocks to cover the CUs
    for (int s = 0; s < m->nsys; s++) {
        g.row_first[s] = int32_t(g.rows.size());
        for (int64_t r = m->sys_row_begin[s]; r < m->sys_row_begin[s + 1]; r += rchunk)
            g.rows.push_back(RowWork{r, std::min(m->sys_row_begin[s + 1], r + rchunk), s, 0});
        g.row_count[s] = int32_t(g.rows.size()) - g.row_first[s];
        g.col_first[s] = int32_t(g.cols.size());
        const int64_t c0 = m->sys_col[s], c1 = c0 + m->cols_per_sys;
        for (int64_t c = c0; c < c1; c += cchunk) g.cols.push_back(ColWork{c, std::min(c1, c + cchunk), s, c == c0});
        g.col_count[s] = int32_t(g.cols.size()) - g.col_first[s];
    }
    g.d_rows = to_device(g.rows, ctx.stream);
    g.d_cols = to_device(g.cols, ctx.stream);
    g.d_row_first = to_device(g.row_first, ctx.stream);
    g.d_row_count = to_device(g.row_count, ctx.stream);
    g.d_col_first = to_device(g.col_first, ctx.stream);
    g.d_col_count = to_device(g.col_count, ctx.stream);
    std::vector<unsigned> zeros(std::max(1, m->nsys), 0u);
    g.d_row_counters = to_device(zeros, ctx.stream);
    g.d_col_counters = to_device(zeros, ctx.stream);
    g.d_partials = DBuf(2 * sizeof(double) * std::max<size_t>(1, std::max(g.rows.size(), g.cols.size())));
    // max |y| per row block and per system, for the fixed-point FP32 scatter
    m->splan.d_ymax_part = DBuf(sizeof(unsigned) * std::max<size_t>(1, g.rows.size()));
    m->splan.d_ymax = DBuf(sizeof(unsigned) * size_t(std::max(1, m->nsys)));
}

}  // namespace cad
