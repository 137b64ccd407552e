// Device buffers, contexts, pattern groups, matrices and their launch plans.
#include "runtime.hpp"

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

// Tiles of (system rows) x (group window).  Row chunks are sized so that the
// LDS accumulation of a tile outweighs its window flush and the grid holds
// several tiles per CU.
std::vector<ScatterWork> make_scatter_work(const ScatterPlan& sp, const std::vector<int64_t>& sys_rows,
                                           const std::vector<uint32_t>& sys_col, uint32_t wmax, int cus,
                                           uint32_t& widest)
{
    std::vector<ScatterWork> out;
    widest = 0;
    const int nsys = int(sys_col.size());
    double total = 0;
    for (int s = 0; s < nsys; s++)
        for (auto& g : sp.groups) {
            const int64_t R = sys_rows[s + 1] - sys_rows[s];
            const uint32_t nwin = (g.col_hi - g.col_lo + wmax - 1) / wmax;
            total += double(R) * g.nslots * nwin;
        }
    const double target_items = 8.0 * std::max(cus, 1);
    const double per_item = std::max(1.0, total / target_items);
    for (int s = 0; s < nsys; s++) {
        const int64_t R = sys_rows[s + 1] - sys_rows[s];
        if (R <= 0) continue;
        for (size_t gi = 0; gi < sp.groups.size(); gi++) {
            const SlotGroup& g = sp.groups[gi];
            for (uint32_t wl = g.col_lo; wl < g.col_hi; wl += wmax) {
                const uint32_t wh = std::min(g.col_hi, wl + wmax);
                const uint32_t W = wh - wl;
                int64_t chunk = int64_t(per_item / g.nslots);
                chunk = std::max<int64_t>(chunk, 4 * int64_t(W) / std::max<int>(1, g.nslots));
                chunk = std::max<int64_t>(chunk, kScatterBlock);
                chunk = (chunk + kScatterBlock - 1) / kScatterBlock * kScatterBlock;
                for (int64_t r0 = 0; r0 < R; r0 += chunk) {
                    ScatterWork w;
                    w.row_begin = sys_rows[s] + r0;
                    w.row_end = sys_rows[s] + std::min(R, r0 + chunk);
                    w.win_lo = sys_col[s] + wl;
                    w.win_hi = sys_col[s] + wh;
                    w.group = int32_t(gi);
                    w.sys = s;
                    out.push_back(w);
                }
                widest = std::max(widest, W);
            }
        }
    }
    return out;
}

}  // namespace

void build_scatter_plan(ScatterPlan& sp, const std::vector<uint32_t>& lo, const std::vector<uint32_t>& hi,
                        const std::vector<int64_t>& sys_rows, const std::vector<uint32_t>& sys_col,
                        const Context& ctx)
{
    make_groups(lo, hi, sp);
    const size_t lds = std::min<size_t>(ctx.max_lds, 128 * 1024);
    const uint32_t w4 = uint32_t(lds / 4), w8 = uint32_t(lds / 8);
    sp.work4 = make_scatter_work(sp, sys_rows, sys_col, w4, ctx.cus, sp.win4);
    sp.work8 = make_scatter_work(sp, sys_rows, sys_col, w8, ctx.cus, sp.win8);
    sp.work_cnt = sp.work4;
    sp.win_cnt = sp.win4;
    sp.d_groups = to_device(sp.groups, ctx.stream);
    sp.d_slot_list = to_device(sp.slot_list, ctx.stream);
    sp.d_work4 = to_device(sp.work4, ctx.stream);
    sp.d_work8 = to_device(sp.work8, ctx.stream);
    sp.d_work_cnt = to_device(sp.work_cnt, ctx.stream);
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
    build_scatter_plan(m->splan, m->slot_lo, m->slot_hi, m->sys_row_begin, m->sys_col, ctx);

    GridPlan& g = m->grid;
    g.rows.clear();
    g.cols.clear();
    g.row_first.assign(m->nsys, 0);
    g.row_count.assign(m->nsys, 0);
    g.col_first.assign(m->nsys, 0);
    g.col_count.assign(m->nsys, 0);
    const int64_t rchunk = int64_t(kRowBlock) * kRowsPerThread;
    const int64_t cchunk = int64_t(kRowBlock) * 4;
    for (int s = 0; s < m->nsys; s++) {
        g.row_first[s] = int32_t(g.rows.size());
        for (int64_t r = m->sys_row_begin[s]; r < m->sys_row_begin[s + 1]; r += rchunk)
            g.rows.push_back(RowWork{r, std::min(m->sys_row_begin[s + 1], r + rchunk), s, 0});
        g.row_count[s] = int32_t(g.rows.size()) - g.row_first[s];
        g.col_first[s] = int32_t(g.cols.size());
        const int64_t c0 = m->sys_col[s], c1 = c0 + m->cols_per_sys;
        for (int64_t c = c0; c < c1; c += cchunk) g.cols.push_back(ColWork{c, std::min(c1, c + cchunk), s, 0});
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
}

}  // namespace cad
