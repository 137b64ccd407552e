// Products, Jacobi and the device-resident solvers (LSQR, mixed-precision
// refined LSQR, PCG / CG on A^T A).
#include "solver.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>

#include "kernels.hpp"

namespace cad {

// ------------------------------------------------------------ helpers -----
template <class F>
void with_view(cad_matrix* m, F&& f)
{
    if (m->mf)
        f(PositionsView{m->d_pos.as<uint64_t>(), m->d_slots.as<DevSlot>(), m->d_sq.as<uint8_t>(),
                        m->d_dense.as<uint32_t>(), m->d_tslots.as<DevTern>(), m->d_tern.as<uint32_t>(),
                        m->d_tw.as<uint4>()});
    else if (m->ibytes == 2)
        f(SlotMajor<uint16_t>{m->d_idx.as<uint16_t>(), m->ld, m->d_base.as<uint32_t>()});
    else
        f(SlotMajor<uint32_t>{m->d_idx.as<uint32_t>(), m->ld, m->d_base.as<uint32_t>()});
}

int grid_for(int64_t n, int bs = 256)
{
    int64_t g = (n + bs - 1) / bs;
    return int(std::max<int64_t>(1, std::min<int64_t>(g, 65536)));
}

template <class T>
void fill(T* p, int64_t n, T v, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(n)), dim3(256), 0, s, p, n, v);
    CAD_HIP(hipGetLastError());
}

// Raises the dynamic-LDS limit of a kernel (once per kernel and size; the
// solvers call prepare_scatter before they capture a graph, so this never
// runs inside a capture).
template <class K>
void allow_lds(K kernel, size_t bytes)
{
    static std::mutex mu;
    static std::map<const void*, size_t> done;
    const void* key = reinterpret_cast<const void*>(kernel);
    std::lock_guard<std::mutex> lock(mu);
    if (bytes > 64 * 1024 && bytes > done[key]) {
        (void)hipFuncSetAttribute(key, hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
        (void)hipGetLastError();
        done[key] = bytes;
    }
}

// The scatter windows: every real-valued scatter accumulates in 8-byte LDS
// words (FP64, or the 64-bit fixed point of the FP32 stored-index scatter), so
// it uses the 8-byte windows whatever its value type (f4 serves the u32
// multiplicity counts of the Jacobi diagonal).
template <class VT>
ScatterLists& lists(ScatterPlan& sp)
{
    return sp.f8;
}

// The partial buffer of a scatter list, sized for 8-byte elements whatever PT
// is: FP32 and FP64 scatters share the list, and a buffer grown for one
// precision would free the one a captured graph of the other still uses.
// (Allocated outside graph capture: every solver calls prepare_scatter first.)
template <class PT>
PT* partials(ScatterLists& L)
{
    const size_t bytes = sizeof(double) * size_t(std::max<int64_t>(1, L.part_elems));
    if (L.part.bytes < bytes) L.part = DBuf(bytes);
    return L.part.as<PT>();
}

// A^T y in two launches: LDS window tiles into partial slices, then the
// column reduction in chunk order (which writes every column of z).
// The first-stage scatter kernel of a view: vector loads of rpt (4, or 8 for
// u16 indices) consecutive rows for stored indices, 4 positions per step for
// a matrix-free matrix, one row per step otherwise; f receives the kernel.
// beta: the instantiation that forms an LSQR iteration's beta itself; ymax:
// the stored-index instantiation that forms each system's max |y| from the
// row-block maxima (FixedScale::ypart), used by FP32 PCG.
template <class VT, class YT, class View, class F>
void with_scatter_kernel(int rpt, bool beta, bool ymax, F&& f)
{
    auto pick = [&](auto kb, auto ky) {
        constexpr bool B = decltype(kb)::value, Y = decltype(ky)::value;
        if constexpr (std::is_same_v<View, SlotMajor<uint16_t>>) {
            if (rpt == 8) return f(k_scatter_tile<VT, YT, uint16_t, 8, B, Y>);
            return f(k_scatter_tile<VT, YT, uint16_t, 4, B, Y>);
        } else if constexpr (std::is_same_v<View, SlotMajor<uint32_t>>) {
            return f(k_scatter_tile<VT, YT, uint32_t, 4, B, Y>);
        } else if constexpr (std::is_same_v<View, PositionsView>) {
            return f(k_scatter_tile_mf<VT, YT, 4, B>);
        } else {
            return f(k_scatter_part<VT, YT, View, B>);
        }
    };
    if (beta) return pick(std::true_type{}, std::false_type{});
    if (ymax) return pick(std::false_type{}, std::true_type{});
    return pick(std::false_type{}, std::false_type{});
}

// A^T y in stages: the max |y| pre-pass of the fixed-point FP32 windows
// (outside an LSQR iteration), the LDS window tiles into partial slices, then
// the column reduction in chunk order (which writes every column of z).
template <class VT, class View>
constexpr bool fixed_point_scatter()
{
    return std::is_same_v<VT, float> &&
           (std::is_same_v<View, SlotMajor<uint16_t>> || std::is_same_v<View, SlotMajor<uint32_t>>);
}

template <class YT>
void launch_absmax(ScatterPlan& sp, const YT* y, const GridPlan& g, int nsys, hipStream_t s, bool systems = true)
{
    if (g.rows.empty()) return;
    hipLaunchKernelGGL(k_absmax_blocks<YT>, dim3(unsigned(g.rows.size())), dim3(kRowBlock), 0, s,
                       g.d_rows.as<RowWork>(), y, sp.d_ymax_part.as<unsigned>());
    if (!systems) {
        CAD_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_absmax_systems, dim3(unsigned(nsys)), dim3(kRowBlock), 0, s, g.d_row_first.as<int32_t>(),
                       g.d_row_count.as<int32_t>(), (const unsigned*)sp.d_ymax_part.as<unsigned>(),
                       sp.d_ymax.as<unsigned>());
    CAD_HIP(hipGetLastError());
}

// yrows: the row blocks whose maxima k_absmax_blocks left in sp.d_ymax_part,
// for the instantiation that forms max |y| itself (FP32 windows of a stored
// matrix only; null: read sp.d_ymax).
template <class VT, class YT, class View>
void launch_scatter_tiles(ScatterPlan& sp, const View& view, int rpt, const uint32_t* sys_col, const YT* y,
                          const LsqrScalars* div, int mode, hipStream_t s, bool y_padded = false,
                          const ActiveLists* al = nullptr, const GridSys& bsum = GridSys{},
                          const GridPlan* yrows = nullptr)
{
    ScatterLists& L = lists<VT>(sp);
    const ScatterWork* work = al ? al->scatter.as<ScatterWork>() : L.d_work.as<ScatterWork>();
    const int nwork = al ? al->nscatter : int(L.work.size());
    if (nwork == 0) return;
    // the fixed-point windows read max |y| per system outside an LSQR
    // iteration, whose terms are bounded by construction
    const bool ymax = yrows && fixed_point_scatter<VT, View>() && mode != kScatterLsqr;
    const FixedScale fx{fixed_point_scatter<VT, View>() && mode != kScatterLsqr ? sp.d_ymax.as<unsigned>() : nullptr,
                        al ? al->cbits : L.cbits, ymax ? sp.d_ymax_part.as<unsigned>() : nullptr,
                        ymax ? yrows->d_row_first.as<int32_t>() : nullptr,
                        ymax ? yrows->d_row_count.as<int32_t>() : nullptr};
    VT* part = al ? al->part.as<VT>() : partials<VT>(L);
    const size_t lds = size_t(L.widest) * sizeof(double);
    with_scatter_kernel<VT, YT, View>(rpt, bsum.partials != nullptr, ymax, [&](auto kern) {
        allow_lds(kern, lds);
        hipLaunchKernelGGL(kern, dim3(unsigned(nwork)), dim3(kScatterBlock), lds, s, view, work,
                           L.d_groups.as<SlotGroup>(), L.d_slot_list.as<uint16_t>(), sys_col, y,
                           const_cast<LsqrScalars*>(div), mode, fx, int(y_padded), part, bsum);
    });
    CAD_HIP(hipGetLastError());
}

template <class VT, class ZT>
void launch_scatter_reduce(ScatterPlan& sp, const LsqrScalars* div, int mode, ZT* z, hipStream_t s)
{
    ScatterLists& L = lists<VT>(sp);
    if (L.reduce.empty()) return;
    hipLaunchKernelGGL((k_scatter_reduce<VT, ZT>), dim3(unsigned(L.reduce.size())), dim3(kRowBlock), 0, s,
                       L.d_reduce.as<ReduceWork>(), (const VT*)partials<VT>(L), z, div, mode);
    CAD_HIP(hipGetLastError());
}

template <class VT, class YT, class ZT, class View>
void launch_scatter(ScatterPlan& sp, const View& view, int rpt, const uint32_t* sys_col, const YT* y,
                    const LsqrScalars* div, int mode, ZT* z, const GridPlan& g, int nsys, hipStream_t s,
                    bool y_padded = false)
{
    if (fixed_point_scatter<VT, View>() && mode != kScatterLsqr && !lists<VT>(sp).work.empty())
        launch_absmax(sp, y, g, nsys, s);
    launch_scatter_tiles<VT, YT>(sp, view, rpt, sys_col, y, div, mode, s, y_padded);
    launch_scatter_reduce<VT>(sp, div, mode, z, s);
}

template <class VT, class YT>
void prepare_scatter(cad_matrix* m)
{
    ScatterLists& L = lists<VT>(m->splan);
    partials<VT>(L);
    const size_t lds = size_t(L.widest) * sizeof(double);
    with_view(m, [&](auto view) {
        for (int v = 0; v < 3; v++)
            with_scatter_kernel<VT, YT, decltype(view)>(m->tune.srpt, v == 1, v == 2,
                                                        [&](auto kern) { allow_lds(kern, lds); });
    });
}

// A^T y for every system (div == nullptr), or, in an LSQR iteration, for the
// iterating systems with y scaled by 1 / beta.  y_padded: y is readable up
// to padded_rows(rows) (the library's own row vectors).
template <class VT, class YT, class ZT>
void scatter(cad_matrix* m, const YT* y, const LsqrScalars* div, ZT* z, hipStream_t s, bool y_padded = false)
{
    const int mode = div ? kScatterLsqr : kScatterAll;
    with_view(m, [&](auto view) {
        launch_scatter<VT, YT, ZT>(m->splan, view, m->tune.srpt, m->d_sys_col.as<uint32_t>(), y, div, mode, z,
                                   m->grid, m->nsys, s, y_padded);
    });
}

// A^T y for the systems being (re)started (active and restart in sel).
template <class VT, class YT, class ZT>
void scatter_restart(cad_matrix* m, const YT* y, const LsqrScalars* sel, ZT* z, hipStream_t s)
{
    with_view(m, [&](auto view) {
        launch_scatter<VT, YT, ZT>(m->splan, view, m->tune.srpt, m->d_sys_col.as<uint32_t>(), y, sel,
                                   kScatterRestart, z, m->grid, m->nsys, s);
    });
}

template <class VT>
GatherPlan& gather_plan(cad_matrix* m)
{
    return sizeof(VT) == 4 ? m->gplan4 : m->gplan8;
}

// The tile gather's rows per thread is a schedule knob: the instantiated
// values are 2, 4 (FP32; and 8 for u16 indices) and 1, 2 (FP64), the ones
// whose registers fit a 1024-thread block without spilling; f receives the
// value as a compile-time constant.
template <class VT, class View, class F>
void with_rpt(int rpt, F&& f)
{
    if constexpr (sizeof(VT) == 4) {
        if (rpt == 2) return f(std::integral_constant<int, 2>{});
        if (rpt == 4) return f(std::integral_constant<int, 4>{});
        if constexpr (std::is_same_v<typename View::index_type, uint16_t>)
            if (rpt == 8) return f(std::integral_constant<int, 8>{});
    } else {
        if (rpt == 1) return f(std::integral_constant<int, 1>{});
        if (rpt == 2) return f(std::integral_constant<int, 2>{});
    }
    throw Error(CAD_ERR_ARGUMENT, "unsupported rows per thread of the tile gather");
}

// Shape of the stored-index tile gather (schedule knobs): threads per block
// (kTileBlock, or half of it with two blocks per CU) and sub-tiles per staging
// (2 or 3 with full blocks, for the variants whose per-thread rows fit 8 bytes
// of entries: the others would spill); f receives both as compile-time
// constants.
template <bool kSubTiles, class F>
void with_tile_shape(int bs, int sub, F&& f)
{
    using std::integral_constant;
    if (bs == kTileBlock) {
        if (sub == 1) return f(integral_constant<int, kTileBlock>{}, integral_constant<int, 1>{});
        if constexpr (kSubTiles) {
            if (sub == 2) return f(integral_constant<int, kTileBlock>{}, integral_constant<int, 2>{});
            if (sub == 3) return f(integral_constant<int, kTileBlock>{}, integral_constant<int, 3>{});
        }
    } else if (bs == kTileBlock / 2 && sub == 1) {
        return f(integral_constant<int, kTileBlock / 2>{}, integral_constant<int, 1>{});
    }
    throw Error(CAD_ERR_ARGUMENT, "unsupported block size / sub-tiles of the tile gather");
}

// the tile gathers that take sub-tiles: FP32 with 2 rows per thread, FP64
// with 1, and at most 4 bytes of entries per slot and thread
template <class VT, class IdxT>
constexpr bool sub_tiles_ok(int rpt)
{
    return rpt * int(sizeof(VT)) == 8 && rpt * int(sizeof(IdxT)) <= 4;
}

// The row blocks a gather launches: the tile plan's or the grid plan's (all
// systems), or those of an ActiveLists built for this value size.
struct RowList {
    const RowWork* p;
    int n;
    GridSys sys;
};

template <class VT>
RowList row_list(cad_matrix* m, const ActiveLists* al)
{
    GatherPlan& gp = gather_plan<VT>(m);
    if (al) {
        const GridSys base = gp.ok && !al->row_kernel ? gp.sys() : m->grid.row_sys();
        return RowList{al->tiles.as<RowWork>(), al->ntiles,
                       GridSys{al->tfirst.as<int32_t>(), al->tcount.as<int32_t>(), base.counters,
                               al->row_kernel ? al->gpart.as<double>() : base.partials}};
    }
    if (gp.ok) return RowList{gp.d_tiles.as<RowWork>(), int(gp.tiles.size()), gp.sys()};
    return RowList{m->grid.d_rows.as<RowWork>(), int(m->grid.rows.size()), m->grid.row_sys()};
}

// Row gather with an epilogue: the LDS-staged tile kernel for pattern
// matrices, the generic row kernel otherwise.  al: launch only the blocks of
// the systems it holds (with the row kernel when it says so).  finish: false
// leaves the epilogue's block partials for the next launch to sum (the LSQR
// iteration's scatter forms beta from them).
template <class VT, class Epi>
void gather(cad_matrix* m, const VT* x, Epi epi, hipStream_t s, const ActiveLists* al = nullptr,
            bool finish = true)
{
    GatherPlan& gp = gather_plan<VT>(m);
    const RowList rl = row_list<VT>(m, al);
    if (rl.n == 0) return;
    with_view(m, [&](auto view) {
        using View = decltype(view);
        using IdxT = typename View::index_type;
        if (gp.ok && !(al && al->row_kernel)) {
            const size_t lds = size_t(gp.lds_elems) * sizeof(VT);
            with_rpt<VT, View>(gp.rpt, [&](auto rc) {
                constexpr int RPT = decltype(rc)::value;
                if constexpr (std::is_void_v<IdxT>) {
                    auto kern = k_gather_tile_mf<VT, RPT, Epi>;
                    allow_lds(kern, lds);
                    hipLaunchKernelGGL(kern, dim3(unsigned(rl.n)), dim3(kTileBlock), lds, s, view, rl.p,
                                       m->d_sys_col.as<uint32_t>(), gp.d_steps.as<GatherStep>(),
                                       int(gp.steps.size()), gp.d_phases.as<PhaseSeg>(), gp.wmax, x, epi, rl.sys);
                } else {
                    with_tile_shape<sub_tiles_ok<VT, IdxT>(RPT)>(gp.bs, gp.sub, [&](auto bc, auto sc) {
                        constexpr int BS = decltype(bc)::value;
                        auto kern = k_gather_tile<VT, RPT, IdxT, Epi, BS, decltype(sc)::value>;
                        allow_lds(kern, lds);
                        hipLaunchKernelGGL(kern, dim3(unsigned(rl.n)), dim3(BS), lds, s, view, rl.p,
                                           m->d_sys_col.as<uint32_t>(), gp.d_steps.as<GatherStep>(),
                                           int(gp.steps.size()), gp.d_phases.as<PhaseSeg>(), gp.wmax, x, epi,
                                           rl.sys);
                    });
                }
            });
        } else {
            hipLaunchKernelGGL((k_gather_rows<VT, decltype(view), Epi>), dim3(unsigned(rl.n)), dim3(kRowBlock), 0,
                               s, view, m->k, rl.p, m->d_sys_col.as<uint32_t>(), x, epi, rl.sys);
        }
    });
    CAD_HIP(hipGetLastError());
    // the epilogue's per-system finish, from the block partials the gather left
    if constexpr (finish_kernel<Epi>()) {
        if (finish) {
            hipLaunchKernelGGL(k_epi_finish<Epi>, dim3(unsigned(m->nsys)), dim3(kRowBlock), 0, s, epi, rl.sys);
            CAD_HIP(hipGetLastError());
        }
    }
}

template <class VT>
void spmv(cad_matrix* m, const VT* x, VT* y, hipStream_t s)
{
    gather<VT>(m, x, EpiStore<VT>{y}, s);
}

// ------------------------------------------------------------ products ----
void matrix_spmv_f32(cad_matrix* m, const float* x, float* y) { spmv<float>(m, x, y, m->ctx->stream); }
void matrix_spmv_f64(cad_matrix* m, const double* x, double* y) { spmv<double>(m, x, y, m->ctx->stream); }

void matrix_spmtv_f32(cad_matrix* m, const float* y, float* z)
{
    scatter<float, float, float>(m, y, nullptr, z, m->ctx->stream);
}

void matrix_spmtv_f64(cad_matrix* m, const double* y, double* z)
{
    scatter<double, double, double>(m, y, nullptr, z, m->ctx->stream);
}

void matrix_targets(cad_matrix* m, const double* xt, uint64_t seed, double sn, double* b)
{
    gather<double>(m, xt, EpiTargets{b, seed, sn}, m->ctx->stream);
}

// The first stage of the Jacobi counts: u32 multiplicity-squared windows
// into the partial slices of the 4-byte scatter lists.
static void launch_jacobi_part(cad_matrix* m, hipStream_t s)
{
    ScatterLists& L = m->splan.f4;
    uint32_t* part = partials<uint32_t>(L);
    if (L.work.empty()) return;
    const size_t lds = size_t(L.widest) * 4;
    with_view(m, [&](auto view) {
        auto kern = k_jacobi_part<decltype(view)>;
        allow_lds(kern, lds);
        hipLaunchKernelGGL(kern, dim3(unsigned(L.work.size())), dim3(kScatterBlock), lds, s, view,
                           L.d_work.as<ScatterWork>(), L.d_groups.as<SlotGroup>(), L.d_slot_list.as<uint16_t>(),
                           m->d_sys_col.as<uint32_t>(), part);
    });
    CAD_HIP(hipGetLastError());
}

const uint32_t* jacobi_counts(cad_matrix* m)
{
    if (m->have_cnt) return m->d_jacobi_cnt.as<uint32_t>();
    hipStream_t s = m->ctx->stream;
    m->d_jacobi_cnt = DBuf(sizeof(uint32_t) * std::max<int64_t>(1, m->cols));
    ScatterPlan& sp = m->splan;
    ScatterLists& L = sp.f4;
    uint32_t* part = partials<uint32_t>(L);
    launch_jacobi_part(m, s);
    hipLaunchKernelGGL((k_scatter_reduce<uint32_t, uint32_t>), dim3(unsigned(L.reduce.size())), dim3(kRowBlock), 0,
                       s, L.d_reduce.as<ReduceWork>(), (const uint32_t*)part, m->d_jacobi_cnt.as<uint32_t>(),
                       (const LsqrScalars*)nullptr, kScatterAll);
    CAD_HIP(hipGetLastError());
    m->have_cnt = true;
    return m->d_jacobi_cnt.as<uint32_t>();
}

void matrix_jacobi(cad_matrix* m, float* d)
{
    const uint32_t* cnt = jacobi_counts(m);
    hipLaunchKernelGGL(k_jacobi_finish, dim3(grid_for(m->cols)), dim3(256), 0, m->ctx->stream, cnt, m->cols, d);
    CAD_HIP(hipGetLastError());
}

// ------------------------------------------------------------ workspace ---
Workspace::~Workspace()
{
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (host_scal) (void)hipHostFree(host_scal);
}

static Workspace& workspace(cad_matrix* m, int key)
{
    auto& w = m->ws[key];
    if (!w) {
        w.reset(new Workspace);
        w->key = key;
        CAD_HIP(hipHostMalloc(&w->host_scal, sizeof(LsqrScalars) * std::max(1, m->nsys) * 2 + 4096));
    }
    return *w;
}

static void ensure(DBuf& b, size_t bytes)
{
    if (b.bytes < bytes) b = DBuf(bytes);
}

// Reads back the per-system scalars (synchronises the stream).
template <class S>
static const S* read_scalars(cad_matrix* m, Workspace& ws, const DBuf& d, hipStream_t s)
{
    const size_t bytes = sizeof(S) * m->nsys;
    CAD_HIP(hipMemcpyAsync(ws.host_scal, d.p, bytes, hipMemcpyDeviceToHost, s));
    CAD_HIP(hipStreamSynchronize(s));
    return static_cast<const S*>(ws.host_scal);
}

static void set_active(cad_matrix* m, Workspace& ws, DBuf& d, const std::vector<int>& active, hipStream_t s)
{
    std::vector<LsqrScalars> h(m->nsys);
    std::memset(h.data(), 0, sizeof(LsqrScalars) * m->nsys);
    for (int i = 0; i < m->nsys; i++) h[i].active = h[i].restart = active[i];
    CAD_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(LsqrScalars) * m->nsys, hipMemcpyHostToDevice, s));
    CAD_HIP(hipStreamSynchronize(s));
    (void)ws;
}

// Column scaling D = 1/|a_c| (1 for empty columns) or all ones.
static void column_scaling(cad_matrix* m, double* D, bool on, hipStream_t s)
{
    if (on) {
        const uint32_t* cnt = jacobi_counts(m);
        hipLaunchKernelGGL(k_colscale, dim3(grid_for(m->cols)), dim3(256), 0, s, cnt, m->cols, D);
        CAD_HIP(hipGetLastError());
    } else {
        fill<double>(D, m->cols, 1.0, s);
    }
}

// Explicit FP64 |A^T(b - A x)| and |b - A x| per system (one FP64 SpMV and one
// FP64 scatter).  Uses its own scalar block so that solver state is untouched.
static void explicit_residual(cad_matrix* m, Workspace& ws, const double* x, const double* b,
                              std::vector<double>& nres, std::vector<double>& rnorm, hipStream_t s)
{
    const GridPlan& g = m->grid;
    ensure(ws.r64, sizeof(double) * m->rows);
    ensure(ws.z64, sizeof(double) * m->cols);
    ensure(ws.scal_chk, sizeof(LsqrScalars) * m->nsys);
    std::vector<int> act(m->nsys, 1);
    set_active(m, ws, ws.scal_chk, act, s);
    LsqrScalars* sc = ws.scal_chk.as<LsqrScalars>();
    gather<double>(m, x, EpiResidual<double>{b, ws.r64.as<double>(), nullptr, sc}, s);
    scatter<double, double, double>(m, ws.r64.as<double>(), nullptr, ws.z64.as<double>(), s);
    hipLaunchKernelGGL(k_cols_norm2<double>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s,
                       g.d_cols.as<ColWork>(), ws.z64.as<double>(), g.col_sys(), sc);
    CAD_HIP(hipGetLastError());
    const LsqrScalars* h = read_scalars<LsqrScalars>(m, ws, ws.scal_chk, s);
    nres.resize(m->nsys);
    rnorm.resize(m->nsys);
    for (int i = 0; i < m->nsys; i++) {
        nres[i] = std::sqrt(h[i].reduce_b);
        rnorm[i] = std::sqrt(h[i].reduce_a);
    }
}

// |A^T b| per system in FP64.
static std::vector<double> atb_norms(cad_matrix* m, Workspace& ws, const double* b, hipStream_t s)
{
    const GridPlan& g = m->grid;
    ensure(ws.z64, sizeof(double) * m->cols);
    ensure(ws.scal_chk, sizeof(LsqrScalars) * m->nsys);
    std::vector<int> act(m->nsys, 1);
    set_active(m, ws, ws.scal_chk, act, s);
    scatter<double, double, double>(m, b, nullptr, ws.z64.as<double>(), s);
    hipLaunchKernelGGL(k_cols_norm2<double>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s,
                       g.d_cols.as<ColWork>(), ws.z64.as<double>(), g.col_sys(), ws.scal_chk.as<LsqrScalars>());
    CAD_HIP(hipGetLastError());
    const LsqrScalars* h = read_scalars<LsqrScalars>(m, ws, ws.scal_chk, s);
    std::vector<double> out(m->nsys);
    for (int i = 0; i < m->nsys; i++) out[i] = std::sqrt(h[i].reduce_b);
    return out;
}

// ------------------------------------------------------------ LSQR --------
template <class VT>
static void lsqr_iteration(cad_matrix* m, Workspace& ws, hipStream_t s, const ActiveLists* al = nullptr)
{
    const GridPlan& g = m->grid;
    LsqrScalars* sc = ws.scal.as<LsqrScalars>();
    // u_raw = A v - alpha u, leaving the block partials of |u_raw|^2, from
    // which the tiles of A^T (u / beta) form beta -- except where the scatter
    // instantiation with that prologue runs short of registers: on a
    // matrix-free matrix (its index programs need them; about 10% slower) and
    // with 8 rows a step (128 VGPRs and a spill): there the gather's finish
    // launch forms beta.  Then the column reduction fused with
    // vbar = D z - beta v (u is padded to whole 16-byte groups), and the
    // rotation with the column update.
    const bool fuse = !m->mf && m->tune.srpt != 8;
    gather<VT>(m, ws.vd.as<VT>(), EpiLsqrU<VT>{ws.u.as<VT>(), sc, VT(0)}, s, al, !fuse);
    ScatterLists& L = lists<VT>(m->splan);
    const GridSys usys = fuse ? row_list<VT>(m, al).sys : GridSys{};
    with_view(m, [&](auto view) {
        launch_scatter_tiles<VT, VT>(m->splan, view, m->tune.srpt, m->d_sys_col.as<uint32_t>(), ws.u.as<VT>(), sc,
                                     kScatterLsqr, s, true, al, usys);
    });
    const GridSys rsys = al ? GridSys{al->rfirst.as<int32_t>(), al->rcount.as<int32_t>(),
                                      L.reduce_sys().counters, L.reduce_sys().partials}
                            : L.reduce_sys();
    const ReduceWork* red = al ? al->reduce.as<ReduceWork>() : L.d_reduce.as<ReduceWork>();
    const int nred = al ? al->nreduce : int(L.reduce.size());
    if (nred > 0)
        hipLaunchKernelGGL(k_lsqr_reduce_v<VT>, dim3(unsigned(nred)), dim3(kRowBlock), 0, s, red,
                           al ? al->part.as<const VT>() : (const VT*)partials<VT>(L), ws.D.as<double>(),
                           ws.v.as<double>(), rsys, sc);
    const ColWork* cw = al ? al->cols.as<ColWork>() : g.d_cols.as<ColWork>();
    const int ncw = al ? al->ncols : int(g.cols.size());
    if (ncw > 0)
        hipLaunchKernelGGL(k_lsqr_step<VT>, dim3(unsigned(ncw)), dim3(kRowBlock), 0, s, cw, rsys, ws.D.as<double>(),
                           ws.v.as<double>(), ws.vd.as<VT>(), ws.w.as<double>(), ws.y.as<double>(), sc);
    CAD_HIP(hipGetLastError());
}

// The iteration lists of the systems in `on` (see ActiveLists).
template <class VT>
static void build_active(cad_matrix* m, ActiveLists& al, const std::vector<int>& on, hipStream_t s)
{
    const int n = m->nsys;
    GatherPlan& gp = gather_plan<VT>(m);
    // A tile of the staged gather costs about the same whatever its rows (it
    // stages every x window of the pattern set into LDS): once the running
    // systems are small, the row kernel (x read through L2, the same row sums
    // in the same order) finishes sooner.
    int64_t running_rows = 0;
    for (int i = 0; i < n; i++)
        if (on[i]) running_rows += m->sys_row_begin[i + 1] - m->sys_row_begin[i];
    al.row_kernel = !gp.ok || running_rows <= kRowKernelRows;
    // the row kernel's blocks: one row a thread (its rows' loads are then all
    // in flight at once), with partials of their own
    std::vector<RowWork> fine;
    std::vector<int32_t> fine_first(n, 0), fine_count(n, 0);
    if (al.row_kernel)
        for (int i = 0; i < n; i++) {
            fine_first[i] = int32_t(fine.size());
            if (on[i])
                for (int64_t r = m->sys_row_begin[i]; r < m->sys_row_begin[i + 1]; r += kRowBlock)
                    fine.push_back(RowWork{r, std::min(m->sys_row_begin[i + 1], r + int64_t(kRowBlock)), i, 0});
            fine_count[i] = int32_t(fine.size()) - fine_first[i];
        }
    const std::vector<RowWork>& tiles = al.row_kernel ? fine : gp.tiles;
    const std::vector<int32_t>& tf = al.row_kernel ? fine_first : gp.first;
    const std::vector<int32_t>& tc = al.row_kernel ? fine_count : gp.count;
    std::vector<RowWork> t;
    std::vector<int32_t> t_first(n, 0), t_count(n, 0);
    std::vector<ColWork> cw;
    al.on.assign(n, 0);
    al.rows = 0;
    for (int i = 0; i < n; i++) {
        if (!on[i]) continue;
        al.on[i] = 1;
        al.rows += m->sys_row_begin[i + 1] - m->sys_row_begin[i];
        t_first[i] = int32_t(t.size());
        t_count[i] = tc[i];
        t.insert(t.end(), tiles.begin() + tf[i], tiles.begin() + tf[i] + tc[i]);
        cw.insert(cw.end(), m->grid.cols.begin() + m->grid.col_first[i],
                  m->grid.cols.begin() + m->grid.col_first[i] + m->grid.col_count[i]);
    }
    // the window tiles planned for these systems alone: how many row chunks
    // a window needs to fill the CUs depends on how many windows run side by
    // side (few in the tail of a batch, where the systems left are small)
    ScatterLists sub;
    plan_scatter_subset(m->splan, m->sys_row_begin, m->sys_col, m->cols_per_sys, m->mf ? 2 : m->ibytes, *m->ctx,
                        m->tune, al.on, sub);
    al.ntiles = int(t.size());
    al.nscatter = int(sub.work.size());
    al.nreduce = int(sub.reduce.size());
    al.ncols = int(cw.size());
    al.tiles = to_device(t, s);
    if (al.row_kernel) al.gpart = DBuf(2 * sizeof(double) * std::max<size_t>(1, t.size()));
    al.tfirst = to_device(t_first, s);
    al.tcount = to_device(t_count, s);
    al.scatter = to_device(sub.work, s);
    al.part = DBuf(sizeof(double) * size_t(std::max<int64_t>(1, sub.part_elems)));
    al.cbits = sub.cbits;
    al.reduce = to_device(sub.reduce, s);
    al.rfirst = to_device(sub.red_first, s);
    al.rcount = to_device(sub.red_count, s);
    al.cols = to_device(cw, s);
    CAD_HIP(hipStreamSynchronize(s));  // the host vectors go out of scope
    al.gen++;
}

// Before a run of iterations of a batch: launch only the systems still
// iterating once they hold less than half the rows of the lists in use (and
// always when a running system is missing from the active lists).
template <class VT>
static const ActiveLists* active_lists(cad_matrix* m, Workspace& ws, const std::vector<int>& running, hipStream_t s)
{
    if (m->nsys <= 1) return nullptr;
    int64_t rows = 0;
    bool missing = false;
    for (int i = 0; i < m->nsys; i++) {
        if (!running[i]) continue;
        rows += m->sys_row_begin[i + 1] - m->sys_row_begin[i];
        if (ws.use_active && !ws.active.on[i]) missing = true;
    }
    const int64_t in_use = ws.use_active ? ws.active.rows : m->rows;
    if (missing || 2 * rows < in_use) {
        build_active<VT>(m, ws.active, running, s);
        ws.use_active = true;
    }
    return ws.use_active ? &ws.active : nullptr;
}

// Kernel attributes of a gather with epilogue Epi (set outside graph capture).
template <class VT, class Epi>
static void prepare_gather(cad_matrix* m)
{
    GatherPlan& gp = gather_plan<VT>(m);
    if (!gp.ok) return;
    with_view(m, [&](auto view) {
        using View = decltype(view);
        using IdxT = typename View::index_type;
        with_rpt<VT, View>(gp.rpt, [&](auto rc) {
            constexpr int RPT = decltype(rc)::value;
            if constexpr (std::is_void_v<IdxT>)
                allow_lds(k_gather_tile_mf<VT, RPT, Epi>, size_t(gp.lds_elems) * sizeof(VT));
            else
                with_tile_shape<sub_tiles_ok<VT, IdxT>(RPT)>(gp.bs, gp.sub, [&](auto bc, auto sc) {
                    allow_lds(k_gather_tile<VT, RPT, IdxT, Epi, decltype(bc)::value, decltype(sc)::value>,
                              size_t(gp.lds_elems) * sizeof(VT));
                });
        });
    });
}

// Runs n LSQR iterations, replaying a captured graph of `batch` iterations
// where possible (al: the active lists the iterations launch, or all).
template <class VT>
static void lsqr_run(cad_matrix* m, Workspace& ws, int n, int batch, bool graph, hipStream_t s,
                     const ActiveLists* al = nullptr)
{
    const int gen = al ? al->gen : -1;
    if (graph && batch > 1 && s != nullptr) {
        if (!ws.exec || ws.graph_iters != batch || ws.graph_prec != int(sizeof(VT)) || ws.graph_gen != gen) {
            if (ws.exec) (void)hipGraphExecDestroy(ws.exec);
            if (ws.graph) (void)hipGraphDestroy(ws.graph);
            ws.exec = nullptr;
            ws.graph = nullptr;
            prepare_scatter<VT, VT>(m);
            prepare_gather<VT, EpiLsqrU<VT>>(m);
            CAD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int i = 0; i < batch; i++) lsqr_iteration<VT>(m, ws, s, al);
            CAD_HIP(hipStreamEndCapture(s, &ws.graph));
            CAD_HIP(hipGraphInstantiate(&ws.exec, ws.graph, nullptr, nullptr, 0));
            ws.graph_iters = batch;
            ws.graph_prec = int(sizeof(VT));
            ws.graph_gen = gen;
        }
        while (n >= batch) {
            CAD_HIP(hipGraphLaunch(ws.exec, s));
            n -= batch;
        }
    }
    for (int i = 0; i < n; i++) lsqr_iteration<VT>(m, ws, s, al);
}

// LSQR start from the current FP64 residual r = b - A x (x64): u = r / |r|,
// v = D A^T r / |D A^T r|, all of A^T r computed in FP64.  Returns
// |A^T r| per system via reduce_b.
template <class VT>
static void lsqr_start_from_residual(cad_matrix* m, Workspace& ws, const double* b, hipStream_t s)
{
    const GridPlan& g = m->grid;
    LsqrScalars* sc = ws.scal.as<LsqrScalars>();
    gather<double>(m, ws.x64.as<double>(), EpiResidual<VT>{b, ws.r64.as<double>(), ws.u.as<VT>(), sc}, s);
    scatter_restart<double, double, double>(m, ws.r64.as<double>(), sc, ws.z64.as<double>(), s);
    hipLaunchKernelGGL(k_lsqr_init_v<double>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s,
                       g.d_cols.as<ColWork>(), ws.D.as<double>(), ws.z64.as<double>(), ws.v.as<double>(), g.col_sys(),
                       sc);
    hipLaunchKernelGGL(k_lsqr_init_w<VT>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s,
                       g.d_cols.as<ColWork>(), ws.D.as<double>(), ws.v.as<double>(), ws.vd.as<VT>(),
                       ws.w.as<double>(), ws.y.as<double>(), (const LsqrScalars*)sc);
    CAD_HIP(hipGetLastError());
}

template <class VT>
static void lsqr_alloc(cad_matrix* m, Workspace& ws, hipStream_t s)
{
    ensure(ws.u, sizeof(VT) * size_t(padded_rows(m->rows)));  // padded: the scatter reads it in 16-byte groups
    ensure(ws.r64, sizeof(double) * m->rows);
    ensure(ws.v, sizeof(double) * m->cols);
    ensure(ws.w, sizeof(double) * m->cols);
    ensure(ws.y, sizeof(double) * m->cols);
    ensure(ws.x64, sizeof(double) * m->cols);
    ensure(ws.D, sizeof(double) * m->cols);
    ensure(ws.z64, sizeof(double) * m->cols);
    ensure(ws.scal, sizeof(LsqrScalars) * m->nsys);
    if (ws.vd.bytes < sizeof(VT) * m->cols || ws.vd_prec != int(sizeof(VT))) {
        ws.vd = DBuf(sizeof(VT) * m->cols);
        ws.z = DBuf(sizeof(VT) * m->cols);
        ws.vd_prec = int(sizeof(VT));
        fill<VT>(ws.z.as<VT>(), m->cols, VT(0), s);
    }
}

// LSQR in one precision (FP64 or FP32 products and vectors) with column
// scaling; stops on the explicit FP64 metric.
template <class VT>
static void lsqr_plain(cad_matrix* m, const double* b, double* x, const cad_solve_options& o, cad_solve_report& rep,
                       SystemDetail& det)
{
    hipStream_t s = m->ctx->stream;
    Workspace& ws = workspace(m, 100 + int(sizeof(VT)));
    lsqr_alloc<VT>(m, ws, s);
    column_scaling(m, ws.D.as<double>(), o.precondition != 0, s);
    const std::vector<double> atb = atb_norms(m, ws, b, s);
    std::vector<int> act(m->nsys, 1);
    set_active(m, ws, ws.scal, act, s);
    fill<double>(ws.x64.as<double>(), m->cols, 0.0, s);
    // x = 0: the residual start is u = b, exactly Math/Solver.h:146-150.
    lsqr_start_from_residual<VT>(m, ws, b, s);

    const int batch = std::max(1, o.check_every);
    std::vector<double> trigger(m->nsys, 1.0), nres, rn;
    std::vector<int> frozen_at(m->nsys, -1);
    int it = 0;
    bool done = (o.max_iterations <= 0);
    while (!done) {
        const int n = std::min(batch, o.max_iterations - it);
        lsqr_run<VT>(m, ws, n, batch, o.use_graph != 0 && n == batch, s);
        it += n;
        const LsqrScalars* h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
        bool want_check = false;
        for (int i = 0; i < m->nsys; i++) {
            if (!h[i].active) continue;
            const double est = h[i].nres0 > 0 ? h[i].nres / h[i].nres0 : 0.0;
            if (!std::isfinite(est)) throw Error(CAD_ERR_BREAKDOWN, "LSQR breakdown (non-finite recurrence)");
            if (o.tol > 0 && est <= o.tol * trigger[i]) want_check = true;
        }
        if (o.verbose && m->nsys == 1)
            std::fprintf(stderr, "[lsqr %s] it %d est %.3e\n", sizeof(VT) == 4 ? "fp32" : "fp64", it,
                         h[0].nres0 > 0 ? h[0].nres / h[0].nres0 : 0.0);
        if (want_check || it >= o.max_iterations) {
            hipLaunchKernelGGL(k_scale_mul, dim3(grid_for(m->cols)), dim3(256), 0, s, ws.D.as<double>(),
                               ws.y.as<double>(), m->cols, ws.x64.as<double>());
            CAD_HIP(hipGetLastError());
            explicit_residual(m, ws, ws.x64.as<double>(), b, nres, rn, s);
            rep.explicit_checks++;
            bool all = true;
            std::vector<int> still(m->nsys, 0);
            h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
            for (int i = 0; i < m->nsys; i++) {
                const double rel = atb[i] > 0 ? nres[i] / atb[i] : nres[i];
                if (o.tol > 0 && rel <= o.tol) continue;
                if (h[i].active) {
                    still[i] = 1;
                    all = false;
                    const double est = h[i].nres0 > 0 ? h[i].nres / h[i].nres0 : 0.0;
                    if (o.tol > 0 && est <= o.tol * trigger[i]) trigger[i] *= std::min(0.5, o.tol / std::max(rel, 1e-300));
                }
            }
            if (all || it >= o.max_iterations) done = true;
            else {
                // freeze converged systems
                for (int i = 0; i < m->nsys; i++)
                    if (!still[i] && h[i].active) {
                        frozen_at[i] = it;
                        LsqrScalars t = h[i];
                        t.active = 0;
                        CAD_HIP(hipMemcpyAsync(ws.scal.as<LsqrScalars>() + i, &t, sizeof(t), hipMemcpyHostToDevice, s));
                    }
                CAD_HIP(hipStreamSynchronize(s));
            }
        }
    }
    if (nres.empty()) {  // no iteration ran: x = 0
        fill<double>(ws.x64.as<double>(), m->cols, 0.0, s);
        explicit_residual(m, ws, ws.x64.as<double>(), b, nres, rn, s);
    }
    CAD_HIP(hipMemcpyAsync(x, ws.x64.p, sizeof(double) * m->cols, hipMemcpyDeviceToDevice, s));
    rep.iterations = it;
    double worst = 0, rsum = 0;
    det.rel.assign(m->nsys, 0.0);
    det.iters.assign(m->nsys, it);
    for (int i = 0; i < m->nsys; i++) {
        det.rel[i] = atb[i] > 0 ? nres[i] / atb[i] : nres[i];
        if (frozen_at[i] >= 0) det.iters[i] = frozen_at[i];
        worst = std::max(worst, det.rel[i]);
        rsum += rn[i] * rn[i];
    }
    rep.rel_normal_residual = worst;
    rep.residual_norm = std::sqrt(rsum);
    rep.converged = o.tol > 0 && worst <= o.tol;
}

// State of a refined LSQR solve shared by its precision stages.
struct RefineState {
    std::vector<double> atb, rel, rn2;
    std::vector<int> act;      // still to solve
    std::vector<int> stalled;  // left by a stage that stopped making progress
    int it = 0, refinements = 0;
};

// The smallest inner LSQR reduction asked of FP32 products (they stall near
// 1e-7 relative; an inner_tol below it is raised to it).
constexpr double kFp32InnerFloor = 1e-7;

// One precision stage of the refined LSQR, from the FP64 x in ws.x64.  Every
// system runs its own refinement loop -- r = b - A x and A^T r in FP64, an
// inner LSQR with VT products on min |A d - r| that stops on the device once
// its estimate fell by the system's target, x += d -- and the systems of a
// batch advance independently: at each host check the systems whose inner
// solve ended are restarted (apply, FP64 residual, new start, selected by the
// restart flag) while the others keep iterating in the same launches.  A
// system leaves the stage when it meets tol (act = 0), when a refinement made
// too little progress (stalled = 1) or after max_refinements.  Too little is
// a cut of the metric by less than 4x with FP32 products (they cannot resolve
// the residual further: the FP64 stage takes over) and by less than 10 %
// with FP64 products.  A refinement that did not lower the metric at all is
// undone (x goes back to the iterate before it), so a stage never hands on a
// worse x than it got.
template <class VT>
static void refine_stage(cad_matrix* m, Workspace& ws, const double* b, const cad_solve_options& o, double floor_tol,
                         int max_refinements, RefineState& st, SystemDetail& det, cad_solve_report& rep)
{
    hipStream_t s = m->ctx->stream;
    const int n = m->nsys;
    const int batch = std::max(1, o.check_every);
    const double keep_factor = floor_tol > 0 ? 0.25 : 0.9;  // stall unless rel <= keep_factor * prev
    const char* prec = sizeof(VT) == 4 ? "fp32" : "fp64";
    std::vector<double> target(n, 0.0), prev(n, 0.0), prev_rn2(n, 0.0), best(n, 1.0);
    std::vector<int> refs(n, 0), stall(n, 0), checks(n, 0), running(n, 0);
    std::vector<int> restart = st.act;
    ensure(ws.xprev, sizeof(double) * m->cols);
    ws.use_active = false;  // every system's blocks until most have finished
    set_active(m, ws, ws.scal, st.act, s);  // all scalars zero, restart = active
    const LsqrScalars* h = nullptr;
    bool first = true;
    auto any_of = [&](const std::vector<int>& v) {
        for (int i = 0; i < n; i++)
            if (v[i]) return true;
        return false;
    };
    // writes the flags: active = running or selected, restart = selected
    auto write_flags = [&](const std::vector<int>& rs) {
        std::vector<LsqrScalars> hv(n);
        if (h) std::memcpy(hv.data(), h, sizeof(LsqrScalars) * n);
        else std::memset(hv.data(), 0, sizeof(LsqrScalars) * n);
        for (int i = 0; i < n; i++) {
            hv[i].active = (running[i] || rs[i]) ? 1 : 0;
            hv[i].restart = rs[i];
        }
        CAD_HIP(hipMemcpyAsync(ws.scal.p, hv.data(), sizeof(LsqrScalars) * n, hipMemcpyHostToDevice, s));
    };
    auto cols_kernel = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(unsigned(m->grid.cols.size())), dim3(kRowBlock), 0, s,
                           m->grid.d_cols.as<ColWork>(), ws.D.as<double>(), ws.y.as<double>(), ws.x64.as<double>(),
                           ws.xprev.as<double>(), (const LsqrScalars*)ws.scal.as<LsqrScalars>());
        CAD_HIP(hipGetLastError());
    };
    // x += D y for the systems in rs (saving x first), then the FP64 residual
    // start of all of them; returns the new metric in st.rel / st.rn2 and
    // undoes, for the systems whose refinement (refs > 0) did not lower the
    // metric, the step just applied.  undone[i] = 1 for those.
    auto apply_and_measure = [&](const std::vector<int>& rs, bool apply, std::vector<int>& undone) {
        undone.assign(n, 0);
        if (apply) {
            write_flags(rs);
            cols_kernel(k_lsqr_apply);
        }
        lsqr_start_from_residual<VT>(m, ws, b, s);
        h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
        rep.explicit_checks++;
        for (int i = 0; i < n; i++) {
            if (!rs[i]) continue;
            st.rel[i] = st.atb[i] > 0 ? std::sqrt(h[i].reduce_b) / st.atb[i] : std::sqrt(h[i].reduce_b);
            st.rn2[i] = h[i].reduce_a;
            if (apply && refs[i] > 0 && st.rel[i] >= prev[i] && !(o.tol > 0 && st.rel[i] <= o.tol)) {
                undone[i] = 1;
                if (o.verbose)
                    std::fprintf(stderr, "[refine %s] sys %d it %d: rel %.3e not below %.3e, step undone\n", prec, i,
                                 st.it, st.rel[i], prev[i]);
                st.rel[i] = prev[i];
                st.rn2[i] = prev_rn2[i];
            }
        }
        if (any_of(undone)) {
            std::vector<int> keep_running = running;
            std::fill(running.begin(), running.end(), 0);
            write_flags(undone);
            cols_kernel(k_lsqr_undo);
            running = keep_running;
        }
    };
    std::vector<int> undone;
    for (;;) {
        if (any_of(restart)) {
            apply_and_measure(restart, !first, undone);
            for (int i = 0; i < n; i++) {
                if (!restart[i]) continue;
                const char* why = nullptr;
                if (o.tol > 0 && st.rel[i] <= o.tol) {
                    st.act[i] = 0;
                    det.iters[i] = st.it;
                    why = "converged";
                } else if (undone[i] || (refs[i] > 0 && st.rel[i] > keep_factor * prev[i])) {
                    st.stalled[i] = 1;
                    why = "stalled";
                } else if (refs[i] >= max_refinements) {
                    why = "out of refinements";  // left to the caller: not converged
                } else {
                    // the last refinement only needs to close the remaining gap
                    // (with a factor 3 of margin); products in VT cannot go much
                    // below floor_tol in one solve
                    target[i] = o.tol > 0 ? std::max(floor_tol, o.tol / (3.0 * st.rel[i])) : floor_tol;
                    prev[i] = st.rel[i];
                    prev_rn2[i] = st.rn2[i];
                    refs[i]++;
                    st.refinements = std::max(st.refinements, refs[i]);
                    running[i] = 1;
                    best[i] = 1.0;
                    stall[i] = checks[i] = 0;
                }
                if (o.verbose)
                    std::fprintf(stderr, "[refine %s] sys %d it %d rel %.3e: %s\n", prec, i, st.it, st.rel[i],
                                 why ? why : "refine");
            }
            // the started systems iterate; targets on the device
            std::vector<LsqrScalars> hv(h, h + n);
            for (int i = 0; i < n; i++) {
                hv[i].active = running[i];
                hv[i].restart = 0;
                if (restart[i]) {
                    hv[i].target = running[i] ? target[i] : 0.0;
                    hv[i].stop = 0;
                }
            }
            CAD_HIP(hipMemcpyAsync(ws.scal.p, hv.data(), sizeof(LsqrScalars) * n, hipMemcpyHostToDevice, s));
            first = false;
            std::fill(restart.begin(), restart.end(), 0);
        }
        if (!any_of(running) || st.it >= o.max_iterations) break;
        const int k = std::min(batch, o.max_iterations - st.it);
        lsqr_run<VT>(m, ws, k, batch, o.use_graph != 0 && k == batch, s, active_lists<VT>(m, ws, running, s));
        st.it += k;
        h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
        // systems whose inner solve ended: met the target on the device, or
        // stuck (the estimate of |A^T r| is not monotone in LSQR and can
        // plateau, so only a best estimate that did not improve by 1% for as
        // many checks as the solve has run, at least 32, counts as stuck)
        for (int i = 0; i < n; i++) {
            if (!running[i]) continue;
            const double est = h[i].nres0 > 0 ? h[i].nres / h[i].nres0 : 0.0;
            if (!std::isfinite(est)) throw Error(CAD_ERR_BREAKDOWN, "LSQR breakdown (non-finite recurrence)");
            checks[i]++;
            if (est < 0.99 * best[i]) {
                best[i] = est;
                stall[i] = 0;
            } else {
                stall[i]++;
            }
            const bool stuck = stall[i] >= std::max(32, checks[i] / 2 + 1) && stall[i] * 2 >= checks[i];
            if (h[i].stop != 0 || est <= target[i] || stuck) {
                if (o.verbose)
                    std::fprintf(stderr, "[refine %s] sys %d it %d: inner solve ended (est %.3e, %s)\n", prec, i,
                                 st.it, est, stuck && h[i].stop == 0 && est > target[i] ? "stuck" : "target");
                running[i] = 0;
                restart[i] = 1;
            }
        }
    }
    // out of iterations with inner solves in flight: keep their corrections
    // (unless the metric rose) and measure where they got
    if (any_of(running)) {
        restart = running;
        std::fill(running.begin(), running.end(), 0);
        apply_and_measure(restart, true, undone);
        for (int i = 0; i < n; i++)
            if (restart[i] && o.tol > 0 && st.rel[i] <= o.tol) {
                st.act[i] = 0;
                det.iters[i] = st.it;
            }
    }
}

// Refined LSQR.  MIXED: FP32 products inside LSQR, FP64 residual
// r = b - A x and A^T r outside; systems whose FP32 refinement stops
// converging (FP32 rounding of A^T u swamps a residual that is tiny against
// |A| |r|) finish with FP64 products from the same x.  FP64: the FP64 stage
// alone from x = 0, whose first inner solve is plain LSQR (Math/Solver.h:146-
// 176) stopped on the device at tol / 3, then confirmed (or refined) by the
// explicit FP64 metric.
static void lsqr_refined(cad_matrix* m, const double* b, double* x, const cad_solve_options& o,
                         cad_solve_report& rep, SystemDetail& det, bool mixed)
{
    hipStream_t s = m->ctx->stream;
    Workspace& ws = workspace(m, mixed ? 200 : 201);
    if (mixed) lsqr_alloc<float>(m, ws, s);
    else lsqr_alloc<double>(m, ws, s);
    column_scaling(m, ws.D.as<double>(), o.precondition != 0, s);
    RefineState st;
    st.atb = atb_norms(m, ws, b, s);
    st.rel.assign(m->nsys, 1.0);
    st.rn2.assign(m->nsys, 0.0);
    st.act.assign(m->nsys, 1);
    st.stalled.assign(m->nsys, 0);
    det.iters.assign(m->nsys, -1);
    fill<double>(ws.x64.as<double>(), m->cols, 0.0, s);
    double* xs = ws.x64.as<double>();
    if (mixed) {
        // an inner target FP32 products cannot reach would leave every inner
        // solve iterating on rounding noise until the iteration budget ends
        refine_stage<float>(m, ws, b, o, std::max(o.inner_tol, kFp32InnerFloor), o.max_refinements, st, det, rep);
        bool finish = false;  // systems the FP32 stage left unconverged
        for (int i = 0; i < m->nsys; i++) finish |= st.act[i] != 0;
        if (finish && st.it < o.max_iterations) {
            Workspace& w8 = workspace(m, 201);
            lsqr_alloc<double>(m, w8, s);
            CAD_HIP(hipMemcpyAsync(w8.D.p, ws.D.p, sizeof(double) * m->cols, hipMemcpyDeviceToDevice, s));
            CAD_HIP(hipMemcpyAsync(w8.x64.p, ws.x64.p, sizeof(double) * m->cols, hipMemcpyDeviceToDevice, s));
            for (int i = 0; i < m->nsys; i++) st.stalled[i] = 0;
            refine_stage<double>(m, w8, b, o, 0.0, o.max_refinements, st, det, rep);
            xs = w8.x64.as<double>();
        }
    } else {
        refine_stage<double>(m, ws, b, o, 0.0, o.max_refinements, st, det, rep);
    }
    CAD_HIP(hipMemcpyAsync(x, xs, sizeof(double) * m->cols, hipMemcpyDeviceToDevice, s));
    double worst = 0, rsum = 0;
    for (int i = 0; i < m->nsys; i++) {
        worst = std::max(worst, st.rel[i]);
        rsum += st.rn2[i];
        if (det.iters[i] < 0) det.iters[i] = st.it;
    }
    det.rel = st.rel;
    rep.iterations = st.it;
    rep.refinements = st.refinements;
    rep.rel_normal_residual = worst;
    rep.residual_norm = std::sqrt(rsum);
    rep.converged = o.tol > 0 && worst <= o.tol;
}

// ------------------------------------------------------------ PCG ---------
template <class VT>
static void pcg_iteration(cad_matrix* m, Workspace& ws, hipStream_t s)
{
    const GridPlan& g = m->grid;
    PcgScalars* sc = ws.pscal.as<PcgScalars>();
    spmv<VT>(m, ws.p.as<VT>(), ws.t.as<VT>(), s);
    // A^T t: the window tiles (after the row blocks' max |t| of the
    // fixed-point windows; each tile takes its system's max of those), then
    // the column sums with the partials of p.q
    ScatterLists& L = lists<VT>(m->splan);
    with_view(m, [&](auto view) {
        const bool fixed = fixed_point_scatter<VT, decltype(view)>() && !L.work.empty();
        if (fixed) launch_absmax(m->splan, ws.t.as<VT>(), g, m->nsys, s, false);
        launch_scatter_tiles<VT, VT>(m->splan, view, m->tune.srpt, m->d_sys_col.as<uint32_t>(), ws.t.as<VT>(),
                                     nullptr, kScatterAll, s, false, nullptr, GridSys{}, fixed ? &g : nullptr);
    });
    if (!L.reduce.empty())
        hipLaunchKernelGGL(k_pcg_reduce_q<VT>, dim3(unsigned(L.reduce.size())), dim3(kRowBlock), 0, s,
                           L.d_reduce.as<ReduceWork>(), (const VT*)partials<VT>(L), ws.p.as<VT>(), ws.q.as<VT>(),
                           L.reduce_sys(), (const PcgScalars*)sc);
    hipLaunchKernelGGL(k_pcg_update<VT>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s,
                       g.d_cols.as<ColWork>(), ws.P.as<VT>(), ws.p.as<VT>(), ws.q.as<VT>(), ws.xr.as<VT>(),
                       ws.rr.as<VT>(), ws.zz.as<VT>(), L.reduce_sys(), g.col_sys(), sc);
    hipLaunchKernelGGL(k_pcg_p<VT>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s, g.d_cols.as<ColWork>(),
                       ws.zz.as<VT>(), ws.p.as<VT>(), ws.q.as<VT>(), g.col_sys(), sc);
    CAD_HIP(hipGetLastError());
}

template <class VT>
static void pcg_solve(cad_matrix* m, const double* b, double* x, const cad_solve_options& o, cad_solve_report& rep,
                      bool jacobi, SystemDetail& det)
{
    hipStream_t s = m->ctx->stream;
    Workspace& ws = workspace(m, 300 + int(sizeof(VT)) + (jacobi ? 0 : 50));
    const int64_t C = m->cols, R = m->rows;
    ensure(ws.P, sizeof(VT) * C);
    ensure(ws.p, sizeof(VT) * C);
    ensure(ws.q, sizeof(VT) * C);
    ensure(ws.zz, sizeof(VT) * C);
    ensure(ws.xr, sizeof(VT) * C);
    ensure(ws.rr, sizeof(VT) * C);
    ensure(ws.t, sizeof(VT) * R);
    ensure(ws.bvt, sizeof(VT) * R);
    ensure(ws.pscal, sizeof(PcgScalars) * m->nsys);
    ensure(ws.jac, sizeof(float) * C);
    const GridPlan& g = m->grid;
    if (jacobi) {
        matrix_jacobi(m, ws.jac.as<float>());
        hipLaunchKernelGGL(k_precond_from_jacobi<VT>, dim3(grid_for(C)), dim3(256), 0, s, ws.jac.as<float>(), C,
                           ws.P.as<VT>());
    } else {
        fill<VT>(ws.P.as<VT>(), C, VT(1), s);
    }
    {
        std::vector<PcgScalars> h(m->nsys);
        std::memset(h.data(), 0, sizeof(PcgScalars) * m->nsys);
        for (auto& e : h) e.active = 1;
        CAD_HIP(hipMemcpyAsync(ws.pscal.p, h.data(), sizeof(PcgScalars) * m->nsys, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL((k_convert<VT, double>), dim3(grid_for(R)), dim3(256), 0, s, ws.bvt.as<VT>(), b, R);
    // rhs = A^T b into rr (x0 = 0 so r = rhs)
    scatter<VT, VT, VT>(m, ws.bvt.as<VT>(), nullptr, ws.rr.as<VT>(), s);
    hipLaunchKernelGGL(k_pcg_init<VT>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s, g.d_cols.as<ColWork>(),
                       ws.P.as<VT>(), ws.rr.as<VT>(), ws.zz.as<VT>(), ws.p.as<VT>(), ws.xr.as<VT>(), ws.q.as<VT>(),
                       g.col_sys(), ws.pscal.as<PcgScalars>());
    CAD_HIP(hipGetLastError());
    const int batch = std::max(1, o.check_every);
    int it = 0;
    while (it < o.max_iterations) {
        const int n = std::min(batch, o.max_iterations - it);
        if (o.use_graph && batch > 1 && n == batch && s != nullptr) {
            // `batch` iterations captured once per workspace, replayed
            if (!ws.exec || ws.graph_iters != batch) {
                if (ws.exec) (void)hipGraphExecDestroy(ws.exec);
                if (ws.graph) (void)hipGraphDestroy(ws.graph);
                ws.exec = nullptr;
                ws.graph = nullptr;
                prepare_scatter<VT, VT>(m);
                prepare_gather<VT, EpiStore<VT>>(m);
                CAD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < batch; i++) pcg_iteration<VT>(m, ws, s);
                CAD_HIP(hipStreamEndCapture(s, &ws.graph));
                CAD_HIP(hipGraphInstantiate(&ws.exec, ws.graph, nullptr, nullptr, 0));
                ws.graph_iters = batch;
                ws.graph_prec = int(sizeof(VT));
            }
            CAD_HIP(hipGraphLaunch(ws.exec, s));
        } else {
            for (int i = 0; i < n; i++) pcg_iteration<VT>(m, ws, s);
        }
        it += n;
        const PcgScalars* h = read_scalars<PcgScalars>(m, ws, ws.pscal, s);
        for (int i = 0; i < m->nsys; i++)
            if (!std::isfinite(h[i].rz) || !std::isfinite(h[i].alpha))
                throw Error(CAD_ERR_BREAKDOWN, "PCG breakdown (non-finite recurrence)");
        if (o.verbose && m->nsys == 1) std::fprintf(stderr, "[pcg] it %d r.r %.6e\n", it, h[0].rr);
    }
    hipLaunchKernelGGL((k_convert<double, VT>), dim3(grid_for(C)), dim3(256), 0, s, x, ws.xr.as<VT>(), C);
    CAD_HIP(hipGetLastError());
    std::vector<double> atb = atb_norms(m, ws, b, s), nres, rn;
    explicit_residual(m, ws, x, b, nres, rn, s);
    double worst = 0, rsum = 0;
    det.rel.assign(m->nsys, 0.0);
    det.iters.assign(m->nsys, it);
    for (int i = 0; i < m->nsys; i++) {
        det.rel[i] = atb[i] > 0 ? nres[i] / atb[i] : nres[i];
        worst = std::max(worst, det.rel[i]);
        rsum += rn[i] * rn[i];
    }
    rep.iterations = it;
    rep.rel_normal_residual = worst;
    rep.residual_norm = std::sqrt(rsum);
    rep.converged = o.tol > 0 && worst <= o.tol;
}

// ------------------------------------------------------------ LSMR --------
// LSMR (Math/Solver.h:219-287, Fong & Saunders 2011, damp = 0) in FP64 on
// the device: the same bidiagonalisation launches as LSQR (gather with the
// u-update, window scatter, column reduction with the v-update), then the
// LSMR rotations (one block per system) and its vector updates; scalars stay
// on the device, check_every iterations are one captured graph, and each
// system stops on the device once |A^T r| (|zetabar|) reaches tol / 3 of
// |A^T b|.  Like the reference's LSMR it is not preconditioned.
static void lsmr_iteration(cad_matrix* m, Workspace& ws, hipStream_t s)
{
    const GridPlan& g = m->grid;
    LsqrScalars* sc = ws.scal.as<LsqrScalars>();
    gather<double>(m, ws.vd.as<double>(), EpiLsqrU<double>{ws.u.as<double>(), sc, 0.0}, s);
    ScatterLists& L = lists<double>(m->splan);
    with_view(m, [&](auto view) {
        launch_scatter_tiles<double, double>(m->splan, view, m->tune.srpt, m->d_sys_col.as<uint32_t>(),
                                             ws.u.as<double>(), sc, kScatterLsqr, s, true);
    });
    if (!L.reduce.empty())
        hipLaunchKernelGGL(k_lsqr_reduce_v<double>, dim3(unsigned(L.reduce.size())), dim3(kRowBlock), 0, s,
                           L.d_reduce.as<ReduceWork>(), (const double*)partials<double>(L), ws.D.as<double>(),
                           ws.v.as<double>(), L.reduce_sys(), sc);
    hipLaunchKernelGGL(k_lsmr_rot, dim3(unsigned(m->nsys)), dim3(kRowBlock), 0, s, L.reduce_sys(), sc,
                       ws.lsmr.as<LsmrState>());
    hipLaunchKernelGGL(k_lsmr_update<double>, dim3(unsigned(g.cols.size())), dim3(kRowBlock), 0, s,
                       g.d_cols.as<ColWork>(), ws.D.as<double>(), ws.v.as<double>(), ws.vd.as<double>(),
                       ws.w.as<double>(), ws.hb.as<double>(), ws.y.as<double>(), (const LsqrScalars*)sc,
                       (const LsmrState*)ws.lsmr.as<LsmrState>());
    CAD_HIP(hipGetLastError());
}

static void lsmr_solve(cad_matrix* m, const double* b, double* x, const cad_solve_options& o, cad_solve_report& rep,
                       SystemDetail& det)
{
    hipStream_t s = m->ctx->stream;
    Workspace& ws = workspace(m, 400);
    lsqr_alloc<double>(m, ws, s);
    ensure(ws.hb, sizeof(double) * m->cols);
    ensure(ws.lsmr, sizeof(LsmrState) * m->nsys);
    fill<double>(ws.D.as<double>(), m->cols, 1.0, s);
    const std::vector<double> atb = atb_norms(m, ws, b, s);
    std::vector<int> act(m->nsys, 1);
    set_active(m, ws, ws.scal, act, s);
    fill<double>(ws.x64.as<double>(), m->cols, 0.0, s);
    // x = 0: u = b / |b|, v = A^T u / |A^T u|, h = v (the LSQR start), hbar = 0
    lsqr_start_from_residual<double>(m, ws, b, s);
    fill<double>(ws.hb.as<double>(), m->cols, 0.0, s);
    hipLaunchKernelGGL(k_lsmr_init, dim3(unsigned((m->nsys + 255) / 256)), dim3(256), 0, s, m->nsys,
                       (const LsqrScalars*)ws.scal.as<LsqrScalars>(), ws.lsmr.as<LsmrState>());
    CAD_HIP(hipGetLastError());
    {
        // the device stop: |A^T r| at tol / 3 of |A^T b| (none when tol = 0)
        const LsqrScalars* h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
        std::vector<LsqrScalars> hv(h, h + m->nsys);
        for (auto& e : hv) {
            e.target = o.tol > 0 ? o.tol / 3.0 : 0.0;
            e.stop = 0;
        }
        CAD_HIP(hipMemcpyAsync(ws.scal.p, hv.data(), sizeof(LsqrScalars) * m->nsys, hipMemcpyHostToDevice, s));
    }
    const int batch = std::max(1, o.check_every);
    int it = 0;
    while (it < o.max_iterations) {
        const int n = std::min(batch, o.max_iterations - it);
        if (o.use_graph && batch > 1 && n == batch && s != nullptr) {
            if (!ws.exec || ws.graph_iters != batch) {
                if (ws.exec) (void)hipGraphExecDestroy(ws.exec);
                if (ws.graph) (void)hipGraphDestroy(ws.graph);
                ws.exec = nullptr;
                ws.graph = nullptr;
                prepare_scatter<double, double>(m);
                prepare_gather<double, EpiLsqrU<double>>(m);
                CAD_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int i = 0; i < batch; i++) lsmr_iteration(m, ws, s);
                CAD_HIP(hipStreamEndCapture(s, &ws.graph));
                CAD_HIP(hipGraphInstantiate(&ws.exec, ws.graph, nullptr, nullptr, 0));
                ws.graph_iters = batch;
                ws.graph_prec = 8;
            }
            CAD_HIP(hipGraphLaunch(ws.exec, s));
        } else {
            for (int i = 0; i < n; i++) lsmr_iteration(m, ws, s);
        }
        it += n;
        const LsqrScalars* h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
        bool running = false;
        for (int i = 0; i < m->nsys; i++) {
            const double est = h[i].nres0 > 0 ? h[i].nres / h[i].nres0 : 0.0;
            if (!std::isfinite(est)) throw Error(CAD_ERR_BREAKDOWN, "LSMR breakdown (non-finite recurrence)");
            running |= h[i].active && h[i].stop == 0;
        }
        if (o.verbose && m->nsys == 1)
            std::fprintf(stderr, "[lsmr] it %d est %.3e\n", it, h[0].nres0 > 0 ? h[0].nres / h[0].nres0 : 0.0);
        if (!running) break;
    }
    CAD_HIP(hipMemcpyAsync(x, ws.y.p, sizeof(double) * m->cols, hipMemcpyDeviceToDevice, s));
    std::vector<double> nres, rn;
    explicit_residual(m, ws, x, b, nres, rn, s);
    const LsqrScalars* h = read_scalars<LsqrScalars>(m, ws, ws.scal, s);
    double worst = 0, rsum = 0;
    det.rel.assign(m->nsys, 0.0);
    det.iters.assign(m->nsys, it);
    for (int i = 0; i < m->nsys; i++) {
        det.rel[i] = atb[i] > 0 ? nres[i] / atb[i] : nres[i];
        det.iters[i] = h[i].iters;
        worst = std::max(worst, det.rel[i]);
        rsum += rn[i] * rn[i];
    }
    rep.iterations = it;
    rep.rel_normal_residual = worst;
    rep.residual_norm = std::sqrt(rsum);
    rep.converged = o.tol > 0 && worst <= o.tol;
}

// ------------------------------------------------------------ entry -------
void solve(cad_matrix* m, const double* b, double* x, const cad_solve_options& o, cad_solve_report& rep,
           SystemDetail* detail)
{
    std::memset(&rep, 0, sizeof(rep));
    SystemDetail det;
    auto t0 = std::chrono::steady_clock::now();
    if (o.solver == CAD_SOLVER_LSQR) {
        if (o.precision == CAD_PREC_FP32) lsqr_plain<float>(m, b, x, o, rep, det);
        else lsqr_refined(m, b, x, o, rep, det, o.precision == CAD_PREC_MIXED);
    } else if (o.solver == CAD_SOLVER_PCG || o.solver == CAD_SOLVER_CG) {
        const bool jac = o.solver == CAD_SOLVER_PCG;
        if (o.precision == CAD_PREC_FP32) pcg_solve<float>(m, b, x, o, rep, jac, det);
        else pcg_solve<double>(m, b, x, o, rep, jac, det);
    } else if (o.solver == CAD_SOLVER_LSMR) {
        lsmr_solve(m, b, x, o, rep, det);
    } else {
        throw Error(CAD_ERR_ARGUMENT, "unknown solver");
    }
    CAD_HIP(hipStreamSynchronize(m->ctx->stream));
    rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (detail) *detail = std::move(det);
}

double rel_normal_residual(cad_matrix* m, const double* b, const double* x, double* rnorm)
{
    hipStream_t s = m->ctx->stream;
    Workspace& ws = workspace(m, 1);
    std::vector<double> atb = atb_norms(m, ws, b, s), nres, rn;
    explicit_residual(m, ws, x, b, nres, rn, s);
    double worst = 0, rsum = 0;
    for (int i = 0; i < m->nsys; i++) {
        worst = std::max(worst, atb[i] > 0 ? nres[i] / atb[i] : nres[i]);
        rsum += rn[i] * rn[i];
    }
    if (rnorm) *rnorm = std::sqrt(rsum);
    return worst;
}

// ------------------------------------------------------------ autotune --
namespace {

struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    EventTimer()
    {
        CAD_HIP(hipEventCreate(&a));
        CAD_HIP(hipEventCreate(&b));
    }
    ~EventTimer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    // average of reps launches after one warm launch, in microseconds
    template <class F>
    double us(F&& f, hipStream_t s, int reps = 3)
    {
        f();
        CAD_HIP(hipEventRecord(a, s));
        for (int i = 0; i < reps; i++) f();
        CAD_HIP(hipEventRecord(b, s));
        CAD_HIP(hipEventSynchronize(b));
        float ms = 0;
        CAD_HIP(hipEventElapsedTime(&ms, a, b));
        return 1e3 * double(ms) / reps;
    }
};

void drop_graphs(cad_matrix* m)
{
    for (auto& kv : m->ws) {
        Workspace& w = *kv.second;
        if (w.exec) (void)hipGraphExecDestroy(w.exec);
        if (w.graph) (void)hipGraphDestroy(w.graph);
        w.exec = nullptr;
        w.graph = nullptr;
        w.graph_iters = 0;
    }
}

}  // namespace

void set_schedule(cad_matrix* m, const TuneParams& tp)
{
    const bool u16 = !m->mf && m->ibytes == 2;
    auto one_of = [](int v, std::initializer_list<int> ok) {
        for (int o : ok)
            if (v == o) return true;
        return false;
    };
    if (!(u16 ? one_of(tp.rpt4, {0, 2, 4, 8}) : one_of(tp.rpt4, {0, 2, 4})) || !one_of(tp.rpt8, {0, 1, 2}))
        throw Error(CAD_ERR_ARGUMENT, "unsupported rows per thread of the tile gather");
    for (int b : {tp.bs4, tp.bs8})
        if (!(b == kTileBlock || (b == kTileBlock / 2 && !m->mf)))
            throw Error(CAD_ERR_ARGUMENT, "unsupported block size of the tile gather");
    // sub-tiles: 1, or 2..3 for the full-block stored-index variants that take
    // them (rpt * value bytes == 8, rpt * index bytes <= 4)
    auto sub_ok = [&](int bs, int sub, int rpt, int vbytes) {
        return sub == 1 || (sub >= 2 && sub <= 3 && bs == kTileBlock && !m->mf && rpt * vbytes == 8 &&
                            rpt * m->ibytes <= 4);
    };
    if (!sub_ok(tp.bs4, tp.sub4, tp.rpt4, 4) || !sub_ok(tp.bs8, tp.sub8, tp.rpt8, 8))
        throw Error(CAD_ERR_ARGUMENT, "unsupported sub-tiles of the tile gather");
    if (!(u16 ? one_of(tp.srpt, {4, 8}) : tp.srpt == 4) || !one_of(tp.scatter_order, {0, 1}) ||
        !(tp.blocks_per_cu > 0 && tp.blocks_per_cu <= 64) || !(tp.partial_cap > 0 && tp.partial_cap <= 1) ||
        !(tp.scatter_balance >= 0 && tp.scatter_balance <= 1))
        throw Error(CAD_ERR_ARGUMENT, "unsupported scatter schedule");
    CAD_HIP(hipStreamSynchronize(m->ctx->stream));
    drop_graphs(m);
    const TuneParams keep = m->tune;
    m->tune = tp;
    build_gather_plan(m, m->gplan4, 4, tp.rpt4, tp.bs4, tp.sub4);
    build_gather_plan(m, m->gplan8, 8, tp.rpt8, tp.bs8, tp.sub8);
    if ((tp.rpt4 > 0 && !m->gplan4.ok) || (tp.rpt8 > 0 && !m->gplan8.ok)) {
        m->tune = keep;
        build_gather_plan(m, m->gplan4, 4, keep.rpt4, keep.bs4, keep.sub4);
        build_gather_plan(m, m->gplan8, 8, keep.rpt8, keep.bs8, keep.sub8);
        throw Error(CAD_ERR_ARGUMENT, "this matrix cannot use the tile gather");
    }
    build_scatter_plan(m->splan, m->slot_lo, m->slot_hi, m->sys_row_begin, m->sys_col, m->cols_per_sys,
                       m->mf ? 2 : m->ibytes, *m->ctx, tp);
    CAD_HIP(hipStreamSynchronize(m->ctx->stream));
}

void kernel_times(cad_matrix* m, int reps, double* us, int n)
{
    hipStream_t s = m->ctx->stream;
    CAD_HIP(hipStreamSynchronize(s));
    const int64_t R = std::max<int64_t>(1, m->rows), Cn = std::max<int64_t>(1, m->cols);
    DBuf x4(4 * Cn), y4(4 * R), z4(4 * Cn);
    fill<float>(x4.as<float>(), Cn, 1.0f, s);
    fill<float>(y4.as<float>(), R, 1.0f, s);
    EventTimer t;
    double v[5] = {0, 0, 0, 0, 0};
    v[0] = t.us([&] { spmv<float>(m, x4.as<float>(), y4.as<float>(), s); }, s, reps);
    with_view(m, [&](auto view) {
        using View = decltype(view);
        ScatterPlan& sp = m->splan;
        partials<float>(lists<float>(sp));
        const uint32_t* sc = m->d_sys_col.as<uint32_t>();
        const float* y = y4.as<float>();
        if (fixed_point_scatter<float, View>())
            v[3] = t.us([&] { launch_absmax(sp, y, m->grid, m->nsys, s); }, s, reps);
        v[1] = t.us([&] { launch_scatter_tiles<float, float>(sp, view, m->tune.srpt, sc, y, nullptr, kScatterAll, s); },
                    s, reps);
        v[2] = t.us([&] { launch_scatter_reduce<float>(sp, nullptr, kScatterAll, z4.as<float>(), s); }, s, reps);
    });
    v[4] = t.us([&] { launch_jacobi_part(m, s); }, s, std::max(1, reps / 10));
    CAD_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < std::min(n, 5); i++) us[i] = v[i];
}

std::string autotune(cad_matrix* m)
{
    hipStream_t s = m->ctx->stream;
    CAD_HIP(hipStreamSynchronize(s));
    drop_graphs(m);
    const int64_t R = std::max<int64_t>(1, m->rows), Cn = std::max<int64_t>(1, m->cols);
    DBuf x4(4 * Cn), y4(4 * R), z4(4 * Cn), x8(8 * Cn), y8(8 * R);
    fill<float>(x4.as<float>(), Cn, 1.0f, s);
    fill<float>(y4.as<float>(), R, 1.0f, s);
    fill<double>(x8.as<double>(), Cn, 1.0, s);
    EventTimer t;
    std::string log;
    char line[160];
    TuneParams& tp = m->tune;
    // gathers: tile rows per thread (0: the row kernel) x threads per block
    // (1024, or 512 with two blocks per CU) x sub-tiles per staging (stored
    // indices only)
    struct GShape { int rpt, bs, sub; };
    auto tune_gather = [&](int vbytes, const std::vector<int>& rpts, int& rpt, int& bs, int& sub, auto&& run) {
        GatherPlan& gp = vbytes == 4 ? m->gplan4 : m->gplan8;
        std::vector<GShape> cands;
        for (int r : rpts) {
            if (r == 0) {
                cands.push_back(GShape{0, kTileBlock, 1});
                continue;
            }
            for (int sb : {1, 2, 3})
                if (sb == 1 || (!m->mf && r * vbytes == 8 && r * m->ibytes <= 4))
                    cands.push_back(GShape{r, kTileBlock, sb});
            if (!m->mf) cands.push_back(GShape{r, kTileBlock / 2, 1});
        }
        GShape best{rpt, bs, sub};
        double best_us = 1e300;
        for (const GShape& c : cands) {
            build_gather_plan(m, gp, vbytes, c.rpt, c.bs, c.sub);
            if (c.rpt > 0 && !gp.ok) continue;
            const double us = t.us(run, s);
            std::snprintf(line, sizeof line, "gather f%d rpt %d block %d sub %d: %.1f us\n", 8 * vbytes, c.rpt, c.bs,
                          c.sub, us);
            log += line;
            if (us < best_us) best_us = us, best = c;
        }
        rpt = best.rpt;
        bs = best.bs;
        sub = best.sub;
        build_gather_plan(m, gp, vbytes, rpt, bs, sub);
    };
    tune_gather(4, (!m->mf && m->ibytes == 2) ? std::vector<int>{4, 8, 2, 0} : std::vector<int>{4, 2, 0}, tp.rpt4,
                tp.bs4, tp.sub4, [&] { spmv<float>(m, x4.as<float>(), y4.as<float>(), s); });
    tune_gather(8, std::vector<int>{2, 1, 0}, tp.rpt8, tp.bs8, tp.sub8,
                [&] { spmv<double>(m, x8.as<double>(), y8.as<double>(), s); });
    // scatter: work order, rows per step, tiles per CU and the partial-traffic
    // cap (FP32 timing)
    {
        struct Cand { int order, srpt; double bpc, cap, bal; };
        std::vector<Cand> cands;
        const std::vector<int> srpts = (!m->mf && m->ibytes == 2) ? std::vector<int>{4, 8} : std::vector<int>{4};
        for (int order : {1, 0})
            for (int sr : srpts)
                for (double bpc : {1.0, 2.0, 4.0})
                    for (double cap : {0.1, 0.25, 1.0})
                        for (double bal : {1.0, 0.5, 0.25})
                            if (order == 0 || bal == 1.0) cands.push_back(Cand{order, sr, bpc, cap, bal});
        const TuneParams keep = tp;
        Cand best{tp.scatter_order, tp.srpt, tp.blocks_per_cu, tp.partial_cap, tp.scatter_balance};
        double best_us = 1e300;
        for (const Cand& c : cands) {
            tp.scatter_order = c.order;
            tp.srpt = c.srpt;
            tp.blocks_per_cu = c.bpc;
            tp.partial_cap = c.cap;
            tp.scatter_balance = c.bal;
            build_scatter_plan(m->splan, m->slot_lo, m->slot_hi, m->sys_row_begin, m->sys_col, m->cols_per_sys,
                               m->mf ? 2 : m->ibytes, *m->ctx, tp);
            const double us =
                t.us([&] { scatter<float, float, float>(m, y4.as<float>(), nullptr, z4.as<float>(), s); }, s);
            std::snprintf(line, sizeof line,
                          "scatter f32 order %d rpt %d blocks/CU %.0f cap %.2f balance %.2f (%zu tiles): %.1f us\n",
                          c.order, c.srpt, c.bpc, c.cap, c.bal, m->splan.f8.work.size(), us);
            log += line;
            // a later candidate must win by 2 %, and a per-window (order 0)
            // one over the chunk-shared XCD-major order, which comes first, by
            // 5 %: that order moves half the HBM bytes (its windows' re-reads
            // hit L2), and a few % of isolated time are not worth twice the
            // traffic (the two also trade places from run to run on the 5M-row
            // matrix, where they are within 3 % of each other)
            const double margin = (c.order == 0 && best.order == 1) ? 0.95 : 0.98;
            if (us < margin * best_us) best_us = us, best = c;
        }
        tp = keep;
        tp.scatter_order = best.order;
        tp.srpt = best.srpt;
        tp.blocks_per_cu = best.bpc;
        tp.partial_cap = best.cap;
        tp.scatter_balance = best.bal;
        build_scatter_plan(m->splan, m->slot_lo, m->slot_hi, m->sys_row_begin, m->sys_col, m->cols_per_sys,
                           m->mf ? 2 : m->ibytes, *m->ctx, tp);
    }
    CAD_HIP(hipStreamSynchronize(s));
    m->tuned = true;
    std::snprintf(line, sizeof line,
                  "chosen: gather f32 rpt %d block %d sub %d, f64 rpt %d block %d sub %d, scatter order %d rpt %d "
                  "blocks/CU %.0f cap %.2f balance %.2f\n",
                  tp.rpt4, tp.bs4, tp.sub4, tp.rpt8, tp.bs8, tp.sub8, tp.scatter_order, tp.srpt, tp.blocks_per_cu,
                  tp.partial_cap, tp.scatter_balance);
    log += line;
    return log;
}

}  // namespace cad

cad_matrix::~cad_matrix() = default;
