// extern "C" boundary of libcassandra_amd.so (declared in include/cassandra_amd.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "evaluator.hpp"
#include "kernels.hpp"
#include "solver.hpp"

// Scratch of cad_eval_backward_f32_noalloc: the FP64 partial column sums of
// its row chunks, chunk after chunk (cols doubles each).  Every device the
// library runs on holds its own copy until the process ends.  In host code the
// name is only a stand-in: the device address comes from hipGetSymbolAddress
// at context set-up (noalloc_scratch_init) and only that address is used.
constexpr int64_t kNoAllocScratchDoubles = int64_t(4) << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace cad {

thread_local std::string g_last_error;

void set_error(const std::string& s) { g_last_error = s; }

// ---- evaluator backward without allocation --------------------------------
// grad_w = A^T g for row-major int32 indices in two launches that need no
// plan: a block per (column window, row chunk) adds g[r] into an FP64 LDS
// window for the slots of the window's pattern block and writes the window to
// the chunk's partial slice of the scratch; then every column sums its chunk
// slices in chunk order (FP64) and rounds once to float.
constexpr int kNoAllocWindow = 16384;     // columns per LDS window (128 KB of doubles)
constexpr int kNoAllocMaxWindows = 128;   // window table passed by value
constexpr int kNoAllocBlock = 1024;
constexpr int kMaxDevices = 64;

struct NoAllocWindow {
    int32_t c0, c1;  // global columns [c0, c1) of one pattern block
    int32_t s0, s1;  // the slots whose entries lie in that block
};
struct NoAllocWindows {
    NoAllocWindow w[kNoAllocMaxWindows];
};

__global__ __launch_bounds__(kNoAllocBlock) void k_noalloc_grad_window(const int32_t* __restrict__ idx, int64_t B,
                                                                       int k, const float* __restrict__ g,
                                                                       NoAllocWindows win, int64_t rows_per_chunk,
                                                                       int64_t cols, double* __restrict__ part)
{
    extern __shared__ double lds[];
    const NoAllocWindow w = win.w[blockIdx.x];
    const int n = w.c1 - w.c0;
    for (int i = threadIdx.x; i < n; i += kNoAllocBlock) lds[i] = 0.0;
    __syncthreads();
    const int64_t r0 = int64_t(blockIdx.y) * rows_per_chunk;
    const int64_t r1 = r0 + rows_per_chunk < B ? r0 + rows_per_chunk : B;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kNoAllocBlock) {
        const double gv = double(g[r]);
        const int32_t* row = idx + r * k;
        for (int s = w.s0; s < w.s1; s++) {
            // entries outside the window (or outside [0, cols)) are not this block's
            const uint32_t c = uint32_t(row[s]) - uint32_t(w.c0);
            if (c < uint32_t(n)) atomicAdd(&lds[c], gv);
        }
    }
    __syncthreads();
    double* out = part + int64_t(blockIdx.y) * cols + w.c0;
    for (int i = threadIdx.x; i < n; i += kNoAllocBlock) out[i] = lds[i];
}

__global__ __launch_bounds__(256) void k_noalloc_grad_reduce(const double* __restrict__ part, int nchunk,
                                                             int64_t cols, float* __restrict__ grad)
{
    for (int64_t c = blockIdx.x * int64_t(256) + threadIdx.x; c < cols; c += int64_t(gridDim.x) * 256) {
        double s = 0.0;
        for (int ch = 0; ch < nchunk; ch++) s += part[int64_t(ch) * cols + c];
        grad[c] = float(s);
    }
}

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current: the device address of the
// scratch array and the LDS limit of the window kernel.
void noalloc_scratch_init(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) throw Error(CAD_ERR_DEVICE, "device ordinal beyond 64");
    if (g_scratch_ptr[device]) return;
    void* ptr = nullptr;
    CAD_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw Error(CAD_ERR_DEVICE, "no device address for the scratch array");
    CAD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_noalloc_grad_window),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(kNoAllocWindow * sizeof(double))));
    g_scratch_ptr[device] = static_cast<double*>(ptr);
}

double* noalloc_scratch(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || !g_scratch_ptr[device])
        throw Error(CAD_ERR_DEVICE, "scratch of the no-allocation entry point not set up on this device");
    return g_scratch_ptr[device];
}


namespace {

template <class F>
int guarded(F&& f)
{
    try {
        f();
        return CAD_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return CAD_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return CAD_ERR_ARGUMENT;
    } catch (...) {
        g_last_error = "unknown error";
        return CAD_ERR_ARGUMENT;
    }
}

void need(bool c, const char* msg, int code = CAD_ERR_ARGUMENT)
{
    if (!c) throw Error(code, msg);
}

int grid1(int64_t n, int bs) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

// Upload the index programs of a pattern group.
void upload_programs(cad_patterns* p)
{
    const PatternGroup& g = *p->group;
    std::vector<DevSlot> slots;
    std::vector<uint8_t> sq;
    for (const SlotProgram& s : g.slots) {
        DevSlot d;
        std::memset(&d, 0, sizeof(d));
        d.kind = uint8_t(s.kind);
        for (int l = 0; l < 3; l++) {
            d.off[l] = uint16_t(sq.size());
            d.len[l] = uint8_t(s.list[l].src.size());
            sq.insert(sq.end(), s.list[l].src.begin(), s.list[l].src.end());
        }
        d.hs = s.hs;
        d.ds = s.ds;
        d.col_offset = s.col_offset;
        d.dense_offset = s.dense_offset;
        slots.push_back(d);
    }
    // the ternary-word programs: per list, its entries {2 (q mod 16), 3^j}
    // grouped by the word q / 16; three zero entries end the array (the
    // kernels load four entries at a time, also at a list's last entry)
    std::vector<DevTern> tslots;
    std::vector<uint32_t> tern;
    for (const SlotProgram& s : g.slots) {
        DevTern t;
        std::memset(&t, 0, sizeof(t));
        for (int l = 0; l < 3; l++)
            append_ternary_list(s.list[l].src.data(), int(s.list[l].src.size()), tern, t, l);
        tslots.push_back(t);
    }
    tern.insert(tern.end(), kTernPadding, 0u);
    hipStream_t st = p->ctx->stream;
    p->d_tslots = to_device(tslots, st);
    p->d_tern = to_device(tern, st);
    p->d_slots = to_device(slots, st);
    sq.resize((sq.size() + 7) / 4 * 4, 0);  // the kernels read the squares as whole 32-bit words
    p->d_sq = to_device(sq, st);
    p->d_dense = to_device(g.dense, st);
    p->slot_lo.clear();
    p->slot_hi.clear();
    for (const SlotProgram& s : g.slots) {
        p->slot_lo.push_back(s.col_offset);
        p->slot_hi.push_back(s.col_offset + s.block_size);
    }
    std::vector<uint32_t> zeros(std::max<size_t>(1, g.slots.size()), 0u);
    p->d_base0 = to_device(zeros, st);
    p->d_syscol0 = to_device(std::vector<uint32_t>{0u}, st);
}

void extract(cad_patterns* p, const uint64_t* pos, int64_t n, void* out, int bytes, int64_t slot_stride,
             int64_t row_stride, int local)
{
    if (n <= 0) return;
    const int k = p->group->k();
    hipStream_t s = p->ctx->stream;
    const dim3 grid(grid1(n, kRowBlock * kExtractRows));
    if (bytes == 2)
        hipLaunchKernelGGL((k_index_extract<uint16_t, kExtractRows>), grid, dim3(kRowBlock), 0, s, pos, n,
                           p->d_slots.as<DevSlot>(), k, p->d_tslots.as<DevTern>(), p->d_tern.as<uint32_t>(),
                           p->d_dense.as<uint32_t>(), static_cast<uint16_t*>(out), slot_stride, row_stride, local);
    else
        hipLaunchKernelGGL((k_index_extract<uint32_t, kExtractRows>), grid, dim3(kRowBlock), 0, s, pos, n,
                           p->d_slots.as<DevSlot>(), k, p->d_tslots.as<DevTern>(), p->d_tern.as<uint32_t>(),
                           p->d_dense.as<uint32_t>(), static_cast<uint32_t*>(out), slot_stride, row_stride, local);
    CAD_HIP(hipGetLastError());
}

cad_matrix* new_matrix(cad_context* ctx, int64_t rows, int k, int64_t cols)
{
    need(rows >= 0 && k > 0 && k <= 1024 && cols > 0, "invalid matrix shape");
    need(cols <= int64_t(0xFFFFFFFFu), "more than 2^32 columns");
    auto* m = new cad_matrix;
    m->ctx = ctx;
    m->rows = rows;
    m->k = k;
    m->cols = cols;
    return m;
}

// The scatter plan of the evaluator backward for B rows of a pattern group
// (cached per B; a training loop sees one or two minibatch sizes).
const ScatterPlan& eval_plan(cad_patterns* p, int64_t B)
{
    auto it = p->eval_plans.find(B);
    if (it != p->eval_plans.end()) return *it->second;
    if (p->eval_plans.size() >= 16) p->eval_plans.clear();
    std::unique_ptr<ScatterPlan> sp(new ScatterPlan);
    // A minibatch has few rows per window: the row chunks are sized to give
    // every CU a tile (partial slices up to 8x the index bytes; the default
    // cap of 1x left a 65,536-row minibatch of 4 patterns 28 tiles), with the
    // 2 MB floor of a large system (the 8 MB one of a small system gave the
    // training step more slices to sum than tiles to spread: 3.7 -> 4.0 ms an
    // epoch).
    TuneParams tp;
    tp.partial_cap = 8.0;
    tp.small_floor_bytes = 2.0 * 1024 * 1024;
    build_scatter_plan(*sp, p->slot_lo, p->slot_hi, std::vector<int64_t>{0, B}, std::vector<uint32_t>{0u},
                       int64_t(p->group->cols()), 4, *p->ctx, tp);
    const ScatterPlan& ref = *sp;
    p->eval_plans[B] = std::move(sp);
    return ref;
}

// The backward's first stage (B > 0): the window tiles of the evaluator's
// scatter plan into its partial slices; returns the plan's lists and slices.
template <class VT, class View>
ScatterLists& eval_backward_windows(cad_patterns* p, const View& view, int64_t B, const VT* g, VT*& part)
{
    hipStream_t s = p->ctx->stream;
    ScatterPlan& sp = const_cast<ScatterPlan&>(eval_plan(p, B));
    ScatterLists& L = sp.f8;  // 8-byte LDS windows (k_scatter_part accumulates in FP64)
    const size_t need = sizeof(double) * size_t(std::max<int64_t>(1, L.part_elems));  // either precision
    if (L.part.bytes < need) L.part = DBuf(need);
    part = L.part.as<VT>();
    if (!L.work.empty()) {
        const size_t lds = size_t(L.widest) * sizeof(double);
        auto kern = k_scatter_part<VT, VT, View>;
        if (lds > 64 * 1024)
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
        hipLaunchKernelGGL(kern, dim3(unsigned(L.work.size())), dim3(kScatterBlock), lds, s,
                           view, L.d_work.as<ScatterWork>(), L.d_groups.as<SlotGroup>(),
                           L.d_slot_list.as<uint16_t>(), p->d_syscol0.as<uint32_t>(), g,
                           (LsqrScalars*)nullptr, kScatterAll, FixedScale{nullptr, 0}, 0, part, GridSys{});
        CAD_HIP(hipGetLastError());
    }
    return L;
}

template <class VT, class View>
void eval_backward_view(cad_patterns* p, const View& view, int64_t B, const VT* g, VT* grad)
{
    hipStream_t s = p->ctx->stream;
    const int64_t C = int64_t(p->group->cols());
    if (B <= 0) {
        hipLaunchKernelGGL(k_fill<VT>, dim3(grid1(C, 256)), dim3(256), 0, s, grad, C, VT(0));
        CAD_HIP(hipGetLastError());
        return;
    }
    VT* part = nullptr;
    ScatterLists& L = eval_backward_windows<VT>(p, view, B, g, part);
    hipLaunchKernelGGL((k_scatter_reduce<VT, VT>), dim3(unsigned(L.reduce.size())), dim3(kRowBlock), 0, s,
                       L.d_reduce.as<ReduceWork>(), (const VT*)part, grad, (const LsqrScalars*)nullptr, kScatterAll);
    CAD_HIP(hipGetLastError());
}

template <class VT>
void eval_backward(cad_patterns* p, const int32_t* idx, int64_t B, const VT* g, VT* grad)
{
    eval_backward_view<VT>(p, RowMajor32{idx, p->group->k()}, B, g, grad);
}

// One Adam step of the Huber-loss fit on the minibatch rows sel[0..B): the
// forward with the loss and its gradient (which also copies the minibatch's
// rows contiguously), the backward over that copy (the evaluator's scatter
// plan) whose column reduction applies the Adam update.
void train_step(cad_patterns* p, const int32_t* idx, const float* target, const int64_t* sel, int64_t B,
                float delta, float lr, float b1, float b2, float eps, int64_t step, float* w, float* m, float* v,
                double* loss)
{
    hipStream_t s = p->ctx->stream;
    const int64_t C = int64_t(p->group->cols());
    const int k = p->group->k();
    const int nblk = grid1(B, kRowBlock);
    auto grow = [](DBuf& b, size_t bytes) {
        if (b.bytes < bytes) b = DBuf(bytes);
    };
    grow(p->d_train_g, sizeof(float) * size_t(std::max<int64_t>(1, B)));
    grow(p->d_train_grad, sizeof(float) * size_t(C));
    grow(p->d_train_part, sizeof(double) * size_t(nblk));
    grow(p->d_train_rows, sizeof(int32_t) * size_t(std::max<int64_t>(1, B)) * size_t(k));
    float* g = p->d_train_g.as<float>();
    float* grad = p->d_train_grad.as<float>();
    double* part = p->d_train_part.as<double>();
    int32_t* rows = p->d_train_rows.as<int32_t>();
    const AdamStep a{lr, b1, b2, eps, 1.0f - std::pow(b1, float(step)), std::sqrt(1.0f - std::pow(b2, float(step)))};
    if (B > 0) {
        hipLaunchKernelGGL(k_train_forward, dim3(nblk), dim3(kRowBlock), 0, s, idx, sel, B, k, w, uint32_t(C), target,
                           delta, g, part, rows);
        // the backward's window tiles over the contiguous copy of the
        // minibatch's rows, then their column sums with the Adam update
        float* gpart = nullptr;
        ScatterLists& L = eval_backward_windows<float>(p, RowMajor32{rows, k}, B, g, gpart);
        hipLaunchKernelGGL(k_adam_reduce, dim3(unsigned(L.reduce.size())), dim3(kRowBlock), 0, s,
                           L.d_reduce.as<ReduceWork>(), (const float*)gpart, w, m, v, a, (const double*)part, nblk, B,
                           loss);
    } else {
        // no rows: a zero gradient (Adam still moves the weights by its moments)
        eval_backward_view<float>(p, RowMajor32{rows, k}, B, g, grad);
        hipLaunchKernelGGL(k_adam, dim3(grid1(C, kRowBlock)), dim3(kRowBlock), 0, s, w, (const float*)grad, m, v, C, a,
                           (const double*)part, 0, int64_t(1), loss);
    }
    CAD_HIP(hipGetLastError());
}

template <class VT>
void eval_forward(cad_patterns* p, const int32_t* idx, int64_t B, const VT* w, VT* out)
{
    if (B <= 0) return;
    hipLaunchKernelGGL(k_eval_forward<VT>, dim3(grid1(B, kRowBlock)), dim3(kRowBlock), 0, p->ctx->stream, idx, B,
                       p->group->k(), w, uint32_t(p->group->cols()), out);
    CAD_HIP(hipGetLastError());
}

}  // namespace
}  // namespace cad

using namespace cad;

extern "C" {

const char* cad_last_error(void) { return g_last_error.c_str(); }
const char* cad_version(void) { return "cassandra_amd 0.1 (gfx950)"; }

int cad_context_create(int device, cad_context** out)
{
    return guarded([&] {
        need(out != nullptr, "null output");
        int n = 0;
        CAD_HIP(hipGetDeviceCount(&n));
        need(device >= 0 && device < n, "no such device", CAD_ERR_DEVICE);
        CAD_HIP(hipSetDevice(device));
        noalloc_scratch_init(device);
        auto* c = new cad_context;
        c->device = device;
        hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            CAD_HIP(e);
        }
        c->stream = c->own;
        hipDeviceProp_t prop;
        CAD_HIP(hipGetDeviceProperties(&prop, device));
        c->cus = prop.multiProcessorCount;
        c->max_lds = prop.sharedMemPerBlock;
        *out = c;
    });
}

int cad_context_destroy(cad_context* ctx)
{
    return guarded([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->stream);
        delete ctx;
    });
}

int cad_context_set_stream(cad_context* ctx, void* stream)
{
    return guarded([&] {
        need(ctx != nullptr, "null context");
        ctx->stream = static_cast<hipStream_t>(stream);
    });
}

int cad_context_reset_stream(cad_context* ctx)
{
    return guarded([&] {
        need(ctx != nullptr, "null context");
        ctx->stream = ctx->own;
    });
}

int cad_context_get_stream(cad_context* ctx, void** out)
{
    return guarded([&] {
        need(ctx && out, "null argument");
        *out = static_cast<void*>(ctx->stream);
    });
}

int cad_context_synchronize(cad_context* ctx)
{
    return guarded([&] { CAD_HIP(hipStreamSynchronize(ctx->stream)); });
}

int cad_malloc(cad_context* ctx, size_t bytes, void** out)
{
    return guarded([&] {
        need(ctx && out, "null argument");
        CAD_HIP(hipSetDevice(ctx->device));
        *out = nullptr;
        if (bytes) CAD_HIP(hipMalloc(out, bytes));
    });
}

int cad_free(cad_context* ctx, void* p)
{
    return guarded([&] {
        (void)ctx;
        if (p) CAD_HIP(hipFree(p));
    });
}

int cad_memcpy_h2d(cad_context* ctx, void* d, const void* h, size_t bytes)
{
    return guarded([&] {
        if (!bytes) return;
        CAD_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
        CAD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cad_memcpy_d2h(cad_context* ctx, void* h, const void* d, size_t bytes)
{
    return guarded([&] {
        if (!bytes) return;
        CAD_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
        CAD_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cad_memcpy_d2d(cad_context* ctx, void* d, const void* s, size_t bytes)
{
    return guarded([&] {
        if (!bytes) return;
        CAD_HIP(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    });
}

// ------------------------------------------------------------ patterns ----
int cad_patterns_create(cad_context* ctx, const uint64_t* masks, int n, cad_patterns** out)
{
    return guarded([&] {
        need(out && (masks || n == 0) && n >= 0, "invalid argument");
        std::vector<uint64_t> v(masks, masks + n);
        std::unique_ptr<PatternGroup> g;
        try {
            g.reset(new PatternGroup(v));
        } catch (const std::runtime_error& e) {
            throw Error(CAD_ERR_PATTERN, e.what());
        }
        need(g->k() > 0, "empty pattern group", CAD_ERR_PATTERN);
        auto* p = new cad_patterns;
        p->ctx = ctx;
        p->group = std::move(g);
        try {
            if (ctx) upload_programs(p);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

int cad_patterns_destroy(cad_patterns* p)
{
    return guarded([&] { delete p; });
}

int cad_patterns_info(const cad_patterns* p, int* k, int64_t* cols, int* npat)
{
    return guarded([&] {
        need(p != nullptr, "null patterns");
        if (k) *k = p->group->k();
        if (cols) *cols = int64_t(p->group->cols());
        if (npat) *npat = int(p->group->patterns.size());
    });
}

int cad_patterns_pattern_info(const cad_patterns* p, int i, int* cls, uint64_t* mask, int64_t* size, int* order,
                              int64_t* offset)
{
    return guarded([&] {
        need(p && i >= 0 && i < int(p->group->patterns.size()), "pattern index out of range");
        const PatternInfo& pi = p->group->patterns[i];
        if (cls) *cls = pi.cls;
        if (mask) *mask = pi.mask;
        if (size) *size = pi.size;
        if (order) *order = pi.order;
        if (offset) *offset = pi.offset;
    });
}

int cad_patterns_variations(const cad_patterns* p, uint64_t* out)
{
    return guarded([&] {
        need(p && out, "null argument");
        int j = 0;
        for (auto& pi : p->group->patterns)
            for (uint64_t v : pi.variations) out[j++] = v;
    });
}

int cad_patterns_indices_host(const cad_patterns* p, const uint64_t* pos, int64_t n, uint32_t* out)
{
    return guarded([&] {
        need(p && (pos || n == 0) && (out || n == 0) && n >= 0, "invalid argument");
        const int k = p->group->k();
        for (int64_t r = 0; r < n; r++) p->group->row(pos[2 * r], pos[2 * r + 1], out + r * k);
    });
}

int cad_index_extract(cad_patterns* p, const uint64_t* d_pos, int64_t n, uint32_t* d_out, int layout)
{
    return guarded([&] {
        need(p && n >= 0 && (n == 0 || (d_pos && d_out)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        need(layout == CAD_LAYOUT_ROW_MAJOR || layout == CAD_LAYOUT_SLOT_MAJOR, "unknown layout");
        const int k = p->group->k();
        if (layout == CAD_LAYOUT_ROW_MAJOR) extract(p, d_pos, n, d_out, 4, 1, k, 0);
        else extract(p, d_pos, n, d_out, 4, n, 1, 0);
    });
}

// ------------------------------------------------------------ matrices ----
namespace {

// Index matrix of positions straight from the device; row_begin (nsys + 1
// entries) splits the rows into independent systems, each with its own copy of
// the pattern group's columns (system s owns columns [s * cols, (s + 1) * cols)).
cad_matrix* matrix_from_positions(cad_patterns* p, const uint64_t* d_pos, const std::vector<int64_t>& row_begin)
{
    need(p->ctx != nullptr, "host-only pattern group (created without a context)");
    const PatternGroup& g = *p->group;
    const int nsys = int(row_begin.size()) - 1;
    const int64_t n = row_begin.back();
    const int64_t cols_per = int64_t(g.cols());
    need(cols_per * nsys <= int64_t(0xFFFFFFFFu), "more than 2^32 columns over all systems");
    std::unique_ptr<cad_matrix> m(new_matrix(p->ctx, n, g.k(), cols_per * nsys));
    m->slot_lo = p->slot_lo;
    m->slot_hi = p->slot_hi;
    uint32_t widest = 0;
    for (size_t s = 0; s < m->slot_lo.size(); s++) widest = std::max(widest, m->slot_hi[s] - m->slot_lo[s]);
    m->ibytes = widest <= 65536 ? 2 : 4;
    m->tune.rpt4 = m->ibytes == 2 ? 8 : 4;
    if (nsys == 1) default_schedule(m->tune, m->ibytes, n);
    m->ld = padded_rows(n);
    m->d_idx = DBuf(size_t(m->ibytes) * size_t(m->ld) * size_t(g.k()));
    CAD_HIP(hipMemsetAsync(m->d_idx.p, 0, m->d_idx.bytes, p->ctx->stream));
    std::vector<uint32_t> base(g.k());
    for (int s = 0; s < g.k(); s++) base[s] = m->ibytes == 2 ? m->slot_lo[s] : 0u;
    m->d_base = to_device(base, p->ctx->stream);
    extract(p, d_pos, n, m->d_idx.p, m->ibytes, m->ld, 1, m->ibytes == 2 ? 1 : 0);
    m->nsys = nsys;
    m->sys_row_begin = row_begin;
    m->cols_per_sys = cols_per;
    m->sys_col.resize(nsys);
    for (int s = 0; s < nsys; s++) m->sys_col[s] = uint32_t(cols_per * s);
    build_plans(m.get());
    CAD_HIP(hipStreamSynchronize(p->ctx->stream));
    return m.release();
}

}  // namespace

int cad_matrix_from_positions(cad_patterns* p, const uint64_t* d_pos, int64_t n, cad_matrix** out)
{
    return guarded([&] {
        need(p && out && n > 0 && d_pos, "invalid argument");
        *out = matrix_from_positions(p, d_pos, {0, n});
    });
}

int cad_matrix_from_positions_systems(cad_patterns* p, const uint64_t* d_pos, const int64_t* row_begin, int nsys,
                                      cad_matrix** out)
{
    return guarded([&] {
        need(p && out && d_pos && row_begin && nsys >= 1, "invalid argument");
        need(row_begin[0] == 0, "row_begin[0] must be 0");
        std::vector<int64_t> rb(row_begin, row_begin + nsys + 1);
        for (int s = 0; s < nsys; s++)
            need(rb[s + 1] > rb[s], "every system needs at least one row", CAD_ERR_SIZE_MISMATCH);
        *out = matrix_from_positions(p, d_pos, rb);
    });
}

int cad_matrix_from_positions_mf(cad_patterns* p, const uint64_t* d_pos, int64_t n, cad_matrix** out)
{
    return guarded([&] {
        need(p && out && n > 0 && d_pos, "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        const PatternGroup& g = *p->group;
        std::unique_ptr<cad_matrix> m(new_matrix(p->ctx, n, g.k(), int64_t(g.cols())));
        hipStream_t s = p->ctx->stream;
        m->mf = true;
        m->ibytes = 0;
        m->slot_lo = p->slot_lo;
        m->slot_hi = p->slot_hi;
        m->d_base = to_device(std::vector<uint32_t>(g.k(), 0u), s);
        auto copy = [&](DBuf& dst, const DBuf& src) {
            dst = DBuf(std::max<size_t>(src.bytes, 16));
            if (src.bytes) CAD_HIP(hipMemcpyAsync(dst.p, src.p, src.bytes, hipMemcpyDeviceToDevice, s));
        };
        copy(m->d_slots, p->d_slots);
        copy(m->d_sq, p->d_sq);
        copy(m->d_dense, p->d_dense);
        copy(m->d_tslots, p->d_tslots);
        copy(m->d_tern, p->d_tern);
        m->d_pos = DBuf(sizeof(uint64_t) * 2 * size_t(n));
        CAD_HIP(hipMemcpyAsync(m->d_pos.p, d_pos, m->d_pos.bytes, hipMemcpyDeviceToDevice, s));
        m->d_tw = DBuf(sizeof(uint32_t) * 4 * size_t(n));
        hipLaunchKernelGGL(k_ternary_rows, dim3(unsigned(std::min<int64_t>(grid1(n, 256), 65536))), dim3(256), 0, s,
                           (const uint64_t*)m->d_pos.as<uint64_t>(), n, m->d_tw.as<uint4>());
        CAD_HIP(hipGetLastError());
        build_plans(m.get());
        CAD_HIP(hipStreamSynchronize(s));
        *out = m.release();
    });
}

int cad_matrix_autotune(cad_matrix* m, char* log, size_t log_size)
{
    return guarded([&] {
        need(m != nullptr, "null matrix");
        const std::string l = autotune(m);
        if (log && log_size) {
            const size_t n = std::min(log_size - 1, l.size());
            std::memcpy(log, l.data(), n);
            log[n] = 0;
        }
    });
}

int cad_matrix_kernel_times(cad_matrix* m, int reps, double* us, int n)
{
    return guarded([&] {
        need(m != nullptr && us != nullptr && n >= 0, "null argument");
        need(reps >= 1, "reps must be positive");
        kernel_times(m, reps, us, n);
    });
}

int cad_matrix_get_schedule(const cad_matrix* m, cad_schedule* out)
{
    return guarded([&] {
        need(m != nullptr && out != nullptr, "null argument");
        const TuneParams& t = m->tune;
        *out = cad_schedule{t.rpt4, t.bs4, t.sub4, t.rpt8, t.bs8, t.sub8, t.srpt, t.scatter_order, t.blocks_per_cu,
                            t.partial_cap, t.scatter_balance};
    });
}

int cad_matrix_set_schedule(cad_matrix* m, const cad_schedule* sc)
{
    return guarded([&] {
        need(m != nullptr && sc != nullptr, "null argument");
        TuneParams t;
        t.rpt4 = sc->gather_f32_rpt;
        t.bs4 = sc->gather_f32_block;
        t.sub4 = sc->gather_f32_sub;
        t.rpt8 = sc->gather_f64_rpt;
        t.bs8 = sc->gather_f64_block;
        t.sub8 = sc->gather_f64_sub;
        t.srpt = sc->scatter_rpt;
        t.scatter_order = sc->scatter_order;
        t.blocks_per_cu = sc->scatter_tiles_per_cu;
        t.partial_cap = sc->scatter_partial_cap;
        t.scatter_balance = sc->scatter_balance;
        set_schedule(m, t);
    });
}

int cad_matrix_systems(const cad_matrix* m, int* nsys, int64_t* cols_per_system, int64_t* row_begin)
{
    return guarded([&] {
        need(m != nullptr, "null matrix");
        if (nsys) *nsys = m->nsys;
        if (cols_per_system) *cols_per_system = m->cols_per_sys;
        if (row_begin)
            for (int s = 0; s <= m->nsys; s++) row_begin[s] = m->sys_row_begin[s];
    });
}

int cad_matrix_create(cad_context* ctx, const uint32_t* h_idx, int64_t rows, int k, int64_t cols, cad_matrix** out)
{
    return guarded([&] {
        need(ctx && out && (h_idx || rows == 0), "invalid argument");
        std::unique_ptr<cad_matrix> m(new_matrix(ctx, rows, k, cols));
        std::vector<uint32_t> lo(k, 0xFFFFFFFFu), hi(k, 0u);
        for (int64_t r = 0; r < rows; r++)
            for (int s = 0; s < k; s++) {
                const uint32_t c = h_idx[r * k + s];
                need(int64_t(c) < cols, "column index out of range", CAD_ERR_SIZE_MISMATCH);
                lo[s] = std::min(lo[s], c);
                hi[s] = std::max(hi[s], c + 1);
            }
        uint32_t widest = 0;
        for (int s = 0; s < k; s++) {
            if (lo[s] > hi[s]) lo[s] = hi[s] = 0;  // no rows
            widest = std::max(widest, hi[s] - lo[s]);
        }
        m->slot_lo = lo;
        m->slot_hi = hi;
        m->ibytes = widest <= 65536 ? 2 : 4;
        default_schedule(m->tune, m->ibytes, rows);
        std::vector<uint32_t> base(k);
        for (int s = 0; s < k; s++) base[s] = m->ibytes == 2 ? lo[s] : 0u;
        m->d_base = to_device(base, ctx->stream);
        // transpose to slot-major (and narrow) on the host; padding rows are 0
        m->ld = padded_rows(rows);
        const size_t ld = size_t(m->ld);
        if (m->ibytes == 2) {
            std::vector<uint16_t> t(std::max<size_t>(1, ld * k), 0);
            for (int64_t r = 0; r < rows; r++)
                for (int s = 0; s < k; s++) t[size_t(s) * ld + r] = uint16_t(h_idx[r * k + s] - lo[s]);
            m->d_idx = to_device(t, ctx->stream);
        } else {
            std::vector<uint32_t> t(std::max<size_t>(1, ld * k), 0);
            for (int64_t r = 0; r < rows; r++)
                for (int s = 0; s < k; s++) t[size_t(s) * ld + r] = h_idx[r * k + s];
            m->d_idx = to_device(t, ctx->stream);
        }
        build_plans(m.get());
        *out = m.release();
    });
}

int cad_matrix_destroy(cad_matrix* m)
{
    return guarded([&] {
        if (!m) return;
        (void)hipStreamSynchronize(m->ctx->stream);
        delete m;
    });
}

int cad_matrix_info(const cad_matrix* m, int64_t* rows, int* k, int64_t* cols, int* ib)
{
    return guarded([&] {
        need(m != nullptr, "null matrix");
        if (rows) *rows = m->rows;
        if (k) *k = m->k;
        if (cols) *cols = m->cols;
        if (ib) *ib = m->ibytes;
    });
}

int cad_matrix_download(const cad_matrix* m, uint32_t* out)
{
    return guarded([&] {
        need(m && (out || m->rows == 0), "null argument");
        const int64_t R = m->rows;
        const int k = m->k;
        std::vector<uint32_t> base(k);
        hipStream_t s = m->ctx->stream;
        CAD_HIP(hipMemcpyAsync(base.data(), m->d_base.p, sizeof(uint32_t) * k, hipMemcpyDeviceToHost, s));
        // global column = system base + slot base + stored entry
        std::vector<uint32_t> rowbase(size_t(std::max<int64_t>(R, 0)), 0u);
        for (int sy = 0; sy < m->nsys; sy++)
            for (int64_t r = m->sys_row_begin[sy]; r < m->sys_row_begin[sy + 1]; r++) rowbase[r] = m->sys_col[sy];
        if (m->mf) {
            // recompute the indices (global columns, slot-major) from the positions
            std::vector<uint32_t> t(size_t(R) * k);
            if (R > 0) {
                DBuf tmp(t.size() * 4);
                hipLaunchKernelGGL((k_index_extract<uint32_t, kExtractRows>), dim3(grid1(R, kRowBlock * kExtractRows)),
                                   dim3(kRowBlock), 0, s,
                                   m->d_pos.as<uint64_t>(), R, m->d_slots.as<DevSlot>(), k,
                                   m->d_tslots.as<DevTern>(), m->d_tern.as<uint32_t>(), m->d_dense.as<uint32_t>(),
                                   tmp.as<uint32_t>(), R, int64_t(1), 0);
                CAD_HIP(hipGetLastError());
                CAD_HIP(hipMemcpyAsync(t.data(), tmp.p, t.size() * 4, hipMemcpyDeviceToHost, s));
                CAD_HIP(hipStreamSynchronize(s));
            }
            for (int64_t r = 0; r < R; r++)
                for (int j = 0; j < k; j++) out[r * k + j] = rowbase[r] + t[size_t(j) * R + r];
        } else if (m->ibytes == 2) {
            const size_t ld = size_t(m->ld);
            std::vector<uint16_t> t(ld * k);
            if (!t.empty()) CAD_HIP(hipMemcpyAsync(t.data(), m->d_idx.p, t.size() * 2, hipMemcpyDeviceToHost, s));
            CAD_HIP(hipStreamSynchronize(s));
            for (int64_t r = 0; r < R; r++)
                for (int j = 0; j < k; j++) out[r * k + j] = rowbase[r] + base[j] + t[size_t(j) * ld + r];
        } else {
            const size_t ld = size_t(m->ld);
            std::vector<uint32_t> t(ld * k);
            if (!t.empty()) CAD_HIP(hipMemcpyAsync(t.data(), m->d_idx.p, t.size() * 4, hipMemcpyDeviceToHost, s));
            CAD_HIP(hipStreamSynchronize(s));
            for (int64_t r = 0; r < R; r++)
                for (int j = 0; j < k; j++) out[r * k + j] = rowbase[r] + base[j] + t[size_t(j) * ld + r];
        }
    });
}

int cad_spmv_f32(cad_matrix* m, const float* x, float* y)
{
    return guarded([&] {
        need(m && x && (y || m->rows == 0), "null argument");
        if (m->rows) matrix_spmv_f32(m, x, y);
    });
}

int cad_spmv_f64(cad_matrix* m, const double* x, double* y)
{
    return guarded([&] {
        need(m && x && (y || m->rows == 0), "null argument");
        if (m->rows) matrix_spmv_f64(m, x, y);
    });
}

int cad_spmtv_f32(cad_matrix* m, const float* y, float* z)
{
    return guarded([&] {
        need(m && z && (y || m->rows == 0), "null argument");
        matrix_spmtv_f32(m, y, z);
    });
}

int cad_spmtv_f64(cad_matrix* m, const double* y, double* z)
{
    return guarded([&] {
        need(m && z && (y || m->rows == 0), "null argument");
        matrix_spmtv_f64(m, y, z);
    });
}

int cad_jacobi(cad_matrix* m, float* d)
{
    return guarded([&] {
        need(m && d, "null argument");
        matrix_jacobi(m, d);
    });
}

// ------------------------------------------------------------ solvers -----
void cad_solve_options_default(cad_solve_options* o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->solver = CAD_SOLVER_LSQR;
    o->precision = CAD_PREC_MIXED;
    o->precondition = 1;
    o->max_iterations = 100000;
    o->tol = 1e-10;
    o->inner_tol = 1e-5;
    o->max_refinements = 20;
    o->check_every = 10;
    o->use_graph = 1;
    o->verbose = 0;
}

namespace {

cad_solve_options checked_options(const cad_solve_options* opt)
{
    cad_solve_options o;
    if (opt) o = *opt;
    else cad_solve_options_default(&o);
    need(o.precision >= 0 && o.precision <= 2, "unknown precision");
    need(!(o.precision == CAD_PREC_MIXED && o.solver != CAD_SOLVER_LSQR), "mixed precision is LSQR only");
    return o;
}

}  // namespace

int cad_solve(cad_matrix* m, const double* b, double* x, const cad_solve_options* opt, cad_solve_report* rep)
{
    return guarded([&] {
        need(m && b && x, "null argument");
        const cad_solve_options o = checked_options(opt);
        cad_solve_report r;
        solve(m, b, x, o, r);
        if (rep) *rep = r;
    });
}

int cad_solve_systems(cad_matrix* m, const double* b, double* x, const cad_solve_options* opt,
                      cad_solve_report* rep, double* rel_per_system, int* iterations_per_system)
{
    return guarded([&] {
        need(m && b && x, "null argument");
        const cad_solve_options o = checked_options(opt);
        cad_solve_report r;
        SystemDetail det;
        solve(m, b, x, o, r, &det);
        if (rep) *rep = r;
        for (int s = 0; s < m->nsys; s++) {
            if (rel_per_system) rel_per_system[s] = det.rel[s];
            if (iterations_per_system) iterations_per_system[s] = det.iters[s];
        }
    });
}

int cad_rel_normal_residual(cad_matrix* m, const double* b, const double* x, double* out, double* rn)
{
    return guarded([&] {
        need(m && b && x && out, "null argument");
        *out = rel_normal_residual(m, b, x, rn);
    });
}

// ------------------------------------------------------------ evaluator ---
int cad_eval_forward_f32(cad_patterns* p, const int32_t* idx, int64_t B, const float* w, float* out)
{
    return guarded([&] {
        need(p && B >= 0 && (B == 0 || (idx && w && out)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        eval_forward<float>(p, idx, B, w, out);
    });
}

int cad_eval_forward_f64(cad_patterns* p, const int32_t* idx, int64_t B, const double* w, double* out)
{
    return guarded([&] {
        need(p && B >= 0 && (B == 0 || (idx && w && out)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        eval_forward<double>(p, idx, B, w, out);
    });
}

int cad_eval_backward_f32(cad_patterns* p, const int32_t* idx, int64_t B, const float* g, float* grad)
{
    return guarded([&] {
        need(p && grad && B >= 0 && (B == 0 || (idx && g)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        eval_backward<float>(p, idx, B, g, grad);
    });
}

int cad_eval_backward_f64(cad_patterns* p, const int32_t* idx, int64_t B, const double* g, double* grad)
{
    return guarded([&] {
        need(p && grad && B >= 0 && (B == 0 || (idx && g)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        eval_backward<double>(p, idx, B, g, grad);
    });
}

int cad_eval_train_step_f32(cad_patterns* p, const int32_t* idx, const float* target, const int64_t* sel, int64_t B,
                            float delta, float lr, float beta1, float beta2, float eps, int64_t step, float* w,
                            float* m, float* v, double* loss)
{
    return guarded([&] {
        need(p && w && m && v && loss && B >= 0 && (B == 0 || (idx && target && sel)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        need(step >= 1, "Adam steps count from 1");
        need(delta > 0.0f && lr > 0.0f && beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f &&
                 eps >= 0.0f,
             "invalid Huber or Adam parameter");
        train_step(p, idx, target, sel, B, delta, lr, beta1, beta2, eps, step, w, m, v, loss);
    });
}

int cad_eval_train_epoch_f32(cad_patterns* p, const int32_t* idx, const float* target, const int64_t* order,
                             int64_t n, int64_t batch, float delta, float lr, float beta1, float beta2, float eps,
                             int64_t step0, float* w, float* m, float* v, double* loss)
{
    return guarded([&] {
        need(p && w && m && v && loss && n >= 0 && batch > 0 && (n == 0 || (idx && target && order)),
             "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        need(step0 >= 1, "Adam steps count from 1");
        need(delta > 0.0f && lr > 0.0f && beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f &&
                 eps >= 0.0f,
             "invalid Huber or Adam parameter");
        int64_t step = step0;
        for (int64_t b0 = 0; b0 < n; b0 += batch, step++)
            train_step(p, idx, target, order + b0, std::min(batch, n - b0), delta, lr, beta1, beta2, eps, step, w, m,
                       v, loss);
    });
}

int cad_eval_backward_f32_noalloc(cad_patterns* p, const int32_t* idx, int64_t B, const float* g, float* grad)
{
    return guarded([&] {
        need(p && grad && B >= 0 && (B == 0 || (idx && g)), "invalid argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        const int64_t cols = int64_t(p->group->cols());
        const int k = p->group->k();
        need(cols <= kNoAllocScratchDoubles,
             "no-allocation backward: more than 4,194,304 columns (one partial slice beyond the 32 MiB scratch)");
        double* part = noalloc_scratch(p->ctx->device);
        hipStream_t s = p->ctx->stream;
        if (B == 0) {
            hipLaunchKernelGGL(k_fill<float>, dim3(grid1(cols, 256)), dim3(256), 0, s, grad, cols, 0.0f);
            CAD_HIP(hipGetLastError());
            return;
        }
        // windows: consecutive slots with the same column block form one
        // pattern block, cut into windows of at most kNoAllocWindow columns
        NoAllocWindows win;
        int nwin = 0;
        for (int s0 = 0; s0 < k;) {
            int s1 = s0 + 1;
            while (s1 < k && p->slot_lo[s1] == p->slot_lo[s0] && p->slot_hi[s1] == p->slot_hi[s0]) s1++;
            for (int64_t c0 = p->slot_lo[s0]; c0 < int64_t(p->slot_hi[s0]); c0 += kNoAllocWindow) {
                need(nwin < kNoAllocMaxWindows, "no-allocation backward: more than 128 column windows");
                win.w[nwin++] = NoAllocWindow{int32_t(c0),
                                              int32_t(std::min<int64_t>(c0 + kNoAllocWindow, p->slot_hi[s0])),
                                              s0, s1};
            }
            s0 = s1;
        }
        // row chunks: enough blocks for two per CU, at least 2,048 rows a
        // chunk, and every chunk's slice inside the scratch
        const int64_t max_chunks = std::min<int64_t>(kNoAllocScratchDoubles / cols, 65535);
        int64_t nchunk = std::max<int64_t>(1, (2 * int64_t(p->ctx->cus) + nwin - 1) / nwin);
        nchunk = std::min(nchunk, std::max<int64_t>(1, (B + 2047) / 2048));
        nchunk = std::min(nchunk, max_chunks);
        const int64_t rows_per_chunk = (B + nchunk - 1) / nchunk;
        nchunk = (B + rows_per_chunk - 1) / rows_per_chunk;
        hipLaunchKernelGGL(k_noalloc_grad_window, dim3(unsigned(nwin), unsigned(nchunk)), dim3(kNoAllocBlock),
                           size_t(kNoAllocWindow) * sizeof(double), s, idx, B, k, g, win, rows_per_chunk, cols, part);
        CAD_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_noalloc_grad_reduce, dim3(unsigned(std::min<int64_t>(grid1(cols, 256), 4096))),
                           dim3(256), 0, s, part, int(nchunk), cols, grad);
        CAD_HIP(hipGetLastError());
    });
}

// ------------------------------------------------------------ tables ------
int cad_patterns_expanded_size(cad_patterns* p, int64_t* n)
{
    return guarded([&] {
        need(p && n, "null argument");
        int64_t total = 0;
        for (const auto& pi : p->group->patterns)
            total += int64_t(pow3(__builtin_popcountll(pi.mask))) * int64_t(pi.variations.size());
        *n = total;
    });
}

int cad_weights_expand(cad_patterns* p, const float* w, float* tables)
{
    return guarded([&] {
        need(p && w && tables, "null argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        weights_expand(p, w, tables);
    });
}

int cad_weights_compact(cad_patterns* p, const float* tables, float* w)
{
    return guarded([&] {
        need(p && w && tables, "null argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        for (const auto& pi : p->group->patterns)
            need(__builtin_popcountll(pi.mask) <= 15, "compaction takes patterns of at most 15 squares");
        weights_compact(p, tables, w);
    });
}

int cad_evaluator_create(cad_patterns* p, const float* w, cad_evaluator** out)
{
    return guarded([&] {
        need(p && w && out, "null argument");
        need(p->ctx != nullptr, "host-only pattern group (created without a context)");
        std::unique_ptr<cad_evaluator> e(new cad_evaluator);
        e->ctx = p->ctx;
        e->patterns = p;
        e->tables = DBuf(sizeof(float) * size_t(std::max<int64_t>(1, expanded_layout(p).size)));
        weights_expand(p, w, e->tables.as<float>());
        *out = e.release();
    });
}

int cad_evaluator_destroy(cad_evaluator* e)
{
    return guarded([&] {
        if (!e) return;
        (void)hipStreamSynchronize(e->ctx->stream);
        delete e;
    });
}

int cad_evaluator_set_weights(cad_evaluator* e, const float* w)
{
    return guarded([&] {
        need(e && w, "null argument");
        weights_expand(e->patterns, w, e->tables.as<float>());
    });
}

int cad_evaluator_eval(cad_evaluator* e, const uint64_t* pos, int64_t n, float* out)
{
    return guarded([&] {
        need(e && n >= 0 && (n == 0 || (pos && out)), "invalid argument");
        eval_tables(e->patterns, e->tables.as<float>(), pos, n, out);
    });
}

int cad_evaluator_eval_host(cad_evaluator* e, const uint64_t* pos, int64_t n, float* out)
{
    return guarded([&] {
        need(e && n >= 0 && (n == 0 || (pos && out)), "invalid argument");
        if (n == 0) return;
        hipStream_t s = e->ctx->stream;
        if (n > e->staged) {
            CAD_HIP(hipStreamSynchronize(s));
            e->d_pos = DBuf(sizeof(uint64_t) * 2 * size_t(n));
            e->d_out = DBuf(sizeof(float) * size_t(n));
            e->staged = n;
        }
        CAD_HIP(hipMemcpyAsync(e->d_pos.p, pos, sizeof(uint64_t) * 2 * size_t(n), hipMemcpyHostToDevice, s));
        eval_tables(e->patterns, e->tables.as<float>(), e->d_pos.as<uint64_t>(), n, e->d_out.as<float>());
        CAD_HIP(hipMemcpyAsync(out, e->d_out.p, sizeof(float) * size_t(n), hipMemcpyDeviceToHost, s));
        CAD_HIP(hipStreamSynchronize(s));
    });
}

// ------------------------------------------------------------ data --------
int cad_random_play(cad_context* ctx, uint64_t seed, int64_t n, int e_lo, int e_hi, uint64_t* out)
{
    return guarded([&] {
        need(ctx && n >= 0 && (n == 0 || out), "invalid argument");
        need(e_lo >= 0 && e_hi <= 60 && e_lo <= e_hi, "empty count outside [0, 60]");
        if (!n) return;
        hipLaunchKernelGGL(k_random_play, dim3(grid1(n, kRowBlock)), dim3(kRowBlock), 0, ctx->stream, seed, n, e_lo,
                           e_hi, out);
        CAD_HIP(hipGetLastError());
    });
}

uint64_t cad_flip(int sym, uint64_t b) { return apply_sym(sym, b); }
uint64_t cad_possible_moves(uint64_t P, uint64_t O) { return possible_moves(P, O); }
uint64_t cad_flips(uint64_t P, uint64_t O, int move) { return (move >= 0 && move < 64) ? flips(P, O, move) : 0; }

int cad_random_play_host(uint64_t seed, int64_t n, int e_lo, int e_hi, uint64_t* out)
{
    return guarded([&] {
        need(n >= 0 && (n == 0 || out), "invalid argument");
        need(e_lo >= 0 && e_hi <= 60 && e_lo <= e_hi, "empty count outside [0, 60]");
        for (int64_t i = 0; i < n; i++) {
            const Pos p = random_play(seed, uint64_t(i), e_lo, e_hi);
            out[2 * i] = p.P;
            out[2 * i + 1] = p.O;
        }
    });
}

int cad_make_targets(cad_matrix* m, uint64_t seed_x, uint64_t seed_n, double sx, double sn, double* xt, double* b)
{
    return guarded([&] {
        need(m && xt && (b || m->rows == 0), "null argument");
        hipStream_t s = m->ctx->stream;
        hipLaunchKernelGGL(k_targets_x, dim3(grid1(m->cols, 256)), dim3(256), 0, s, seed_x, m->cols, sx, xt);
        CAD_HIP(hipGetLastError());
        if (!m->rows) return;
        matrix_targets(m, xt, seed_n, sn, b);
        CAD_HIP(hipGetLastError());
    });
}

}  // extern "C"
