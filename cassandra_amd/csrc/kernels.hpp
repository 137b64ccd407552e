// HIP kernels of the pattern-evaluator fit (gfx950 / CDNA4).
//
// Layout of a stored matrix (fixed k entries per row, implicit value 1):
//   slot-major: idx[s * ld + r], r < rows, s < k, column = sys_base + base[s] + idx
//   (u16 when every slot's column range is narrower than 65536, else u32).
// Every row-parallel kernel walks a RowWork list, every column-parallel kernel a
// ColWork list and the scatter kernels ScatterWork / ReduceWork lists, so one
// launch covers any number of independent systems (block-diagonal batches).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "board.hpp"
#include "device.hpp"

namespace cad {
// Kernels have internal linkage: every translation unit instantiates the ones
// it launches.
namespace {

constexpr int kRowBlock = 256;   // threads of the row / column kernels
constexpr int kRowsPerThread = 4;  // rows per thread of the generic row work
constexpr int kScatterBlock = 1024;

// ------------------------------------------------------------ index views --
template <class IdxT>
struct SlotMajor {
    using index_type = IdxT;
    const IdxT* __restrict__ idx;
    int64_t ld;
    const uint32_t* __restrict__ base;  // per slot
    __device__ __forceinline__ uint32_t col(int s, int64_t r) const
    {
        return base[s] + uint32_t(idx[int64_t(s) * ld + r]);
    }
    // Slot s's entries, for a loop over rows (s the same for the whole wave:
    // the pointer and base stay in scalar registers, and the rows' loads of
    // several slots can all be in flight together).
    struct Slot {
        const IdxT* __restrict__ p;
        uint32_t base;
        __device__ __forceinline__ uint32_t col(int64_t r) const { return base + uint32_t(p[r]); }
    };
    __device__ __forceinline__ Slot slot(int s) const
    {
        return Slot{idx + int64_t(s) * ld, uint32_t(__builtin_amdgcn_readfirstlane(base[s]))};
    }
};

template <class View>
struct is_slot_major : std::false_type {};
template <class IdxT>
struct is_slot_major<SlotMajor<IdxT>> : std::true_type {};
template <class View>
struct SlotRefOf {
    using type = int;
};
template <class IdxT>
struct SlotRefOf<SlotMajor<IdxT>> {
    using type = typename SlotMajor<IdxT>::Slot;
};

// Row-major int32 indices as torch hands them over ([B, k], contiguous).
struct RowMajor32 {
    const int32_t* __restrict__ idx;
    int k;
    __device__ __forceinline__ uint32_t col(int s, int64_t r) const
    {
        return uint32_t(idx[r * k + s]);
    }
};

// ------------------------------------------------------------ reductions ---
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over the block; the result is valid in thread 0.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < BS / 64; i++) t += sh[i];
    return t;
}

// Single-pass grid reduction: every block of a system stores its partial, the
// last block to arrive (ticket on the system's counter) sees all partials and
// returns true.  The counter is re-armed by that last block, so the kernel can
// be replayed (graphs) without clearing it.
__device__ __forceinline__ bool arrive_last(unsigned* counter, unsigned nblocks)
{
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        unsigned t = atomicAdd(counter, 1u);
        last = (t == nblocks - 1);
        if (last) atomicExch(counter, 0u);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// Sum of one system's block partials (slot 0), returned to every thread:
// each thread adds its strided share in block order, then the block's fixed
// reduction tree, so every block that forms the sum gets the same value.
template <int BS>
__device__ __forceinline__ double system_sum(const GridSys& g, int sys, double* sh)
{
    double s = 0.0;
    const int first = g.blk_first[sys], n = g.blk_count[sys];
    for (int i = threadIdx.x; i < n; i += BS) s += g.partials[2 * (first + i)];
    s = block_sum<BS>(s, sh);
    if (threadIdx.x == 0) sh[0] = s;
    __syncthreads();
    return sh[0];
}

// The same for both slots of the partials.
template <int BS>
__device__ __forceinline__ void system_sum2(const GridSys& g, int sys, double* sh, double& s0, double& s1)
{
    double a0 = 0.0, a1 = 0.0;
    const int first = g.blk_first[sys], n = g.blk_count[sys];
    for (int i = threadIdx.x; i < n; i += BS) {
        a0 += g.partials[2 * (first + i)];
        a1 += g.partials[2 * (first + i) + 1];
    }
    a0 = block_sum<BS>(a0, sh);
    a1 = block_sum<BS>(a1, sh);
    if (threadIdx.x == 0) {
        sh[0] = a0;
        sh[1] = a1;
    }
    __syncthreads();
    s0 = sh[0];
    s1 = sh[1];
}

// ------------------------------------------------------------ fill --------
template <class T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = v;
}

template <class TO, class TI>
__global__ void k_convert(TO* __restrict__ out, const TI* __restrict__ in, int64_t n)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = TO(in[i]);
}

// ------------------------------------------------------------ index --------
// Pattern/Indexer.cpp InsertIndices for every row: one thread per position.
// out[s * slot_stride + r * row_stride] = (local ? 0 : col_offset) + index.
__device__ __forceinline__ uint32_t list_value(const uint8_t* __restrict__ sq, int off, int len,
                                               uint64_t P, uint64_t O)
{
    // the square numbers are the same for the whole wave: read them as 32-bit
    // words (scalar loads; the array is padded to whole words) and take the
    // byte, rather than one vector byte load per square
    const uint32_t* __restrict__ w = reinterpret_cast<const uint32_t*>(sq);
    uint32_t v = 0;
    for (int j = len - 1; j >= 0; --j) {
        const int b = off + j;
        const int q = int((w[b >> 2] >> (8 * (b & 3))) & 0xFFu);
        v = v * 3u + uint32_t((P >> q) & 1u) + 2u * uint32_t((O >> q) & 1u);
    }
    return v;
}

__device__ __forceinline__ uint32_t slot_index(const DevSlot& s, const uint8_t* __restrict__ sq,
                                               const uint32_t* __restrict__ dense, uint64_t P, uint64_t O)
{
    const uint32_t a = list_value(sq, s.off[0], s.len[0], P, O);
    if (s.kind == 0) return a;
    if (s.kind == 3) return dense[s.dense_offset + a];
    const uint32_t b = list_value(sq, s.off[1], s.len[1], P, O);
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    uint32_t idx = lo * s.hs + hi - lo * (lo + 1u) / 2u;
    if (s.kind == 2) idx = idx * s.ds + list_value(sq, s.off[2], s.len[2], P, O);
    return idx;
}



// The four ternary words of a position: square q's digit P_q + 2 O_q in bits
// 2 (q mod 16) of word q / 16 (each 16-bit half-board's bits spread to every
// other bit).
__device__ __forceinline__ uint32_t spread16(uint32_t x)
{
    x &= 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

__device__ __forceinline__ void ternary_words(uint64_t P, uint64_t O, uint32_t (&T)[4])
{
#pragma unroll
    for (int w = 0; w < 4; w++)
        T[w] = spread16(uint32_t(P >> (16 * w))) | (spread16(uint32_t(O >> (16 * w))) << 1);
}

// Read-only memory at a wave-uniform address through the constant address
// space: scalar loads into scalar registers.  Only for arrays no kernel
// writes (programs uploaded before the launch).  Scalar loads and LDS
// operations share one wait counter, so a wait for one waits for the other.
typedef const __attribute__((address_space(4))) uint32_t* ConstWords;

// A descriptor that is the same for the whole wave, read as whole 32-bit
// words into scalar registers (a struct with byte or half-word fields would
// otherwise be read with vector loads and full waits).  Vector loads and
// lane-to-scalar moves: their wait does not also wait for the LDS operations
// in flight, as a scalar load's would.
template <class T>
__device__ __forceinline__ T uniform_load(const T* p)
{
    static_assert(sizeof(T) % 4 == 0, "whole words");
    const uint32_t* __restrict__ w = reinterpret_cast<const uint32_t*>(p);
    uint32_t r[sizeof(T) / 4];
#pragma unroll
    for (int i = 0; i < int(sizeof(T) / 4); i++) r[i] = __builtin_amdgcn_readfirstlane(w[i]);
    T t;
    __builtin_memcpy(&t, r, sizeof(T));
    return t;
}

// Element i of a uint16 array the whole wave reads (the array is padded to
// whole 32-bit words).
__device__ __forceinline__ int uniform_u16(const uint16_t* p, int i)
{
    const uint32_t w = __builtin_amdgcn_readfirstlane(reinterpret_cast<const uint32_t*>(p)[i >> 1]);
    return int((w >> (16 * (i & 1))) & 0xFFFFu);
}

// Four entries of a ternary-word program (two words each: the digit's shift,
// then 3^j) at a wave-uniform address, through the constant address space so
// that they load into scalar registers with one 32-byte load (the kernels
// never write the arrays read this way).
struct TernEntries {
    uint32_t w[8];
};
__device__ __forceinline__ TernEntries scalar_load_entries(const uint32_t* p)
{
    const ConstWords c = (ConstWords)(p);
    TernEntries e;
#pragma unroll
    for (int i = 0; i < 8; i++) e.w[i] = c[i];
    return e;
}

// Adds entry q's term, digit x 3^j, to RPT rows (word W of each row's
// ternary words): one bit-field extract (shift straight from the entry) and
// one 24-bit multiply-add a row.
template <int RPT, int W>
__device__ __forceinline__ void tern_term(const TernEntries& e, int q, const uint32_t (&T)[RPT][4],
                                          uint32_t (&v)[RPT])
{
#pragma unroll
    for (int i = 0; i < RPT; i++) v[i] += __umul24((T[i][W] >> e.w[2 * q]) & 3u, e.w[2 * q + 1]);
}

// The n entries of word W from entry `off` on: whole groups of four, then the
// last one to three behind uniform branches (no padding entries to evaluate).
template <int RPT, int W>
__device__ __forceinline__ uint32_t tern_word(const uint32_t* __restrict__ ent, uint32_t off, uint32_t n,
                                              const uint32_t (&T)[RPT][4], uint32_t (&v)[RPT])
{
    for (uint32_t g = n >> 2; g > 0; g--, off += 4) {
        const TernEntries e = scalar_load_entries(ent + 2 * off);
#pragma unroll
        for (int q = 0; q < 4; q++) tern_term<RPT, W>(e, q, T, v);
    }
    const uint32_t r = n & 3u;
    if (r) {
        const TernEntries e = scalar_load_entries(ent + 2 * off);
        tern_term<RPT, W>(e, 0, T, v);
        if (r > 1) tern_term<RPT, W>(e, 1, T, v);
        if (r > 2) tern_term<RPT, W>(e, 2, T, v);
    }
    return off + r;
}

// The value sum_j 3^j (P[q_j] + 2 O[q_j]) of one square list for RPT rows,
// from the list's ternary-word program (3^j < 2^24: the host builds no list of
// more than 16 squares).  The entries are the same for the whole wave.
template <int RPT>
__device__ __forceinline__ void tern_values(const uint32_t* __restrict__ ent, uint32_t off, uint32_t nent,
                                            const uint32_t (&T)[RPT][4], uint32_t (&v)[RPT])
{
#pragma unroll
    for (int i = 0; i < RPT; i++) v[i] = 0u;
    off = tern_word<RPT, 0>(ent, off, nent & 0xFFu, T, v);
    off = tern_word<RPT, 1>(ent, off, (nent >> 8) & 0xFFu, T, v);
    off = tern_word<RPT, 2>(ent, off, (nent >> 16) & 0xFFu, T, v);
    tern_word<RPT, 3>(ent, off, nent >> 24, T, v);
}

// slot_index for RPT rows from their ternary words (the same values).
template <int RPT>
__device__ __forceinline__ void slot_index_rows(const DevSlot& s, const DevTern& t, const uint32_t* __restrict__ ent,
                                                const uint32_t* __restrict__ dense, const uint32_t (&T)[RPT][4],
                                                uint32_t (&idx)[RPT])
{
    tern_values<RPT>(ent, t.off[0], t.nent[0], T, idx);
    if (s.kind == 0) return;
    if (s.kind == 3) {
#pragma unroll
        for (int i = 0; i < RPT; i++) idx[i] = dense[s.dense_offset + idx[i]];
        return;
    }
    uint32_t b[RPT];
    tern_values<RPT>(ent, t.off[1], t.nent[1], T, b);
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        const uint32_t lo = idx[i] < b[i] ? idx[i] : b[i], hi = idx[i] < b[i] ? b[i] : idx[i];
        idx[i] = lo * s.hs + hi - lo * (lo + 1u) / 2u;
    }
    if (s.kind == 2) {
        tern_values<RPT>(ent, t.off[2], t.nent[2], T, b);
#pragma unroll
        for (int i = 0; i < RPT; i++) idx[i] = idx[i] * s.ds + b[i];
    }
}

// The ternary words of n positions (a matrix-free matrix keeps them, so the
// products need not recompute them).
__global__ void k_ternary_rows(const uint64_t* __restrict__ pos, int64_t n, uint4* __restrict__ tw)
{
    for (int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; r < n; r += int64_t(gridDim.x) * blockDim.x) {
        uint32_t T[4];
        ternary_words(pos[2 * r], pos[2 * r + 1], T);
        tw[r] = make_uint4(T[0], T[1], T[2], T[3]);
    }
}

// Pattern/Indexer.cpp InsertIndices for every row, RPT positions a thread
// (rows r0 + i kRowBlock, so that each slot's stores of a wave stay
// contiguous): out[s * slot_stride + r * row_stride] = (local ? 0 :
// col_offset) + index, each slot's index from the position's ternary words.
template <class OutT, int RPT>
__global__ __launch_bounds__(kRowBlock) void k_index_extract(
    const uint64_t* __restrict__ pos, int64_t n, const DevSlot* __restrict__ slots, int k,
    const DevTern* __restrict__ tslots, const uint32_t* __restrict__ tern, const uint32_t* __restrict__ dense,
    OutT* __restrict__ out, int64_t slot_stride, int64_t row_stride, int local)
{
    const int64_t r0 = blockIdx.x * int64_t(kRowBlock * RPT) + threadIdx.x;
    if (r0 >= n) return;
    uint32_t T[RPT][4];
    bool ok[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        const int64_t r = r0 + int64_t(i) * kRowBlock;
        ok[i] = r < n;
        const int64_t rr = ok[i] ? r : r0;  // a valid row's words for the lanes past the end
        ternary_words(pos[2 * rr], pos[2 * rr + 1], T[i]);
    }
    for (int s = 0; s < k; s++) {
        const DevSlot sl = uniform_load(slots + s);
        uint32_t v[RPT];
        slot_index_rows<RPT>(sl, uniform_load(tslots + s), tern, dense, T, v);
        const uint32_t base = local ? 0u : sl.col_offset;
#pragma unroll
        for (int i = 0; i < RPT; i++)
            if (ok[i]) out[int64_t(s) * slot_stride + (r0 + int64_t(i) * kRowBlock) * row_stride] = OutT(base + v[i]);
    }
}

// Positions a thread of k_index_extract.
constexpr int kExtractRows = 8;

// Matrix-free view of a pattern matrix: the column of slot s in row r is
// recomputed from the row's position (16 bytes) with the slot's program
// instead of being read from a stored index matrix.
struct PositionsView {
    using index_type = void;  // no stored indices
    const uint64_t* __restrict__ pos;
    const DevSlot* __restrict__ slots;
    const uint8_t* __restrict__ sq;
    const uint32_t* __restrict__ dense;
    const DevTern* __restrict__ tslots;  // the ternary-word programs of the same slots
    const uint32_t* __restrict__ tern;
    const uint4* __restrict__ tw;        // every row's ternary words (computed once, at creation)
    __device__ __forceinline__ uint32_t col(int s, int64_t r) const
    {
        const DevSlot sl = slots[s];
        return sl.col_offset + slot_index(sl, sq, dense, pos[2 * r], pos[2 * r + 1]);
    }
};

// ------------------------------------------------------------ gathers ------
// y = A x with the in-row sum taken sequentially in slot order from zero, the
// order of Math/MatrixCSR.cpp:59-62 (bit-exact with it in FP32).  What is
// done with each row sum is an epilogue:
//   bool begin(sys)             per block; false skips the block (inactive system)
//   double row(r, acc)          store the row; returns its share of a reduction
//   void finish(sys, total)     last block of a system (kReduce only)
template <class VT>
struct EpiStore {
    VT* __restrict__ y;
    static constexpr bool kReduce = false;
    __device__ bool begin(int) { return true; }
    __device__ double row(int64_t r, VT a) { y[r] = a; return 0.0; }
    // RPT consecutive rows from a (a multiple of RPT), all of them valid:
    // 16-byte stores where y + a is 16-byte aligned
    template <int RPT>
    __device__ double rows(int64_t a, const VT (&acc)[RPT])
    {
        constexpr int Q = 16 / sizeof(VT);
        if constexpr (RPT % Q == 0) {
            if ((reinterpret_cast<uintptr_t>(y + a) & 15) == 0) {
                using V4 = std::conditional_t<sizeof(VT) == 4, float4, double2>;
#pragma unroll
                for (int i = 0; i < RPT; i += Q) {
                    V4 v;
                    __builtin_memcpy(&v, &acc[i], 16);
                    *reinterpret_cast<V4*>(y + a + i) = v;
                }
                return 0.0;
            }
        }
#pragma unroll
        for (int i = 0; i < RPT; i++) y[a + i] = acc[i];
        return 0.0;
    }
    __device__ void finish(int, double) {}
};

// LSQR step 1 (Math/Solver.h:161): u <- A v - alpha u.  u is kept unnormalised
// as u_raw with u = u_raw / beta; the gather source is vd = D v.  The new beta
// is |u_raw| (FP64 reduction).
template <class VT>
struct EpiLsqrU {
    VT* __restrict__ u;
    LsqrScalars* scal;
    VT f;
    static constexpr bool kReduce = true;
    static constexpr bool kFinishKernel = true;
    __device__ bool begin(int sys)
    {
        const LsqrScalars* st = &scal[sys];
        if (!st->active || st->stop) return false;
        const double beta = st->beta;
        f = VT(beta != 0.0 ? st->alpha / beta : 0.0);
        return true;
    }
    __device__ double row(int64_t r, VT a)
    {
        const VT v = a - f * u[r];
        u[r] = v;
        return double(v) * double(v);
    }
    // RPT consecutive valid rows from a: 16-byte loads and stores of u where
    // aligned; the squares add up in row order, as row() would
    template <int RPT>
    __device__ double rows(int64_t a, const VT (&acc)[RPT])
    {
        VT o[RPT];
        constexpr int Q = 16 / sizeof(VT);
        bool vec = false;
        if constexpr (RPT % Q == 0) vec = (reinterpret_cast<uintptr_t>(u + a) & 15) == 0;
        using V4 = std::conditional_t<sizeof(VT) == 4, float4, double2>;
        if constexpr (RPT % Q == 0) {
            if (vec) {
#pragma unroll
                for (int i = 0; i < RPT; i += Q) {
                    const V4 t = *reinterpret_cast<const V4*>(u + a + i);
                    __builtin_memcpy(&o[i], &t, 16);
                }
            }
        }
        if (!vec) {
#pragma unroll
            for (int i = 0; i < RPT; i++) o[i] = u[a + i];
        }
        double red = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            o[i] = acc[i] - f * o[i];
            red += double(o[i]) * double(o[i]);
        }
        if constexpr (RPT % Q == 0) {
            if (vec) {
#pragma unroll
                for (int i = 0; i < RPT; i += Q) {
                    V4 t;
                    __builtin_memcpy(&t, &o[i], 16);
                    *reinterpret_cast<V4*>(u + a + i) = t;
                }
                return red;
            }
        }
#pragma unroll
        for (int i = 0; i < RPT; i++) u[a + i] = o[i];
        return red;
    }
    // beta = |u_raw|; its square is kept for the rotation, which adds it to
    // the Frobenius estimate (as the scatter that forms beta itself does)
    __device__ void finish(int sys, double s)
    {
        scal[sys].beta = sqrt(s);
        scal[sys].beta2 = s;
    }
};

// FP64 residual of the original system: r = b - A x into r64, |r|^2 into
// reduce_a and beta = |r|; when u is given it also receives r converted to VT
// (the start vector of a refinement solve).
template <class VT>
struct EpiResidual {
    const double* __restrict__ b;
    double* __restrict__ r64;
    VT* __restrict__ u;
    LsqrScalars* scal;
    static constexpr bool kReduce = true;
    static constexpr bool kFinishKernel = true;
    __device__ bool begin(int sys) { return scal[sys].active != 0 && scal[sys].restart != 0; }
    __device__ double row(int64_t r, double a)
    {
        const double v = b[r] - a;
        r64[r] = v;
        if (u) u[r] = VT(v);
        return v * v;
    }
    template <int RPT>
    __device__ double rows(int64_t a, const double (&acc)[RPT])
    {
        double red = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++) red += row(a + i, acc[i]);
        return red;
    }
    __device__ void finish(int sys, double s)
    {
        scal[sys].reduce_a = s;
        scal[sys].beta = sqrt(s);
        scal[sys].anorm2 = 0.0;
    }
};

// b = clip(2 * floor((A x_true + sn * N) / 2 + 1/2), -64, 64)
struct EpiTargets {
    double* __restrict__ b;
    uint64_t seed;
    double sn;
    static constexpr bool kReduce = false;
    __device__ bool begin(int) { return true; }
    __device__ double row(int64_t r, double a)
    {
        Stream rng(seed, uint64_t(r));
        const double v = __dadd_rn(a, __dmul_rn(sn, rng.ih12()));
        const double e = 2.0 * floor(v * 0.5 + 0.5);
        b[r] = fmin(64.0, fmax(-64.0, e));
        return 0.0;
    }
    template <int RPT>
    __device__ double rows(int64_t a, const double (&acc)[RPT])
    {
        double red = 0.0;
#pragma unroll
        for (int i = 0; i < RPT; i++) red += row(a + i, acc[i]);
        return red;
    }
    __device__ void finish(int, double) {}
};

// Block reduction of the epilogue shares, then the per-system grid finish.
// An epilogue with kFinishKernel only stores its block partials; the
// per-system finish runs in k_epi_finish, the next launch.  That avoids the
// device-scope fences of the last-block ticket (on gfx950 an L2 write-back
// and invalidate in every block), which the tile gather pays once per tile.
template <class Epi>
constexpr bool finish_kernel()
{
    if constexpr (Epi::kReduce)
        return Epi::kFinishKernel;
    else
        return false;
}

template <int BS, class Epi>
__device__ __forceinline__ void reduce_tail(Epi& epi, double acc, int sys, const GridSys& g)
{
    if constexpr (Epi::kReduce) {
        __shared__ double sh[BS / 64];
        acc = block_sum<BS>(acc, sh);
        if (threadIdx.x == 0) g.partials[2 * blockIdx.x] = acc;
        if constexpr (finish_kernel<Epi>()) return;
        if (arrive_last(&g.counters[sys], g.blk_count[sys])) {
            double s = 0.0;
            for (int i = threadIdx.x; i < g.blk_count[sys]; i += BS) s += g.partials[2 * (g.blk_first[sys] + i)];
            s = block_sum<BS>(s, sh);
            if (threadIdx.x == 0) epi.finish(sys, s);
        }
    }
}

// One block per system: the sum of the system's block partials (the order
// of the last-block ticket's sum) handed to the epilogue's finish.  Systems
// whose gather blocks returned early (begin() false) are skipped the same way.
template <class Epi>
__global__ __launch_bounds__(kRowBlock) void k_epi_finish(Epi epi, GridSys g)
{
    __shared__ double sh[kRowBlock / 64];
    const int sys = blockIdx.x;
    if (!epi.begin(sys)) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < g.blk_count[sys]; i += kRowBlock) s += g.partials[2 * (g.blk_first[sys] + i)];
    s = block_sum<kRowBlock>(s, sh);
    if (threadIdx.x == 0) epi.finish(sys, s);
}

template <class VT, class View>
__device__ __forceinline__ VT row_sum(const View& m, int k, int64_t r, uint32_t cb, const VT* __restrict__ x)
{
    VT acc = VT(0);
#pragma unroll 8
    for (int s = 0; s < k; s++) acc += x[cb + m.col(s, r)];
    return acc;
}

// Generic gather: one row per thread, x read straight from global memory
// (any matrix).
template <class VT, class View, class Epi>
__global__ __launch_bounds__(kRowBlock) void k_gather_rows(View m, int k, const RowWork* __restrict__ work,
                                                           const uint32_t* __restrict__ sys_col,
                                                           const VT* __restrict__ x, Epi epi, GridSys g)
{
    const RowWork w = work[blockIdx.x];
    if (!epi.begin(w.sys)) return;
    const uint32_t cb = sys_col[w.sys];
    double acc = 0.0;
    for (int64_t r = w.row_begin + threadIdx.x; r < w.row_end; r += kRowBlock)
        acc += epi.row(r, row_sum<VT>(m, k, r, cb, x));
    reduce_tail<kRowBlock>(epi, acc, w.sys, g);
}

// RPT consecutive entries of one slot column in one aligned vector load.
template <class IdxT, int RPT>
struct IdxVec;
template <>
struct IdxVec<uint16_t, 1> {
    using T = uint16_t;
    static __device__ __forceinline__ uint32_t get(T v, int) { return v; }
};
template <>
struct IdxVec<uint32_t, 1> {
    using T = uint32_t;
    static __device__ __forceinline__ uint32_t get(T v, int) { return v; }
};
template <>
struct IdxVec<uint16_t, 2> {
    using T = uint32_t;
    static __device__ __forceinline__ uint32_t get(T v, int i) { return (v >> (16 * i)) & 0xFFFFu; }
};
template <>
struct IdxVec<uint16_t, 4> {
    using T = uint2;
    static __device__ __forceinline__ uint32_t get(T v, int i)
    {
        const uint32_t w = i < 2 ? v.x : v.y;
        return (w >> (16 * (i & 1))) & 0xFFFFu;
    }
};
template <>
struct IdxVec<uint16_t, 8> {
    using T = uint4;
    static __device__ __forceinline__ uint32_t get(T v, int i)
    {
        const uint32_t w = i < 2 ? v.x : i < 4 ? v.y : i < 6 ? v.z : v.w;
        return (w >> (16 * (i & 1))) & 0xFFFFu;
    }
};
template <>
struct IdxVec<uint32_t, 2> {
    using T = uint2;
    static __device__ __forceinline__ uint32_t get(T v, int i) { return i == 0 ? v.x : v.y; }
};
template <>
struct IdxVec<uint32_t, 4> {
    using T = uint4;
    static __device__ __forceinline__ uint32_t get(T v, int i)
    {
        return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
    }
};

// One 16-byte line from global memory into LDS: lane i of the wave writes
// lds + 16 i (lds is the same for the whole wave).  Issued as an instruction
// the compiler does not see as an LDS write, so its own waits for LDS reads
// do not wait for this load -- nor for every vector load issued after it
// (vmcnt counts in order); the caller waits for it (s_waitcnt vmcnt(0)) and
// synchronises the block before any thread reads those bytes, and the block
// barriers before and after a staging keep the LDS accesses in order around
// it.  Declared without a memory clobber: one would mark every later load of
// the kernel's read-only tables as possibly overwritten and turn their scalar
// loads into vector loads with full vmcnt waits.  M0 (the LDS address of
// the load) is declared clobbered, so the compiler assumes no value of its
// own survives in it across the asm (clang warns that a reserved register is
// not saved for the caller, which is not wanted here).
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void load_line_to_lds(const void* src, void* lds)
{
    const uint32_t la = __builtin_amdgcn_readfirstlane(
        uint32_t(reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) void*)lds)));
    asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(la), "v"(src) : "m0");
}
#pragma clang diagnostic pop

// Copies the W elements at src into LDS at dst + h, h = the element offset of
// src inside its 16-byte line (dst is 16-byte aligned), and returns h: the
// 16-byte lines wholly inside [src, src + W) go from memory straight into LDS
// (global_load_lds_dwordx4: a wave's 64 lanes fill 1 KB of consecutive LDS,
// no registers held, every load of the thread in flight at once), the ragged
// head and tail through registers.  Reads nothing outside [src, src + W).
// The caller waits for the loads (vmcnt) and synchronises the block before
// the LDS is read.
template <class VT, int BS>
__device__ __forceinline__ uint32_t stage_direct(VT* __restrict__ dst, const VT* __restrict__ src, uint32_t W)
{
    constexpr uint32_t Q = 16 / sizeof(VT);
    const uint32_t h = uint32_t(reinterpret_cast<uintptr_t>(src) / sizeof(VT)) & (Q - 1);
    const uint32_t j0 = min(W, (Q - h) & (Q - 1));  // first element of a whole line
    const uint32_t nq = (W - j0) / Q;               // whole lines
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const VT* g = src + j0;
    VT* l = dst + h + j0;  // 16-byte aligned
    for (uint32_t t0 = wave * 64; t0 < nq; t0 += BS) {
        if (t0 + lane < nq) load_line_to_lds(g + (t0 + lane) * Q, l + t0 * Q);
    }
    if (threadIdx.x < Q) {
        const uint32_t a = threadIdx.x, b = j0 + nq * Q + threadIdx.x;
        if (a < j0) dst[h + a] = src[a];
        if (b < W) dst[h + b] = src[b];
    }
    return h;
}

// LDS-staged gather for pattern matrices.  A block owns a tile of rows and
// walks the plan's steps (runs of at most kMaxSegSlots slots of one segment,
// a segment being a run of slots over one column block) in slot order.  The
// columns of consecutive segments that fit the LDS together are staged in one
// phase, and the random reads of x then hit LDS instead of pulling an L2 line
// per element.  Thread t owns the RPT consecutive rows starting at the tile's
// first row rounded down to RPT plus RPT * t, so one aligned vector load brings
// a slot's entries of all its rows; the next step's entry loads go out once a
// step's staging has landed, so they are in flight during its sums.  Segments wider than the LDS (windows of
// wmax columns) are staged window by window: each slot's value is kept in
// registers until every window was seen, then the segment's slots are added
// in slot order, so every row is the sequential sum in slot order
// (Math/MatrixCSR.cpp:59-62).
// Threads of a tile-gather block: kTileBlock (one block and the whole LDS per
// CU), or half of it, with two blocks of half the LDS per CU whose staging
// and sums overlap each other.
constexpr int kTileBlock = 1024;

// Step k's descriptor in scalar registers (the same for the whole block).
__device__ __forceinline__ GatherStep step_at(const GatherStep* __restrict__ steps, int k)
{
    const uint32_t* __restrict__ w = reinterpret_cast<const uint32_t*>(steps + k);
    uint32_t r[sizeof(GatherStep) / 4];
#pragma unroll
    for (int i = 0; i < int(sizeof(GatherStep) / 4); i++) r[i] = __builtin_amdgcn_readfirstlane(w[i]);
    GatherStep st;
    __builtin_memcpy(&st, r, sizeof(st));
    return st;
}

// A step's entries of this thread's rows (threads past the tile read row 0
// through vec; the sums ignore them).  With Uncond every slot up to
// kMaxSegSlots is loaded (slots past the step's last repeat it): no branch
// around the loads, so the compiler can count them and need not wait for all
// loads in flight before a use; otherwise only the step's slots.
template <class IdxT, int RPT, bool Uncond>
__device__ __forceinline__ void load_step(typename IdxVec<IdxT, RPT>::T (&q)[kMaxSegSlots],
                                          const typename IdxVec<IdxT, RPT>::T* __restrict__ vec, int64_t vld,
                                          const GatherStep& st)
{
    const int last = st.nslots - 1;
#pragma unroll
    for (int j = 0; j < kMaxSegSlots; j++) {
        if (Uncond)
            q[j] = vec[int64_t(st.slot_first + min(j, last)) * vld];
        else if (j <= last)
            q[j] = vec[int64_t(st.slot_first + j) * vld];
    }
}

// element offset of p inside its 16-byte line
template <class VT>
__device__ __forceinline__ uint32_t line_head(const VT* p)
{
    return uint32_t(reinterpret_cast<uintptr_t>(p) / sizeof(VT)) & (16 / sizeof(VT) - 1);
}

// One tile per block (no sub-tiles): the body of k_gather_tile<..., 1>.
template <class VT, int RPT, class IdxT, class Epi, int BS>
__device__ __forceinline__ void gather_tile_single(const SlotMajor<IdxT>& m, const RowWork* __restrict__ work,
                                                   const uint32_t* __restrict__ sys_col,
                                                   const GatherStep* __restrict__ steps, int nsteps,
                                                   const PhaseSeg* __restrict__ phases, uint32_t wmax,
                                                   const VT* __restrict__ x, Epi& epi, const GridSys& g)
{
    using V = IdxVec<IdxT, RPT>;
    using VecT = typename V::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    VT* xs = reinterpret_cast<VT*>(smem_raw);
    const RowWork w = work[blockIdx.x];
    if (!epi.begin(w.sys)) return;
    const VT* __restrict__ xb = x + sys_col[w.sys];
    const int64_t a = (w.row_begin & ~int64_t(RPT - 1)) + int64_t(RPT) * threadIdx.x;
    const bool live = a < w.row_end;
    bool ok[RPT];
    VT acc[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        ok[i] = live && a + i >= w.row_begin && a + i < w.row_end;
        acc[i] = VT(0);
    }
    // slot s of this thread's rows: vec[s * vld]
    const VecT* __restrict__ vec = reinterpret_cast<const VecT*>(m.idx + (live ? a : 0));
    const int64_t vld = m.ld / RPT;
    // The next step's entries are loaded once this step's staging has landed,
    // so they are in flight while this step's sums run.  A segment wider than
    // the LDS holds its slots' values of every row until its last window; with
    // more than 8 bytes of entries per slot the registers do not allow holding
    // the next step's entries too, and the step after it loads its own.
    constexpr bool kPrefetchWide = sizeof(VecT) <= 8;
    constexpr int kWide = wide_step_slots(int(sizeof(VecT)));  // slots of a wide step (the plan's split)
    VecT q[kMaxSegSlots], qn[kMaxSegSlots];
    GatherStep st = nsteps > 0 ? step_at(steps, 0) : GatherStep{};
    bool have = false;  // q holds this step's entries (loaded by the step before)
    for (int k = 0; k < nsteps; k++) {
        const bool more = k + 1 < nsteps;
        const GatherStep nx = step_at(steps, more ? k + 1 : k);
        if (!have) load_step<IdxT, RPT, true>(q, vec, vld, st);
        have = false;
        if (st.nwin > 1) {
            // a segment wider than the LDS, at most kWide slots per step
            VT v[RPT][kWide];
#pragma unroll
            for (int j = 0; j < kWide; j++)
#pragma unroll
                for (int i = 0; i < RPT; i++) v[i][j] = VT(0);
            const bool rev = st.wflags & kWideReverse;
            for (uint32_t n = 0; n < st.nwin; n++) {
                const uint32_t wl = (rev ? st.nwin - 1 - n : n) * wmax;
                const uint32_t W = min(wmax, st.width - wl);
                // the next step's entries load with the last window's staging
                if constexpr (kPrefetchWide)
                    if (more && n + 1 == st.nwin) {
                        load_step<IdxT, RPT, true>(qn, vec, vld, nx);
                        have = true;
                    }
                uint32_t h;
                if (n == 0 && (st.wflags & kWideStaged)) {
                    // staged by the step before (same segment, same window)
                    h = line_head(xb + st.col_lo + wl);
                } else {
                    __syncthreads();
                    h = stage_direct<VT, BS>(xs, xb + st.col_lo + wl, W);
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                }
#pragma unroll
                for (int j = 0; j < kWide; j++) {
                    if (j < st.nslots) {
                        // 32-bit offset: base - col_lo wraps for u32 storage
                        // (base 0) and the sum with the entry unwraps it
                        const uint32_t off = st.off[j] - wl;
#pragma unroll
                        for (int i = 0; i < RPT; i++) {
                            const uint32_t c = off + V::get(q[j], i);
                            const bool in = ok[i] && c < W;
                            const VT t = xs[h + (in ? c : 0u)];
                            v[i][j] = in ? t : v[i][j];
                        }
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kWide; j++)
                if (j < st.nslots) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) acc[i] += v[i][j];
                }
            if constexpr (kPrefetchWide)
                if (have) {
#pragma unroll
                    for (int j = 0; j < kMaxSegSlots; j++) q[j] = qn[j];
                }
        } else {
            if (st.phase_count > 0) {
                __syncthreads();
                for (int p = st.phase_first; p < st.phase_first + st.phase_count; p++) {
                    const PhaseSeg ps = phases[p];
                    stage_direct<VT, BS>(xs + ps.lds_off, xb + ps.col_lo, ps.width);
                }
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            } else {
                // this step's entries must be in before the next step's loads
                // go out: the compiler's own wait for them would be a vmcnt(0)
                // after those
                __builtin_amdgcn_s_waitcnt(0);
            }
            if (more) {
                load_step<IdxT, RPT, true>(qn, vec, vld, nx);
                have = true;
            }
            const uint32_t h = line_head(xb + st.col_lo);
#pragma unroll
            for (int j = 0; j < kMaxSegSlots; j++) {
                if (j < st.nslots) {
                    const uint32_t off = st.off[j] + h;
#pragma unroll
                    for (int i = 0; i < RPT; i++)
                    {
                        const VT t = xs[ok[i] ? off + V::get(q[j], i) : 0u];
                        acc[i] += ok[i] ? t : VT(0);
                    }
                }
            }
            // (the copy in each branch: qn is not live across a wide step
            // that does not load it)
            if (have) {
#pragma unroll
                for (int j = 0; j < kMaxSegSlots; j++) q[j] = qn[j];
            }
        }
        st = nx;
    }
    // all RPT rows valid (every thread but the tile's first and last): the
    // epilogue's vector form; else row by row
    bool all = true;
#pragma unroll
    for (int i = 0; i < RPT; i++) all = all && ok[i];
    double red = 0.0;
    if (all) {
        red = epi.template rows<RPT>(a, acc);
    } else {
#pragma unroll
        for (int i = 0; i < RPT; i++)
            if (ok[i]) red += epi.row(a + i, acc[i]);
    }
    reduce_tail<BS>(epi, red, w.sys, g);
}

// S sub-tiles per block: the body of k_gather_tile<..., S > 1>.
template <class VT, int RPT, class IdxT, class Epi, int BS, int S>
__device__ __forceinline__ void gather_tile_sub(const SlotMajor<IdxT>& m, const RowWork* __restrict__ work,
                                                const uint32_t* __restrict__ sys_col,
                                                const GatherStep* __restrict__ steps, int nsteps,
                                                const PhaseSeg* __restrict__ phases, uint32_t wmax,
                                                const VT* __restrict__ x, Epi& epi, const GridSys& g)
{
    using V = IdxVec<IdxT, RPT>;
    using VecT = typename V::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    VT* xs = reinterpret_cast<VT*>(smem_raw);
    const RowWork w = work[blockIdx.x];
    if (!epi.begin(w.sys)) return;
    const VT* __restrict__ xb = x + sys_col[w.sys];
    // S sub-tiles of BS * RPT rows: in sub-tile u the thread owns rows
    // a0 + u * kSub + i, i < RPT; every staging serves all of them
    constexpr int64_t kSub = int64_t(RPT) * BS;
    const int64_t a0 = (w.row_begin & ~int64_t(RPT - 1)) + int64_t(RPT) * threadIdx.x;
    bool okr[S][RPT];  // that row is in the tile
    VT acc[S][RPT];
#pragma unroll
    for (int u = 0; u < S; u++)
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            const int64_t r = a0 + u * kSub + i;
            okr[u][i] = r >= w.row_begin && r < w.row_end;
            acc[u][i] = VT(0);
        }
    auto ok = [&](int u, int i) { return okr[u][i]; };
    // slot s of sub-tile u's rows: vec_of(u)[s * vld] (rows past the tile read row 0)
    const int64_t vld = m.ld / RPT;
    auto vec_of = [&](int u) {
        const int64_t au = a0 + u * kSub;
        return reinterpret_cast<const VecT*>(m.idx + (au < w.row_end ? au : 0));
    };
    // Where the registers allow it (up to 8 bytes of entries per slot) the
    // entries of the next (step, sub-tile) are loaded while this one sums;
    // otherwise each sub-tile loads its own (the first ahead of the staging).
    constexpr bool kPrefetch = sizeof(VecT) <= 8;
    VecT q[kMaxSegSlots], qn[kPrefetch ? kMaxSegSlots : 1];
    GatherStep st = nsteps > 0 ? step_at(steps, 0) : GatherStep{};
    if (kPrefetch && nsteps > 0) load_step<IdxT, RPT, true>(q, vec_of(0), vld, st);
    for (int k = 0; k < nsteps; k++) {
        const bool more = k + 1 < nsteps;
        const GatherStep nx = step_at(steps, more ? k + 1 : k);
        // the pair after (k, u) is (k, u + 1), or (k + 1, 0)
        auto prefetch_after = [&](int u) {
            if constexpr (kPrefetch) {
                if (u + 1 < S)
                    load_step<IdxT, RPT, true>(qn, vec_of(u + 1), vld, st);
                else if (more)
                    load_step<IdxT, RPT, true>(qn, vec_of(0), vld, nx);
            }
        };
        auto advance = [&](int u) {
            if constexpr (kPrefetch)
                if (u + 1 < S || more) {
#pragma unroll
                    for (int j = 0; j < kMaxSegSlots; j++) q[j] = qn[j];
                }
        };
        if (st.nwin > 1) {
            // a segment wider than the LDS, at most kMaxSegSlots slots: each
            // sub-tile walks its windows
#pragma unroll
            for (int u = 0; u < S; u++) {
                if constexpr (!kPrefetch) load_step<IdxT, RPT, false>(q, vec_of(u), vld, st);
                VT v[RPT][kMaxSegSlots];
#pragma unroll
                for (int j = 0; j < kMaxSegSlots; j++)
#pragma unroll
                    for (int i = 0; i < RPT; i++) v[i][j] = VT(0);
                for (uint32_t wl = 0; wl < st.width; wl += wmax) {
                    const uint32_t W = min(wmax, st.width - wl);
                    // the next entries load with the last window's staging
                    if (wl + W >= st.width) prefetch_after(u);
                    __syncthreads();
                    const uint32_t h = stage_direct<VT, BS>(xs, xb + st.col_lo + wl, W);
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
#pragma unroll
                    for (int j = 0; j < kMaxSegSlots; j++) {
                        if (j < st.nslots) {
                            // 32-bit offset: base - col_lo wraps for u32 storage
                            // (base 0) and the sum with the entry unwraps it
                            const uint32_t off = st.off[j] - wl;
#pragma unroll
                            for (int i = 0; i < RPT; i++) {
                                const uint32_t c = off + V::get(q[j], i);
                                const bool in = ok(u, i) && c < W;
                                const VT t = xs[h + (in ? c : 0u)];
                                v[i][j] = in ? t : v[i][j];
                            }
                        }
                    }
                }
#pragma unroll
                for (int j = 0; j < kMaxSegSlots; j++)
                    if (j < st.nslots) {
#pragma unroll
                        for (int i = 0; i < RPT; i++) acc[u][i] += v[i][j];
                    }
                advance(u);
            }
        } else {
            if constexpr (!kPrefetch) load_step<IdxT, RPT, false>(q, vec_of(0), vld, st);
            if (st.phase_count > 0) {
                __syncthreads();
                for (int p = st.phase_first; p < st.phase_first + st.phase_count; p++) {
                    const PhaseSeg ps = phases[p];
                    stage_direct<VT, BS>(xs + ps.lds_off, xb + ps.col_lo, ps.width);
                }
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
            const uint32_t h = line_head(xb + st.col_lo);
#pragma unroll
            for (int u = 0; u < S; u++) {
                if constexpr (!kPrefetch) {
                    if (u > 0) load_step<IdxT, RPT, false>(q, vec_of(u), vld, st);
                } else if (u > 0 || st.phase_count == 0) {
                    // this pair's entries (loaded during the last pair's sums)
                    // must be in before the next pair's loads go out: the
                    // compiler's own wait for them would be a vmcnt(0) after
                    __builtin_amdgcn_s_waitcnt(0);
                }
                prefetch_after(u);
#pragma unroll
                for (int j = 0; j < kMaxSegSlots; j++) {
                    if (j < st.nslots) {
                        const uint32_t off = st.off[j] + h;
#pragma unroll
                        for (int i = 0; i < RPT; i++) {
                            const VT t = xs[ok(u, i) ? off + V::get(q[j], i) : 0u];
                            acc[u][i] += ok(u, i) ? t : VT(0);
                        }
                    }
                }
                advance(u);
            }
        }
        st = nx;
    }
    // a sub-tile's RPT rows all in the tile (every thread but the tile's first
    // and last): the epilogue's vector form; else row by row
    double red = 0.0;
#pragma unroll
    for (int u = 0; u < S; u++) {
        const int64_t au = a0 + u * kSub;
        bool all = true;
#pragma unroll
        for (int i = 0; i < RPT; i++) all = all && okr[u][i];
        if (all) {
            red += epi.template rows<RPT>(au, acc[u]);
        } else {
#pragma unroll
            for (int i = 0; i < RPT; i++)
                if (ok(u, i)) red += epi.row(au + i, acc[u][i]);
        }
    }
    reduce_tail<BS>(epi, red, w.sys, g);
}

template <class VT, int RPT, class IdxT, class Epi, int BS, int S>
__global__ __launch_bounds__(BS) void k_gather_tile(SlotMajor<IdxT> m, const RowWork* __restrict__ work,
                                                    const uint32_t* __restrict__ sys_col,
                                                    const GatherStep* __restrict__ steps, int nsteps,
                                                    const PhaseSeg* __restrict__ phases, uint32_t wmax,
                                                    const VT* __restrict__ x, Epi epi, GridSys g)
{
    if constexpr (S == 1)
        gather_tile_single<VT, RPT, IdxT, Epi, BS>(m, work, sys_col, steps, nsteps, phases, wmax, x, epi, g);
    else
        gather_tile_sub<VT, RPT, IdxT, Epi, BS, S>(m, work, sys_col, steps, nsteps, phases, wmax, x, epi, g);
}

// The LDS-staged gather of a matrix-free pattern matrix: the stored tile
// kernel's plan (steps over the slots, narrow segments staged side by side in
// one phase, wide segments window by window), with each thread's RPT rows'
// ternary words in registers and every slot's column computed by its program
// where the stored kernel loads it.  A wide step computes its slots' columns
// once, before the window loop.
template <class VT, int RPT, class Epi>
__global__ __launch_bounds__(kTileBlock) void k_gather_tile_mf(PositionsView m, const RowWork* __restrict__ work,
                                                               const uint32_t* __restrict__ sys_col,
                                                               const GatherStep* __restrict__ steps, int nsteps,
                                                               const PhaseSeg* __restrict__ phases, uint32_t wmax,
                                                               const VT* __restrict__ x, Epi epi, GridSys g)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    VT* xs = reinterpret_cast<VT*>(smem_raw);
    const RowWork w = work[blockIdx.x];
    if (!epi.begin(w.sys)) return;
    const VT* __restrict__ xb = x + sys_col[w.sys];
    const int64_t r0 = w.row_begin + threadIdx.x;
    const int nval = int(max<int64_t>(0, min<int64_t>(w.row_end - r0, int64_t(RPT) * kTileBlock)));
    uint32_t T[RPT][4];
    VT acc[RPT];
#pragma unroll
    for (int i = 0; i < RPT; i++) {
        acc[i] = VT(0);
        uint4 t = make_uint4(0u, 0u, 0u, 0u);
        if (i * kTileBlock < nval) t = m.tw[r0 + int64_t(i) * kTileBlock];
        T[i][0] = t.x;
        T[i][1] = t.y;
        T[i][2] = t.z;
        T[i][3] = t.w;
    }
    for (int k = 0; k < nsteps; k++) {
        const GatherStep st = step_at(steps, k);
        if (st.nwin > 1) {
            // a segment wider than the LDS: the step's columns (segment-local),
            // then its windows, each slot's value kept until the last window
            uint32_t c[RPT][kMaxSegSlots];
            VT v[RPT][kMaxSegSlots];
#pragma unroll
            for (int j = 0; j < kMaxSegSlots; j++) {
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    v[i][j] = VT(0);
                    c[i][j] = 0xFFFFFFFFu;
                }
                if (j < st.nslots) {
                    const DevSlot sl = uniform_load(m.slots + st.slot_first + j);
                    uint32_t ci[RPT];
                    slot_index_rows<RPT>(sl, uniform_load(m.tslots + st.slot_first + j), m.tern, m.dense, T, ci);
#pragma unroll
                    for (int i = 0; i < RPT; i++)
                        if (i * kTileBlock < nval) c[i][j] = st.off[j] + sl.col_offset + ci[i];
                }
            }
            const bool rev = st.wflags & kWideReverse;
            for (uint32_t n = 0; n < st.nwin; n++) {
                const uint32_t wl = (rev ? st.nwin - 1 - n : n) * wmax;
                const uint32_t W = min(wmax, st.width - wl);
                uint32_t h;
                if (n == 0 && (st.wflags & kWideStaged)) {
                    h = line_head(xb + st.col_lo + wl);  // staged by the step before
                } else {
                    __syncthreads();
                    h = stage_direct<VT, kTileBlock>(xs, xb + st.col_lo + wl, W);
                    __builtin_amdgcn_s_waitcnt(0);
                    __syncthreads();
                }
#pragma unroll
                for (int j = 0; j < kMaxSegSlots; j++) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) {
                        const uint32_t cl = c[i][j] - wl;
                        if (cl < W) v[i][j] = xs[h + cl];
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kMaxSegSlots; j++)
                if (j < st.nslots) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) acc[i] += v[i][j];
                }
        } else {
            if (st.phase_count > 0) {
                __syncthreads();
                for (int p = st.phase_first; p < st.phase_first + st.phase_count; p++) {
                    const PhaseSeg ps = uniform_load(phases + p);
                    stage_direct<VT, kTileBlock>(xs + ps.lds_off, xb + ps.col_lo, ps.width);
                }
                __builtin_amdgcn_s_waitcnt(0);
                __syncthreads();
            }
            const uint32_t h = line_head(xb + st.col_lo);
            for (int j = 0; j < st.nslots; j++) {
                const DevSlot sl = uniform_load(m.slots + st.slot_first + j);
                // (read from the descriptor in memory: a run-time index into
                // the copy in registers would put the step in scratch)
                const uint32_t off = __builtin_amdgcn_readfirstlane(steps[k].off[j]) + h + sl.col_offset;
                uint32_t ci[RPT];
                slot_index_rows<RPT>(sl, uniform_load(m.tslots + st.slot_first + j), m.tern, m.dense, T, ci);
#pragma unroll
                for (int i = 0; i < RPT; i++)
                    if (i * kTileBlock < nval) acc[i] += xs[off + ci[i]];
            }
        }
    }
    double red = 0.0;
#pragma unroll
    for (int i = 0; i < RPT; i++)
        if (i * kTileBlock < nval) red += epi.row(r0 + int64_t(i) * kTileBlock, acc[i]);
    reduce_tail<kTileBlock>(epi, red, w.sys, g);
}

// Rows init of LSQR: u_raw = src (converted), beta = |src| (FP64).
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_lsqr_init_u(const RowWork* __restrict__ work,
                                                           const double* __restrict__ src, VT* __restrict__ u,
                                                           GridSys g, LsqrScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const RowWork w = work[blockIdx.x];
    LsqrScalars* st = &scal[w.sys];
    if (!st->active) return;
    double acc = 0.0;
    for (int64_t r = w.row_begin + threadIdx.x; r < w.row_end; r += kRowBlock) {
        const double v = src[r];
        u[r] = VT(v);
        acc += v * v;
    }
    acc = block_sum<kRowBlock>(acc, sh);
    if (threadIdx.x == 0) g.partials[2 * blockIdx.x] = acc;
    if (arrive_last(&g.counters[w.sys], g.blk_count[w.sys])) {
        double s = 0.0;
        for (int i = threadIdx.x; i < g.blk_count[w.sys]; i += kRowBlock) s += g.partials[2 * (g.blk_first[w.sys] + i)];
        s = block_sum<kRowBlock>(s, sh);
        if (threadIdx.x == 0) {
            st->beta = sqrt(s);
            st->anorm2 = 0.0;
        }
    }
}

// ------------------------------------------------------------ scatter ------
// Which systems a scatter serves, and its scale:
//   kScatterAll      every system, scale 1 (plain products)
//   kScatterLsqr     the systems iterating (active, not stopped), y scaled by 1 / beta
//   kScatterRestart  the systems being (re)started (active, restart), scale 1
constexpr int kScatterAll = 0, kScatterLsqr = 1, kScatterRestart = 2;

__device__ __forceinline__ bool scatter_selected(const LsqrScalars* div, int mode, int sys)
{
    if (mode == kScatterAll) return true;
    const LsqrScalars& d = div[sys];
    if (!d.active) return false;
    return mode == kScatterLsqr ? d.stop == 0 : d.restart != 0;
}

template <class VT>
__device__ __forceinline__ VT scatter_scale(const LsqrScalars* div, int mode, int sys)
{
    if (mode != kScatterLsqr) return VT(1);
    const double d = div[sys].beta;
    return VT(d != 0.0 ? 1.0 / d : 0.0);
}

// The same, where an LSQR iteration hands over its gather's |u_raw|^2 block
// partials (bsum.partials non-null): beta = sqrt of their sum is formed here,
// by every block from the same partials in the same order (so all blocks
// agree), and stored with its square for the kernels after; the scatter then
// reads no beta from the state.  Only the kBeta instantiations of the scatter
// kernels carry this code (the others compile as if it did not exist).
template <class VT, int BS, bool kBeta>
__device__ __forceinline__ VT scatter_scale(LsqrScalars* div, int mode, int sys, const GridSys& bsum)
{
    if constexpr (!kBeta) return scatter_scale<VT>(div, mode, sys);
    if (mode != kScatterLsqr || !bsum.partials) return scatter_scale<VT>(div, mode, sys);
    __shared__ double sh[BS / 64];
    const double s = system_sum<BS>(bsum, sys, sh);
    const double d = sqrt(s);
    if (threadIdx.x == 0) {
        div[sys].beta = d;
        div[sys].beta2 = s;
    }
    return VT(d != 0.0 ? 1.0 / d : 0.0);
}

// Integer accumulation of an FP32 scatter window.  Each term y * scale is
// rounded to a multiple of 2^-s and the window adds the int64 multiples with
// LDS integer atomics (ds_add_u64), which run about 1.8 times as fast as the
// FP64 ones on gfx950 and make the sums independent of the atomics' order
// (the FP32 A^T y is bit-reproducible).  s is chosen so that no column sum
// can overflow: |term| < 2^ey and fewer than 2^cbits terms per column and
// work item, so every sum stays below 2^(62 - 1); each term keeps at least 32
// bits (rounding error <= 2^(ey - 33) of the largest term, far below FP32
// rounding of the sums).  ymax holds, per system, the bits of max |y| (a
// non-negative float, NaN and inf compare above every finite value); null in
// an LSQR iteration, where |u_r / beta| < 2 by construction.
// The kYmax scatter instantiation forms each system's max |y| itself from
// the per-row-block maxima of k_absmax_blocks (ypart, the system's blocks at
// yfirst[sys], ycount[sys]) instead of reading ymax.
struct FixedScale {
    const unsigned* ymax;
    int cbits;
    const unsigned* ypart;
    const int32_t* yfirst;
    const int32_t* ycount;
};

// 2^s for a system whose max |y| has float bits b (bounded: b is given; else
// |y * sc| < 2), or 0 where the FP64 window must be used (non-finite values,
// or fewer than 32 bits left for the largest term).
__device__ __forceinline__ double fixed_scale_bits(int cbits, bool bounded, unsigned b, double sc)
{
    int ey;
    if (bounded) {
        if (b >= 0x7f800000u) return 0.0;
        ey = b == 0 ? -126 : ilogbf(__uint_as_float(b)) + 1;
    } else {
        if (!isfinite(sc)) return 0.0;
        ey = 1;
    }
    const int sh = 62 - cbits - ey;
    if (sh + ey < 32) return 0.0;
    return ldexp(1.0, sh);
}

// The max of a system's row-block maxima (float bits), in every thread.
template <int BS>
__device__ __forceinline__ unsigned system_max(const FixedScale& fx, int sys, unsigned* sh)
{
    unsigned mb = 0;
    const int first = fx.yfirst[sys], n = fx.ycount[sys];
    for (int i = threadIdx.x; i < n; i += BS) mb = max(mb, fx.ypart[first + i]);
    for (int o = 32; o > 0; o >>= 1) mb = max(mb, unsigned(__shfl_xor(int(mb), o, 64)));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < BS / 64; i++) mb = max(mb, sh[i]);
        sh[0] = mb;
    }
    __syncthreads();
    return sh[0];
}

// The same for system sys, with its bits read from fx.ymax where given.
__device__ __forceinline__ double fixed_scale(const FixedScale& fx, int sys, double sc)
{
    return fixed_scale_bits(fx.cbits, fx.ymax != nullptr, fx.ymax ? fx.ymax[sys] : 0u, sc);
}

template <class WT>
__device__ __forceinline__ WT window_term(double v, double fscale)
{
    if constexpr (std::is_same_v<WT, double>)
        return v;
    else
        return WT(__double2ll_rn(v * fscale));
}

// max |VT(y)| as float bits (non-negative floats order as unsigned integers;
// NaN and inf above every finite value), the ymax of FixedScale, in two
// launches without atomics: one value per row block, then one per system.
template <class YT>
__global__ __launch_bounds__(kRowBlock) void k_absmax_blocks(const RowWork* __restrict__ work,
                                                             const YT* __restrict__ y, unsigned* __restrict__ part)
{
    __shared__ unsigned sh[kRowBlock / 64];
    const RowWork w = work[blockIdx.x];
    unsigned mb = 0;
    for (int64_t r = w.row_begin + threadIdx.x; r < w.row_end; r += kRowBlock)
        mb = max(mb, __float_as_uint(fabsf(float(y[r]))));
    for (int o = 32; o > 0; o >>= 1) mb = max(mb, unsigned(__shfl_xor(int(mb), o, 64)));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kRowBlock / 64; i++) mb = max(mb, sh[i]);
        part[blockIdx.x] = mb;
    }
}

__global__ __launch_bounds__(kRowBlock) void k_absmax_systems(const int32_t* __restrict__ first,
                                                              const int32_t* __restrict__ count,
                                                              const unsigned* __restrict__ part,
                                                              unsigned* __restrict__ bits)
{
    __shared__ unsigned sh[kRowBlock / 64];
    const int s = blockIdx.x;
    unsigned mb = 0;
    for (int i = threadIdx.x; i < count[s]; i += kRowBlock) mb = max(mb, part[first[s] + i]);
    for (int o = 32; o > 0; o >>= 1) mb = max(mb, unsigned(__shfl_xor(int(mb), o, 64)));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = mb;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kRowBlock / 64; i++) mb = max(mb, sh[i]);
        bits[s] = mb;
    }
}

// A^T (y * scale), stage 1: each block owns a column window of one slot group
// in LDS, scatters a chunk of rows into it and stores the window to its slice
// of the partial buffer -- no global atomics.  This generic form (the
// evaluator's row-major indices) accumulates in FP64 whatever VT is: LDS
// atomic adds on doubles run several times faster than on floats on this GPU,
// and the sums are exact for integer data either way.
// scale = 1 / div[sys].beta when div is given (0 when beta == 0), else 1.
template <class VT, class YT, class View, bool kBeta = false>
__global__ __launch_bounds__(kScatterBlock) void k_scatter_part(View m, const ScatterWork* __restrict__ work,
                                                                const SlotGroup* __restrict__ groups,
                                                                const uint16_t* __restrict__ slot_list,
                                                                const uint32_t* __restrict__ sys_col,
                                                                const YT* __restrict__ y, LsqrScalars* div,
                                                                int mode, FixedScale fx, int y_padded,
                                                                VT* __restrict__ part, GridSys bsum)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* win = reinterpret_cast<double*>(smem_raw);
    const ScatterWork w = work[blockIdx.x];
    if (!scatter_selected(div, mode, w.sys)) return;
    const uint32_t W = w.win_hi - w.win_lo;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) win[i] = 0.0;
    __syncthreads();
    const SlotGroup gr = groups[w.group];
    const uint32_t cb = sys_col[w.sys];
    const VT sc = scatter_scale<VT, kScatterBlock, kBeta>(div, mode, w.sys, bsum);
    for (int64_t r = w.row_begin + threadIdx.x; r < w.row_end; r += kScatterBlock) {
        const double yv = double(VT(y[r]) * sc);
        for (int j = 0; j < gr.nslots; j++) {
            const int s = slot_list[gr.first + j];
            const uint32_t uc = cb + m.col(s, r) - w.win_lo;
            if (uc < W) atomicAdd(&win[uc], yv);
        }
    }
    __syncthreads();
    VT* __restrict__ out = part + w.part_off;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) out[i] = VT(win[i]);
}

// The rows of one scatter work item for a group of at most NS slots (slot j
// of the group at vec + slot[j], its window offset off[j]; slots ns..NS-1
// repeat slot 0's loads and add nothing).  Every load is unconditional (rows
// clamped into the chunk), and the next step's y and entries are loaded
// before this step's atomics, so their latency overlaps the LDS work.
template <int NS, class WT, bool VY, class VT, class YT, class IdxT, int RPT>
__device__ __forceinline__ void scatter_rows(const SlotMajor<IdxT>& m, const ScatterWork& w, int ns,
                                             const int64_t (&slot)[kMaxSegSlots],
                                             const uint32_t (&off)[kMaxSegSlots], const YT* __restrict__ y,
                                             VT sc, double fscale, uint32_t W, bool full, WT* win)
{
    using V = IdxVec<IdxT, RPT>;
    using VecT = typename V::T;
    if (w.row_end <= w.row_begin) return;
    const int64_t last = w.row_end - 1, alast = last & ~int64_t(RPT - 1);
    const int64_t stride = int64_t(RPT) * kScatterBlock;
    auto fetch = [&](int64_t at, VecT (&q)[NS], YT (&yy)[RPT]) {
        const VecT* __restrict__ vec = reinterpret_cast<const VecT*>(m.idx + min(at, alast));
#pragma unroll
        for (int j = 0; j < NS; j++) q[j] = vec[slot[j]];
        if constexpr (VY) {
            // y readable through the padding to a multiple of 64 rows: the
            // RPT values in 16-byte loads (rows past the chunk are masked)
            constexpr int kPer = 16 / int(sizeof(YT));
            using Y16 = std::conditional_t<sizeof(YT) == 4, float4, double2>;
            const Y16* __restrict__ yv = reinterpret_cast<const Y16*>(y + min(at, alast));
#pragma unroll
            for (int i = 0; i < RPT; i += kPer) {
                const Y16 t = yv[i / kPer];
                __builtin_memcpy(&yy[i], &t, 16);
            }
        } else {
#pragma unroll
            for (int i = 0; i < RPT; i++) yy[i] = y[min(at + i, last)];
        }
    };
    int64_t a = (w.row_begin & ~int64_t(RPT - 1)) + int64_t(RPT) * threadIdx.x;
    VecT q[NS];
    YT yr[RPT];
    fetch(a, q, yr);
    for (; a < w.row_end; a += stride) {
        VecT qn[NS];
        YT yn[RPT];
        fetch(a + stride, qn, yn);
        WT yv[RPT];
        bool ok[RPT];
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            ok[i] = a + i >= w.row_begin && a + i < w.row_end;
            yv[i] = window_term<WT>(double(VT(yr[i]) * sc), fscale);
        }
        // A wave whose rows all lie inside the chunk, over a group whose
        // columns all lie in the window, adds without per-lane conditions
        // (entries past the window, which valid data never holds, are held
        // inside it rather than skipped)
        if (__all(full && a >= w.row_begin && a + RPT <= w.row_end)) {
#pragma unroll
            for (int j = 0; j < NS; j++) {
                if (j < ns) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) atomicAdd(&win[min(off[j] + V::get(q[j], i), W - 1)], yv[i]);
                }
            }
        } else {
#pragma unroll
            for (int j = 0; j < NS; j++) {
                if (j < ns) {
#pragma unroll
                    for (int i = 0; i < RPT; i++) {
                        const uint32_t uc = off[j] + V::get(q[j], i);
                        if (ok[i] && uc < W) atomicAdd(&win[uc], yv[i]);
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < NS; j++) q[j] = qn[j];
#pragma unroll
        for (int i = 0; i < RPT; i++) yr[i] = yn[i];
    }
}

// Stage 1 for stored indices: as k_scatter_part, with each thread taking RPT
// consecutive rows per step (counted from the chunk's first row rounded down
// to RPT) so one aligned vector load brings a slot's entries of all of them.
// For groups of up to kMaxSegSlots slots the slot numbers and window offsets
// are hoisted out of the row loop and the loads are software-pipelined
// (scatter_rows).
template <class VT, class YT, class IdxT, int RPT, bool kBeta = false, bool kYmax = false>
__global__ __launch_bounds__(kScatterBlock) void k_scatter_tile(SlotMajor<IdxT> m,
                                                                const ScatterWork* __restrict__ work,
                                                                const SlotGroup* __restrict__ groups,
                                                                const uint16_t* __restrict__ slot_list,
                                                                const uint32_t* __restrict__ sys_col,
                                                                const YT* __restrict__ y, LsqrScalars* div,
                                                                int mode, FixedScale fx, int y_padded,
                                                                VT* __restrict__ part, GridSys bsum)
{
    using V = IdxVec<IdxT, RPT>;
    using VecT = typename V::T;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* win = reinterpret_cast<double*>(smem_raw);  // 8-byte window words (FP64 or fixed point)
    const ScatterWork w = work[blockIdx.x];
    if (!scatter_selected(div, mode, w.sys)) return;
    const uint32_t W = w.win_hi - w.win_lo;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) win[i] = 0.0;
    __syncthreads();
    const SlotGroup gr = groups[w.group];
    const int ns = gr.nslots;
    const uint32_t cb = sys_col[w.sys];
    const VT sc = scatter_scale<VT, kScatterBlock, kBeta>(div, mode, w.sys, bsum);
    const int64_t vld = m.ld / RPT;
    // FP32 windows accumulate in fixed point where the scale allows it
    double fscale = 0.0;
    if constexpr (kYmax) {
        __shared__ unsigned shm[kScatterBlock / 64];
        const unsigned b = system_max<kScatterBlock>(fx, w.sys, shm);
        if (std::is_same_v<VT, float> && ns <= kMaxSegSlots) fscale = fixed_scale_bits(fx.cbits, true, b, double(sc));
    } else {
        fscale = std::is_same_v<VT, float> && ns <= kMaxSegSlots ? fixed_scale(fx, w.sys, double(sc)) : 0.0;
    }
    if (ns <= kMaxSegSlots) {
        int64_t slot[kMaxSegSlots];
        uint32_t off[kMaxSegSlots];
        const int s0 = slot_list[gr.first];
#pragma unroll
        for (int j = 0; j < kMaxSegSlots; j++) {
            const int sj = j < ns ? int(slot_list[gr.first + j]) : s0;
            slot[j] = int64_t(sj) * vld;
            off[j] = cb + m.base[sj] - w.win_lo;
        }
        // every column of the group inside this window
        const bool full = w.win_lo == cb + gr.col_lo && gr.col_hi - gr.col_lo <= W;
        auto rows_y = [&](auto* wn, auto vy) {
            using WT = std::remove_pointer_t<decltype(wn)>;
            constexpr bool VY = decltype(vy)::value;
            if (ns == 4)
                scatter_rows<4, WT, VY, VT, YT, IdxT, RPT>(m, w, ns, slot, off, y, sc, fscale, W, full, wn);
            else if (ns == 8)
                scatter_rows<8, WT, VY, VT, YT, IdxT, RPT>(m, w, ns, slot, off, y, sc, fscale, W, full, wn);
            else if (ns <= 2)
                scatter_rows<2, WT, VY, VT, YT, IdxT, RPT>(m, w, ns, slot, off, y, sc, fscale, W, full, wn);
            else if (ns <= 4)
                scatter_rows<4, WT, VY, VT, YT, IdxT, RPT>(m, w, ns, slot, off, y, sc, fscale, W, full, wn);
            else
                scatter_rows<8, WT, VY, VT, YT, IdxT, RPT>(m, w, ns, slot, off, y, sc, fscale, W, full, wn);
        };
        // 16-byte y loads need RPT values to fill them and y padded
        constexpr bool kVecY = RPT * sizeof(YT) % 16 == 0;
        auto rows = [&](auto* wn) {
            if constexpr (kVecY) {
                if (y_padded) {
                    rows_y(wn, std::true_type{});
                    return;
                }
            }
            rows_y(wn, std::false_type{});
        };
        if constexpr (std::is_same_v<VT, float>) {
            if (fscale > 0.0)
                rows(reinterpret_cast<unsigned long long*>(win));
            else
                rows(win);
        } else {
            rows(win);
        }
    } else {
        for (int64_t a = (w.row_begin & ~int64_t(RPT - 1)) + int64_t(RPT) * threadIdx.x; a < w.row_end;
             a += int64_t(RPT) * kScatterBlock) {
            bool ok[RPT];
            double yv[RPT];
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                ok[i] = a + i >= w.row_begin && a + i < w.row_end;
                const double t = double(VT(y[min(a + i, w.row_end - 1)]) * sc);
                yv[i] = ok[i] ? t : 0.0;
            }
            const VecT* __restrict__ vec = reinterpret_cast<const VecT*>(m.idx + a);
            for (int j = 0; j < ns; j++) {
                const int sj = slot_list[gr.first + j];
                const VecT q = vec[int64_t(sj) * vld];
                const uint32_t o = cb + m.base[sj] - w.win_lo;
#pragma unroll
                for (int i = 0; i < RPT; i++) {
                    const uint32_t uc = o + V::get(q, i);
                    if (ok[i] && uc < W) atomicAdd(&win[uc], yv[i]);
                }
            }
        }
    }
    __syncthreads();
    VT* __restrict__ out = part + w.part_off;
    if (fscale > 0.0) {
        const unsigned long long* wu = reinterpret_cast<const unsigned long long*>(win);
        const double inv = 1.0 / fscale;  // a power of two: exact
        for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) out[i] = VT(double((long long)wu[i]) * inv);
    } else {
        for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) out[i] = VT(win[i]);
    }
}

// Stage 1 for a matrix-free matrix: as k_scatter_tile, with each thread's RPT
// positions loaded once per step and every slot's column computed by its
// program (the DevSlot is the same for the whole block).
template <class VT, class YT, int RPT, bool kBeta = false>
__global__ __launch_bounds__(kScatterBlock) void k_scatter_tile_mf(PositionsView m,
                                                                   const ScatterWork* __restrict__ work,
                                                                   const SlotGroup* __restrict__ groups,
                                                                   const uint16_t* __restrict__ slot_list,
                                                                   const uint32_t* __restrict__ sys_col,
                                                                   const YT* __restrict__ y, LsqrScalars* div,
                                                                   int mode, FixedScale fx, int y_padded,
                                                                   VT* __restrict__ part, GridSys bsum)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* win = reinterpret_cast<double*>(smem_raw);
    const ScatterWork w = uniform_load(work + blockIdx.x);
    if (!scatter_selected(div, mode, w.sys)) return;
    const uint32_t W = w.win_hi - w.win_lo;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) win[i] = 0.0;
    __syncthreads();
    const SlotGroup gr = uniform_load(groups + w.group);
    const uint32_t cb = sys_col[w.sys];
    const VT sc = scatter_scale<VT, kScatterBlock, kBeta>(div, mode, w.sys, bsum);
    const int64_t last = w.row_end - 1;
    for (int64_t a = (w.row_begin & ~int64_t(RPT - 1)) + int64_t(RPT) * threadIdx.x; a < w.row_end;
         a += int64_t(RPT) * kScatterBlock) {
        bool ok[RPT];
        double yv[RPT];
        uint32_t T[RPT][4];
#pragma unroll
        for (int i = 0; i < RPT; i++) {
            ok[i] = a + i >= w.row_begin && a + i < w.row_end;
            const int64_t r = min(a + i, last);  // clamped: no branch around the loads
            const double t = double(VT(y[r]) * sc);
            yv[i] = ok[i] ? t : 0.0;
            const uint4 tw = m.tw[r];
            T[i][0] = tw.x;
            T[i][1] = tw.y;
            T[i][2] = tw.z;
            T[i][3] = tw.w;
        }
        for (int j = 0; j < gr.nslots; j++) {
            const int sj = uniform_u16(slot_list, gr.first + j);
            const DevSlot sl = uniform_load(m.slots + sj);
            const uint32_t o = cb + sl.col_offset - w.win_lo;
            uint32_t ci[RPT];
            slot_index_rows<RPT>(sl, uniform_load(m.tslots + sj), m.tern, m.dense, T, ci);
#pragma unroll
            for (int i = 0; i < RPT; i++) {
                const uint32_t uc = o + ci[i];
                if (ok[i] && uc < W) atomicAdd(&win[uc], yv[i]);
            }
        }
    }
    __syncthreads();
    VT* __restrict__ out = part + w.part_off;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) out[i] = VT(win[i]);
}

// Stage 2: z[c] = sum of the window's chunk slices, in chunk order
// (the order of the slices is fixed; the sums inside a slice took the LDS
// atomics' arrival order); columns outside every window get 0.
// Column c's sum over the slices of its window: integer counts stay
// integer; real sums add up in FP64, in chunk order, eight slices loaded
// ahead of their additions.
template <class T>
__device__ __forceinline__ std::conditional_t<std::is_integral_v<T>, T, double>
reduce_column(const T* __restrict__ part, const ReduceWork& w, int64_t c)
{
    using AccT = std::conditional_t<std::is_integral_v<T>, T, double>;
    const T* __restrict__ p = part + w.part_off + (c - w.col_begin);
    AccT s = AccT(0);
    int j = 0;
    for (; j + 8 <= w.nchunk; j += 8) {
        T t[8];
#pragma unroll
        for (int q = 0; q < 8; q++) t[q] = p[int64_t(j + q) * w.stride];
#pragma unroll
        for (int q = 0; q < 8; q++) s += AccT(t[q]);
    }
    for (; j < w.nchunk; j++) s += AccT(p[int64_t(j) * w.stride]);
    return s;
}

template <class T, class ZT>
__global__ __launch_bounds__(kRowBlock) void k_scatter_reduce(const ReduceWork* __restrict__ work,
                                                              const T* __restrict__ part, ZT* __restrict__ z,
                                                              const LsqrScalars* div, int mode)
{
    const ReduceWork w = work[blockIdx.x];
    if (!scatter_selected(div, mode, w.sys)) return;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) z[c] = ZT(reduce_column(part, w, c));
}

// Jacobi of A^T A (Math/MatrixCSR.cpp:21-47): diag_c = sum over rows of the
// squared multiplicity of c in the row.  A row's columns of one slot group
// (repeats can only occur within a group, of at most kMaxSegSlots slots) are
// loaded once into registers, and the occurrences of a column add up to its
// mult^2.  Counts are exact 32-bit integers.
template <class View>
__global__ __launch_bounds__(kScatterBlock) void k_jacobi_part(View m, const ScatterWork* __restrict__ work,
                                                               const SlotGroup* __restrict__ groups,
                                                               const uint16_t* __restrict__ slot_list,
                                                               const uint32_t* __restrict__ sys_col,
                                                               uint32_t* __restrict__ part)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    uint32_t* win = reinterpret_cast<uint32_t*>(smem_raw);
    const ScatterWork w = work[blockIdx.x];
    const uint32_t W = w.win_hi - w.win_lo;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) win[i] = 0u;
    __syncthreads();
    const SlotGroup gr = uniform_load(groups + w.group);
    const uint32_t cb = sys_col[w.sys];
    const int ns = gr.nslots;
    // four rows a thread per step, all their entries loaded before any add
    constexpr int kRows = 4;
    // stored indices: the group's slots are looked up once, outside the row
    // loop (slots past the group's end repeat its first, unused)
    using SlotRef = std::conditional_t<is_slot_major<View>::value, typename SlotRefOf<View>::type, int>;
    SlotRef sref[kMaxSegSlots];
    if constexpr (is_slot_major<View>::value) {
#pragma unroll
        for (int j = 0; j < kMaxSegSlots; j++) sref[j] = m.slot(uniform_u16(slot_list, gr.first + (j < ns ? j : 0)));
    }
    const int64_t last = w.row_end - 1;
    for (int64_t r0 = w.row_begin + threadIdx.x; r0 < w.row_end; r0 += kRows * kScatterBlock) {
        uint32_t c[kRows][kMaxSegSlots];
#pragma unroll
        for (int u = 0; u < kRows; u++) {
            const int64_t r = r0 + int64_t(u) * kScatterBlock;
            const bool valid = r < w.row_end;
#pragma unroll
            for (int j = 0; j < kMaxSegSlots; j++) {
                if constexpr (is_slot_major<View>::value) {
                    // every load issued (a clamped row, the group's first slot
                    // past its end), then the unused ones dropped: no branch
                    // around a load, so all of them are in flight together
                    const uint32_t v = cb + sref[j].col(min(r, last));
                    c[u][j] = (valid && j < ns) ? v : 0xFFFFFFFFu;
                } else {
                    c[u][j] = (valid && j < ns) ? cb + m.col(slot_list[gr.first + j], r) : 0xFFFFFFFFu;
                }
            }
        }
        // a column that m slots of the row hold must gain m^2 = 1 + 3 + ...
        // + (2m - 1): its k-th occurrence adds 2k - 1, i.e. one plus twice
        // the number of earlier slots holding it
#pragma unroll
        for (int u = 0; u < kRows; u++) {
#pragma unroll
            for (int j = 0; j < kMaxSegSlots; j++) {
                const uint32_t uc = c[u][j] - w.win_lo;
                if (c[u][j] == 0xFFFFFFFFu || uc >= W) continue;
                uint32_t before = 0;
#pragma unroll
                for (int t = 0; t < j; t++) before += c[u][t] == c[u][j] ? 1u : 0u;
                atomicAdd(&win[uc], 2u * before + 1u);
            }
        }
    }
    __syncthreads();
    uint32_t* __restrict__ out = part + w.part_off;
    for (uint32_t i = threadIdx.x; i < W; i += kScatterBlock) out[i] = win[i];
}

// 1 / diag in float, exactly as inv() of the reference (1/0 = inf).
__global__ void k_jacobi_finish(const uint32_t* __restrict__ cnt, int64_t n, float* __restrict__ d)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        d[i] = 1.0f / float(cnt[i]);
}

// Column scaling of LSQR: D = 1/|a_c| = sqrt(jacobi), 1 for empty columns.
__global__ void k_colscale(const uint32_t* __restrict__ cnt, int64_t n, double* __restrict__ D)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        D[i] = cnt[i] ? 1.0 / sqrt(double(cnt[i])) : 1.0;
}

// ------------------------------------------------------------ LSQR cols ----
// LSQR step 2 (Math/Solver.h:162), fused with the column reduction of
// A^T (u / beta): per column z = the slices' sum rounded to VT (as
// k_scatter_reduce writes it), then vbar = D z - beta v and the block
// partials of |vbar|^2, one block per reduce item (g: the reduce items of
// each system); k_lsqr_step (LSQR) or k_lsmr_rot (LSMR) finishes alpha.  The
// system's first block also keeps rhobar, phibar and stop aside for
// k_lsqr_step, whose blocks read them while its first block updates them.
// (A grid-wide last-block ticket here would cost every block an L2
// write-back and invalidate on gfx950.)
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_lsqr_reduce_v(const ReduceWork* __restrict__ work,
                                                             const VT* __restrict__ part,
                                                             const double* __restrict__ D, double* __restrict__ v,
                                                             GridSys g, LsqrScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ReduceWork w = work[blockIdx.x];
    LsqrScalars* st = &scal[w.sys];
    if (!st->active) return;
    if (threadIdx.x == 0 && int(blockIdx.x) == g.blk_first[w.sys]) {
        st->rhobar0 = st->rhobar;
        st->phibar0 = st->phibar;
        st->stop0 = st->stop;
    }
    const double beta = st->beta;
    double acc = 0.0;
    if (st->stop == 0)
        for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
            const double zc = double(VT(reduce_column(part, w, c)));
            const double vb = D[c] * zc - beta * v[c];
            v[c] = vb;
            acc += vb * vb;
        }
    acc = block_sum<kRowBlock>(acc, sh);
    if (threadIdx.x == 0) g.partials[2 * blockIdx.x] = acc;
}

// The LSQR rotation and update in one launch (Math/Solver.h:165-175).  Every
// block of a system forms alpha = |vbar| from k_lsqr_reduce_v's partials and
// the rotation (rho, c, s, theta, phi, phibar) from the rhobar, phibar and
// stop that k_lsqr_reduce_v kept aside -- the same values in every block --
// and updates its columns: v = vbar / alpha, vd = D v, y += (phi/rho) w,
// w = v - (theta/rho) w.  The system's first column block stores the new
// state, with the device-side stop; the fields it writes are read by no
// block of this launch.
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_lsqr_step(const ColWork* __restrict__ work, GridSys g,
                                                         const double* __restrict__ D, double* __restrict__ v,
                                                         VT* __restrict__ vd, double* __restrict__ w_,
                                                         double* __restrict__ y, LsqrScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ColWork w = work[blockIdx.x];
    LsqrScalars* st = &scal[w.sys];
    if (!st->active) return;
    const double s = system_sum<kRowBlock>(g, w.sys, sh);
    const bool writer = w.first != 0 && threadIdx.x == 0;
    if (st->stop0 != 0) {
        // the iteration after the stop: no further step
        if (writer) {
            st->stop = 2;
            st->step_x = 0.0;
            st->step_w = 0.0;
        }
        return;
    }
    const double beta = st->beta, rhobar = st->rhobar0, phibar = st->phibar0;
    const double alpha = sqrt(s);
    const double rho = sqrt(rhobar * rhobar + beta * beta);
    const double c = rho != 0.0 ? rhobar / rho : 0.0;
    const double sn = rho != 0.0 ? beta / rho : 0.0;
    const double theta = sn * alpha;
    const double phi = c * phibar;
    const double sx = rho != 0.0 ? phi / rho : 0.0;
    const double sw = rho != 0.0 ? theta / rho : 0.0;
    if (writer) {
        const double phibar_new = sn * phibar;
        st->rhobar = -c * alpha;
        st->phibar = phibar_new;
        st->step_x = sx;
        st->step_w = sw;
        st->alpha = alpha;
        st->nres = phibar_new * alpha * fabs(c);
        st->rnorm = phibar_new;
        st->anorm2 += st->beta2;
        st->anorm2 += s;
        st->iters += 1;
        // met the target: this iteration's step is the last (a rank-deficient
        // system that continued would iterate on rounding noise)
        if (st->target > 0.0 && st->nres <= st->target * st->nres0) st->stop = 1;
    }
    const double ia = alpha != 0.0 ? 1.0 / alpha : 0.0;
    for (int64_t c0 = w.col_begin + threadIdx.x; c0 < w.col_end; c0 += kRowBlock) {
        const double vn = v[c0] * ia;
        v[c0] = vn;
        vd[c0] = VT(D[c0] * vn);
        const double wc = w_[c0];
        y[c0] += sx * wc;
        w_[c0] = vn - sw * wc;
    }
}

// LSQR start on the columns (Math/Solver.h:146-150), from an FP64 A^T u_raw:
//   vbar = D z / beta, alpha = |vbar|; also |z|^2 = |A^T src|^2 into reduce_b.
template <class ZT>
__global__ __launch_bounds__(kRowBlock) void k_lsqr_init_v(const ColWork* __restrict__ work,
                                                           const double* __restrict__ D, const ZT* __restrict__ z,
                                                           double* __restrict__ v, GridSys g, LsqrScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ColWork w = work[blockIdx.x];
    LsqrScalars* st = &scal[w.sys];
    if (!st->active || !st->restart) return;
    const double ib = st->beta != 0.0 ? 1.0 / st->beta : 0.0;
    double a0 = 0.0, a1 = 0.0;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const double zc = double(z[c]);
        const double vb = D[c] * zc * ib;
        v[c] = vb;
        a0 += vb * vb;
        a1 += zc * zc;
    }
    a0 = block_sum<kRowBlock>(a0, sh);
    a1 = block_sum<kRowBlock>(a1, sh);
    if (threadIdx.x == 0) {
        g.partials[2 * blockIdx.x] = a0;
        g.partials[2 * blockIdx.x + 1] = a1;
    }
    if (arrive_last(&g.counters[w.sys], g.blk_count[w.sys])) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = threadIdx.x; i < g.blk_count[w.sys]; i += kRowBlock) {
            s0 += g.partials[2 * (g.blk_first[w.sys] + i)];
            s1 += g.partials[2 * (g.blk_first[w.sys] + i) + 1];
        }
        s0 = block_sum<kRowBlock>(s0, sh);
        s1 = block_sum<kRowBlock>(s1, sh);
        if (threadIdx.x == 0) {
            const double alpha = sqrt(s0);
            st->alpha = alpha;
            st->rhobar = alpha;
            st->phibar = st->beta;
            st->nres0 = alpha * st->beta;
            st->nres = st->nres0;
            st->rnorm = st->beta;
            st->anorm2 = s0 + st->beta * 0.0;
            st->reduce_b = s1;
            st->iters = 0;
            st->stop = 0;
        }
    }
}

// v = vbar/alpha, vd = D v, w = v, y = 0.
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_lsqr_init_w(const ColWork* __restrict__ work,
                                                           const double* __restrict__ D, double* __restrict__ v,
                                                           VT* __restrict__ vd, double* __restrict__ w_,
                                                           double* __restrict__ y, const LsqrScalars* scal)
{
    const ColWork w = work[blockIdx.x];
    const LsqrScalars* st = &scal[w.sys];
    if (!st->active || !st->restart) return;
    const double ia = st->alpha != 0.0 ? 1.0 / st->alpha : 0.0;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const double vn = v[c] * ia;
        v[c] = vn;
        vd[c] = VT(D[c] * vn);
        w_[c] = vn;
        y[c] = 0.0;
    }
}

// x += D y (back to the unscaled space), y = 0, keeping the x before in
// xprev; for the selected systems (active and restart).
__global__ __launch_bounds__(kRowBlock) void k_lsqr_apply(const ColWork* __restrict__ work,
                                                          const double* __restrict__ D, double* __restrict__ y,
                                                          double* __restrict__ x, double* __restrict__ xprev,
                                                          const LsqrScalars* scal)
{
    const ColWork w = work[blockIdx.x];
    if (!scal[w.sys].active || !scal[w.sys].restart) return;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const double xc = x[c];
        xprev[c] = xc;
        x[c] = xc + D[c] * y[c];
        y[c] = 0.0;
    }
}

// Undoes the last k_lsqr_apply of the selected systems: x = xprev.
__global__ __launch_bounds__(kRowBlock) void k_lsqr_undo(const ColWork* __restrict__ work,
                                                         const double* __restrict__ D, double* __restrict__ y,
                                                         double* __restrict__ x, double* __restrict__ xprev,
                                                         const LsqrScalars* scal)
{
    (void)D;
    (void)y;
    const ColWork w = work[blockIdx.x];
    if (!scal[w.sys].active || !scal[w.sys].restart) return;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) x[c] = xprev[c];
}

// ------------------------------------------------------------ LSMR ------
// The symmetric Givens rotation of Math/Solver.h:193-217 (scipy's sym_ortho).
__device__ __forceinline__ void sym_ortho(double a, double b, double& c, double& s, double& r)
{
    const double sa = a < 0 ? -1.0 : (a > 0 ? 1.0 : 0.0), sb = b < 0 ? -1.0 : (b > 0 ? 1.0 : 0.0);
    if (b == 0.0) {
        c = sa;
        s = 0.0;
        r = fabs(a);
    } else if (a == 0.0) {
        c = 0.0;
        s = sb;
        r = fabs(b);
    } else if (fabs(b) > fabs(a)) {
        const double tau = a / b;
        s = sb / sqrt(1.0 + tau * tau);
        c = s * tau;
        r = b / s;
    } else {
        const double tau = b / a;
        c = sa / sqrt(1.0 + tau * tau);
        s = c * tau;
        r = a / c;
    }
}

// LSMR's rotations (Math/Solver.h:255-275), one block per system: alpha =
// |vbar| from k_lsqr_reduce_v's partials, then the two rotations, zeta and
// the update coefficients; nres = |zetabar| (= |A^T r_k|) and the device-side
// stop as in k_lsqr_step.
__global__ __launch_bounds__(kRowBlock) void k_lsmr_rot(GridSys g, LsqrScalars* scal, LsmrState* lsmr)
{
    __shared__ double sh[kRowBlock / 64];
    const int sys = blockIdx.x;
    LsqrScalars* st = &scal[sys];
    if (!st->active) return;
    double sum = 0.0;
    for (int i = threadIdx.x; i < g.blk_count[sys]; i += kRowBlock) sum += g.partials[2 * (g.blk_first[sys] + i)];
    sum = block_sum<kRowBlock>(sum, sh);
    if (threadIdx.x != 0) return;
    LsmrState* ls = &lsmr[sys];
    if (st->stop != 0) {
        st->stop = 2;
        return;
    }
    const double beta = st->beta, alpha = sqrt(sum);
    const double rho_old = ls->rho, rhobar_old = ls->rhobar;
    double c, s, rho;
    sym_ortho(ls->alphabar, beta, c, s, rho);
    const double theta_new = s * alpha;
    ls->alphabar = c * alpha;
    const double thetabar = ls->sbar * rho;
    double cbar, sbar, rhobar;
    sym_ortho(ls->cbar * rho, theta_new, cbar, sbar, rhobar);
    const double zeta = cbar * ls->zetabar;
    ls->zetabar = -sbar * ls->zetabar;
    ls->cbar = cbar;
    ls->sbar = sbar;
    ls->rho = rho;
    ls->rhobar = rhobar;
    ls->k_hb = -thetabar * rho / (rho_old * rhobar_old);
    ls->k_x = zeta / (rho * rhobar);
    ls->k_h = rho != 0.0 ? -theta_new / rho : 0.0;
    st->alpha = alpha;
    st->nres = fabs(ls->zetabar);
    st->anorm2 += st->beta2;
    st->anorm2 += sum;
    st->iters += 1;
    if (st->target > 0.0 && st->nres <= st->target * st->nres0) st->stop = 1;
}

// LSMR's vector updates (Math/Solver.h:277-283): v = vbar / alpha, vd = D v,
// hbar = h + k_hb hbar, x += k_x hbar, h = v + k_h h.
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_lsmr_update(const ColWork* __restrict__ work,
                                                           const double* __restrict__ D, double* __restrict__ v,
                                                           VT* __restrict__ vd, double* __restrict__ h,
                                                           double* __restrict__ hb, double* __restrict__ x,
                                                           const LsqrScalars* scal, const LsmrState* lsmr)
{
    const ColWork w = work[blockIdx.x];
    const LsqrScalars* st = &scal[w.sys];
    if (!st->active || st->stop > 1) return;
    const LsmrState* ls = &lsmr[w.sys];
    const double ia = st->alpha != 0.0 ? 1.0 / st->alpha : 0.0;
    const double khb = ls->k_hb, kx = ls->k_x, kh = ls->k_h;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const double vn = v[c] * ia;
        v[c] = vn;
        vd[c] = VT(D[c] * vn);
        const double hc = h[c];
        const double hbn = hc + khb * hb[c];
        hb[c] = hbn;
        x[c] += kx * hbn;
        h[c] = vn + kh * hc;
    }
}

// LSMR's start after the LSQR start of each system (x = 0): the rotation
// state alphabar = alpha, zetabar = alpha beta, rho = rhobar = cbar = 1,
// sbar = 0 (Math/Solver.h:238-252).
__global__ void k_lsmr_init(int nsys, const LsqrScalars* scal, LsmrState* lsmr)
{
    const int sys = blockIdx.x * blockDim.x + threadIdx.x;
    if (sys >= nsys) return;
    LsmrState* ls = &lsmr[sys];
    ls->alphabar = scal[sys].alpha;
    ls->zetabar = scal[sys].alpha * scal[sys].beta;
    ls->rho = ls->rhobar = ls->cbar = 1.0;
    ls->sbar = 0.0;
    ls->k_hb = ls->k_x = ls->k_h = 0.0;
}

// |z|^2 per system into reduce_b (used for |A^T b| and |A^T r|).
template <class ZT>
__global__ __launch_bounds__(kRowBlock) void k_cols_norm2(const ColWork* __restrict__ work, const ZT* __restrict__ z,
                                                          GridSys g, LsqrScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ColWork w = work[blockIdx.x];
    LsqrScalars* st = &scal[w.sys];
    if (!st->active) return;
    double acc = 0.0;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const double zc = double(z[c]);
        acc += zc * zc;
    }
    acc = block_sum<kRowBlock>(acc, sh);
    if (threadIdx.x == 0) g.partials[2 * blockIdx.x] = acc;
    if (arrive_last(&g.counters[w.sys], g.blk_count[w.sys])) {
        double s = 0.0;
        for (int i = threadIdx.x; i < g.blk_count[w.sys]; i += kRowBlock) s += g.partials[2 * (g.blk_first[w.sys] + i)];
        s = block_sum<kRowBlock>(s, sh);
        if (threadIdx.x == 0) st->reduce_b = s;
    }
}

// out = D y (the unscaled solution of a column-scaled LSQR).
__global__ void k_scale_mul(const double* __restrict__ D, const double* __restrict__ y, int64_t n,
                            double* __restrict__ out)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = D[i] * y[i];
}

// ------------------------------------------------------------ PCG ---------
// PCG on A^T A with a diagonal preconditioner (Math/Solver.h:78-117); the
// FP64 scalars live on the device.  An iteration is five launches: the gather
// t = A p, the window tiles of A^T t, then
//   k_pcg_reduce_q  q = A^T t (the slices' sum per column) and the block
//                   partials of p.q, one block per reduce item;
//   k_pcg_update    alpha = rz / pq, formed by every block from those partials
//                   (the same value in each), then x, r, zz = P r and the
//                   partials of r.zz and r.r;
//   k_pcg_p         beta = rz' / rz formed the same way, then p = zz + beta p.
// A system's first column block stores the scalars.  k_pcg_update keeps rz
// aside as rz0, so that k_pcg_p's blocks read the old rz while its first
// block stores the new one.  faithful: the scalars are rounded to VT before
// use, as in the float solver.
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_pcg_reduce_q(const ReduceWork* __restrict__ work,
                                                            const VT* __restrict__ part, const VT* __restrict__ p,
                                                            VT* __restrict__ q, GridSys g, const PcgScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ReduceWork w = work[blockIdx.x];
    if (!scal[w.sys].active) return;
    double acc = 0.0;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const VT qc = VT(reduce_column(part, w, c));
        q[c] = qc;
        acc += double(p[c]) * double(qc);
    }
    acc = block_sum<kRowBlock>(acc, sh);
    if (threadIdx.x == 0) g.partials[2 * blockIdx.x] = acc;
}

// gq: k_pcg_reduce_q's partials; g: this launch's
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_pcg_update(const ColWork* __restrict__ work, const VT* __restrict__ P,
                                                          const VT* __restrict__ p, const VT* __restrict__ q,
                                                          VT* __restrict__ x, VT* __restrict__ r, VT* __restrict__ zz,
                                                          GridSys gq, GridSys g, PcgScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ColWork w = work[blockIdx.x];
    PcgScalars* st = &scal[w.sys];
    if (!st->active) return;
    const double pq = system_sum<kRowBlock>(gq, w.sys, sh);
    const double alpha = pq != 0.0 ? st->rz / pq : 0.0;
    if (w.first && threadIdx.x == 0) {
        st->pq = pq;
        st->alpha = alpha;
        st->rz0 = st->rz;
    }
    const VT a = VT(alpha);
    double a0 = 0.0, a1 = 0.0;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        x[c] += a * p[c];
        const VT rc = r[c] - a * q[c];
        r[c] = rc;
        const VT zc = P[c] * rc;
        zz[c] = zc;
        a0 += double(rc) * double(zc);
        a1 += double(rc) * double(rc);
    }
    a0 = block_sum<kRowBlock>(a0, sh);
    a1 = block_sum<kRowBlock>(a1, sh);
    if (threadIdx.x == 0) {
        g.partials[2 * blockIdx.x] = a0;
        g.partials[2 * blockIdx.x + 1] = a1;
    }
}

// rz' (slot 0 of k_pcg_update's partials), rr (slot 1), beta = rz' / rz;
// p = zz + beta p; q = 0 for the next scatter.
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_pcg_p(const ColWork* __restrict__ work, const VT* __restrict__ zz,
                                                     VT* __restrict__ p, VT* __restrict__ q, GridSys g,
                                                     PcgScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ColWork w = work[blockIdx.x];
    PcgScalars* st = &scal[w.sys];
    if (!st->active) return;
    double rz, rr;
    system_sum2<kRowBlock>(g, w.sys, sh, rz, rr);
    const double beta = st->rz0 != 0.0 ? rz / st->rz0 : 0.0;
    if (w.first && threadIdx.x == 0) {
        st->beta = beta;
        st->rz = rz;
        st->rr = rr;
        st->iters += 1;
    }
    const VT b = VT(beta);
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        p[c] = zz[c] + b * p[c];
        q[c] = VT(0);
    }
}

// PCG start: r = rhs (A^T b, x0 = 0), zz = P r, p = zz, rz = r.zz, q = 0.
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_pcg_init(const ColWork* __restrict__ work, const VT* __restrict__ P,
                                                        VT* __restrict__ rhs_r, VT* __restrict__ zz, VT* __restrict__ p,
                                                        VT* __restrict__ x, VT* __restrict__ q, GridSys g,
                                                        PcgScalars* scal)
{
    __shared__ double sh[kRowBlock / 64];
    const ColWork w = work[blockIdx.x];
    PcgScalars* st = &scal[w.sys];
    if (!st->active) return;
    double a0 = 0.0, a1 = 0.0;
    for (int64_t c = w.col_begin + threadIdx.x; c < w.col_end; c += kRowBlock) {
        const VT rc = rhs_r[c];
        const VT zc = P[c] * rc;
        zz[c] = zc;
        p[c] = zc;
        x[c] = VT(0);
        q[c] = VT(0);
        a0 += double(rc) * double(zc);
        a1 += double(rc) * double(rc);
    }
    a0 = block_sum<kRowBlock>(a0, sh);
    a1 = block_sum<kRowBlock>(a1, sh);
    if (threadIdx.x == 0) {
        g.partials[2 * blockIdx.x] = a0;
        g.partials[2 * blockIdx.x + 1] = a1;
    }
    if (arrive_last(&g.counters[w.sys], g.blk_count[w.sys])) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = threadIdx.x; i < g.blk_count[w.sys]; i += kRowBlock) {
            s0 += g.partials[2 * (g.blk_first[w.sys] + i)];
            s1 += g.partials[2 * (g.blk_first[w.sys] + i) + 1];
        }
        s0 = block_sum<kRowBlock>(s0, sh);
        s1 = block_sum<kRowBlock>(s1, sh);
        if (threadIdx.x == 0) {
            st->rz = s0;
            st->rr = s1;
            st->iters = 0;
        }
    }
}

// DiagonalPreconditioner (Math/Solver.h:29-34): non-normal entries become 1.
template <class VT>
__global__ void k_precond_from_jacobi(const float* __restrict__ jac, int64_t n, VT* __restrict__ P)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const float d = jac[i];
        P[i] = VT(__builtin_isnormal(d) ? d : 1.0f);
    }
}

// ------------------------------------------------------------ evaluator ----
// Forward of the differentiable evaluator: out[r] = sum_s w[idx[r,s]] in slot
// order (Pattern/Evaluator.cpp:48-54 sums the same terms in the same order).
template <class VT>
__global__ __launch_bounds__(kRowBlock) void k_eval_forward(const int32_t* __restrict__ idx, int64_t B, int k,
                                                            const VT* __restrict__ w, uint32_t cols,
                                                            VT* __restrict__ out)
{
    const int64_t r = blockIdx.x * int64_t(kRowBlock) + threadIdx.x;
    if (r >= B) return;
    VT acc = VT(0);
    for (int s = 0; s < k; s++) {
        const uint32_t c = uint32_t(idx[r * k + s]);
        if (c < cols) acc += w[c];
    }
    out[r] = acc;
}

// The Huber loss of d = pred - target (0.5 d^2 inside [-delta, delta],
// delta (|d| - delta/2) outside) and the gradient of the batch mean,
// clamp(d, -delta, delta) / B.
__device__ __forceinline__ double huber_term(float d, float delta, int64_t B, float& grad)
{
    const float ad = fabsf(d);
    grad = fminf(fmaxf(d, -delta), delta) / float(B);
    return ad < delta ? 0.5 * double(d) * double(d) : double(delta) * (double(ad) - 0.5 * double(delta));
}

// One Huber-loss training step, forward half: for the minibatch rows sel[b]
// the prediction (the evaluator's gather-sum, slot order from zero), the loss
// and g[b] (huber_term); the block sums of the losses (FP64) go to
// part[blockIdx.x], and each row is copied to rows[b * k ..] (contiguous, for
// the backward).  A thread takes one row, two entries per 8-byte load where k
// is even.
__global__ __launch_bounds__(kRowBlock) void k_train_forward(const int32_t* __restrict__ idx,
                                                             const int64_t* __restrict__ sel, int64_t B, int k,
                                                             const float* __restrict__ w, uint32_t cols,
                                                             const float* __restrict__ target, float delta,
                                                             float* __restrict__ g, double* __restrict__ part,
                                                             int32_t* __restrict__ rows)
{
    __shared__ double sh[kRowBlock / 64];
    const int64_t b = blockIdx.x * int64_t(kRowBlock) + threadIdx.x;
    double loss = 0.0;
    if (b < B) {
        const int64_t r = sel[b];
        float acc = 0.0f;
        if ((k & 1) == 0) {
            const int2* __restrict__ src = reinterpret_cast<const int2*>(idx + r * k);
            int2* __restrict__ dst = reinterpret_cast<int2*>(rows + b * k);
#pragma unroll 3
            for (int j = 0; j < k / 2; j++) {
                const int2 c = src[j];
                dst[j] = c;
                if (uint32_t(c.x) < cols) acc += w[c.x];
                if (uint32_t(c.y) < cols) acc += w[c.y];
            }
        } else {
            const int32_t* __restrict__ src = idx + r * k;
            int32_t* __restrict__ dst = rows + b * k;
            for (int j = 0; j < k; j++) {
                const int32_t c = src[j];
                dst[j] = c;
                if (uint32_t(c) < cols) acc += w[c];
            }
        }
        float gr;
        loss = huber_term(acc - target[r], delta, B, gr);
        g[b] = gr;
    }
    loss = block_sum<kRowBlock>(loss, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = loss;
}

// The step's Adam update (torch.optim.Adam without weight decay or amsgrad:
// m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) g^2, w -= lr / (1 - b1^t) * m /
// (sqrt(v) / sqrt(1 - b2^t) + eps)); block 0 also adds the step's mean loss
// (the forward's block sums, reduced in a fixed order, / B) to *loss.
struct AdamStep {
    float lr, b1, b2, eps, bias1, bias2_sqrt;
};
__device__ __forceinline__ void adam_update(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                            int64_t c, float gc, const AdamStep& a)
{
    const float mc = a.b1 * m[c] + (1.0f - a.b1) * gc;
    const float vc = a.b2 * v[c] + (1.0f - a.b2) * gc * gc;
    m[c] = mc;
    v[c] = vc;
    w[c] -= (a.lr / a.bias1) * mc / (sqrtf(vc) / a.bias2_sqrt + a.eps);
}

// The loss of the step: the forward's block partials, in a fixed order, / B.
__device__ __forceinline__ void add_loss(const double* __restrict__ part, int nparts, int64_t B,
                                         double* __restrict__ loss)
{
    __shared__ double sh[kRowBlock / 64];
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += kRowBlock) s += part[i];
    s = block_sum<kRowBlock>(s, sh);
    if (threadIdx.x == 0) *loss += s / double(B);
}

__global__ __launch_bounds__(kRowBlock) void k_adam(float* __restrict__ w, const float* __restrict__ grad,
                                                    float* __restrict__ m, float* __restrict__ v, int64_t n,
                                                    AdamStep a, const double* __restrict__ part, int nparts,
                                                    int64_t B, double* __restrict__ loss)
{
    for (int64_t c = blockIdx.x * int64_t(kRowBlock) + threadIdx.x; c < n; c += int64_t(gridDim.x) * kRowBlock)
        adam_update(w, m, v, c, grad[c], a);
    if (blockIdx.x == 0) add_loss(part, nparts, B, loss);
}

// The backward's column reduction with the Adam step fused: per column the
// gradient is the window slices' sum rounded to float (as k_scatter_reduce
// writes it), then the step of k_adam -- one launch fewer a step, the same
// values.  One block per reduce item (the items cover every column once).
__global__ __launch_bounds__(kRowBlock) void k_adam_reduce(const ReduceWork* __restrict__ work,
                                                           const float* __restrict__ gpart, float* __restrict__ w,
                                                           float* __restrict__ m, float* __restrict__ v, AdamStep a,
                                                           const double* __restrict__ part, int nparts, int64_t B,
                                                           double* __restrict__ loss)
{
    const ReduceWork rw = work[blockIdx.x];
    for (int64_t c = rw.col_begin + threadIdx.x; c < rw.col_end; c += kRowBlock)
        adam_update(w, m, v, c, float(reduce_column(gpart, rw, c)), a);
    if (blockIdx.x == 0) add_loss(part, nparts, B, loss);
}

// ------------------------------------------------------------ data ---------
__global__ __launch_bounds__(kRowBlock) void k_random_play(uint64_t seed, int64_t n, int e_lo, int e_hi,
                                                           uint64_t* __restrict__ out)
{
    const int64_t i = blockIdx.x * int64_t(kRowBlock) + threadIdx.x;
    if (i >= n) return;
    const Pos p = random_play(seed, uint64_t(i), e_lo, e_hi);
    out[2 * i] = p.P;
    out[2 * i + 1] = p.O;
}

__global__ void k_targets_x(uint64_t seed, int64_t cols, double sx, double* __restrict__ xt)
{
    for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < cols; c += int64_t(gridDim.x) * blockDim.x) {
        Stream s(seed, uint64_t(c));
        xt[c] = __dmul_rn(sx, s.ih12());
    }
}

}  // namespace
}  // namespace cad
