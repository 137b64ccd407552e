// Plain-old-data structures shared by the host runtime and the HIP kernels.
#pragma once
#include <cstdint>

namespace cad {

// One slot of the index program (see patterns.hpp SlotProgram), packed for
// uniform (scalar) loads on the device.
struct DevSlot {
    uint8_t kind;        // 0: a, 1: pair(a,b), 2: pair(a,b)*ds + d, 3: dense[a]
    uint8_t len[3];      // square-list lengths
    uint16_t off[3];     // offsets into the square array
    uint16_t pad;
    uint32_t hs, ds;
    uint32_t col_offset;
    uint32_t dense_offset;
    uint32_t pad2;
};
static_assert(sizeof(DevSlot) == 32, "DevSlot layout");

// The same slot program in the form the matrix-free products evaluate: each
// square list as entries of a ternary-word program.  A row's position is held
// as four 32-bit "ternary words" (square q's digit P_q + 2 O_q in bits
// 2(q mod 16) of word q / 16); an entry is two 32-bit words {shift, 3^j}: the
// digit of word w at that shift, weighed by 3^j.  List l's entries start at
// entry off[l], grouped by word (word w: (nent[l] >> 8w) & 255 entries), with
// no padding between lists; the array ends with three zero entries, so that a
// load of four entries at any list's last entry stays inside it.
struct DevTern {
    uint32_t off[3];
    uint32_t nent[3];
    uint32_t pad[2];
};
static_assert(sizeof(DevTern) == 32, "DevTern layout");

// A block of rows of one system, for the row-parallel kernels (SpMV and the
// rows-length vector kernels).  Blocks of one system are contiguous in the
// work list.
struct RowWork {
    int64_t row_begin, row_end;  // global rows
    int32_t sys;
    int32_t pad;
};

// A block of columns of one system, for the cols-length vector kernels.
struct ColWork {
    int64_t col_begin, col_end;  // global columns
    int32_t sys;
    int32_t first;               // 1 on the first column block of its system
};

// A scatter tile of A^T y: rows [row_begin,row_end) x the slots of `group`,
// accumulating the column window [win_lo, win_hi) in LDS and storing it to
// partial[part_off ..] (one slice per row chunk of the window).
struct ScatterWork {
    int64_t row_begin, row_end;  // global rows
    int64_t part_off;            // this tile's slice of the partial buffer
    uint32_t win_lo, win_hi;     // global columns
    int32_t group;               // index into the slot-group table
    int32_t sys;
};

// A column range inside one scatter window (or outside every window, nchunk
// = 0): z[c] = sum over the window's chunk slices.
struct ReduceWork {
    int64_t col_begin, col_end;  // global columns
    int64_t part_off;            // slice offset of col_begin in chunk 0
    int32_t nchunk;              // chunk slices of the window
    uint32_t stride;             // slice length (window width)
    int32_t sys;
    int32_t pad;
};

// A run of consecutive slots whose columns lie in [col_lo, col_lo + width)
// of a system (one pattern block), for the LDS-staged gather.  Blocks wider
// than the LDS are staged in nwin windows of wmax columns.
struct GatherSeg {
    uint32_t col_lo, width;
    uint16_t slot_first, nslots;
    uint16_t nwin, pad;
};

// Slots of one gather step (and of a slot group that the scatter tile
// handles with hoisted slot offsets).
constexpr int kMaxSegSlots = 8;

// Slots of one step of a segment wider than the LDS in the stored-index tile
// gather: the step holds every slot's value of each of its rows until the last
// window, and with more than 8 bytes of entries per slot (rows per thread x
// index bytes) the registers hold the values of half as many slots.
__host__ __device__ constexpr int wide_step_slots(int entry_bytes)
{
    return entry_bytes <= 8 ? kMaxSegSlots : kMaxSegSlots / 2;
}

// One step of the stored-index tile gather: up to kMaxSegSlots consecutive
// slots of one segment.  A step with phase_count > 0 first stages the phase's
// segments (phases[phase_first ..]) into LDS side by side; the segment's
// columns start at lds_off there.  A step with nwin > 1 is a segment wider than
// the LDS, staged window by window by the step itself.  off[j] is slot
// (slot_first + j)'s offset from its stored entry to the staged column:
// base - col_lo (+ lds_off for a phased step), modulo 2^32.  64 bytes, read
// by one scalar load.  A wide segment split into several steps is walked
// window 0 up, then down, and so on (wflags): each step after the first starts
// on the window its predecessor ended on, which is still staged.
constexpr uint16_t kWideReverse = 1;  // windows from the last to the first
constexpr uint16_t kWideStaged = 2;   // the first window visited is in LDS already
struct GatherStep {
    uint32_t off[kMaxSegSlots];
    uint16_t slot_first, nslots;
    int16_t phase_first, phase_count;
    uint32_t col_lo;   // the segment's first column (system-local)
    uint32_t lds_off;  // the segment's offset in the staged phase
    uint32_t width;    // the segment's width (windows of a wide segment)
    uint16_t nwin, wflags;
    uint32_t pad2[2];
};
static_assert(sizeof(GatherStep) == 64, "GatherStep is one 64-byte scalar load");

// A segment of a staging phase: columns [col_lo, col_lo + width) go to
// LDS [lds_off, lds_off + width).
struct PhaseSeg {
    uint32_t col_lo, width, lds_off, pad;
};

// Slots whose column ranges overlap: a column index can repeat inside a row
// only among the slots of one group.
struct SlotGroup {
    uint32_t col_lo, col_hi;  // union of the slots' column ranges (local to a system)
    uint16_t nslots, first;   // slots: slot_list[first .. first+nslots)
};

// One variation of the expanded-table evaluator: its mask's squares in
// ascending order (FastIndex digits), 3^n, and where its table starts.
struct EvalSlot {
    uint8_t sq[16];
    uint8_t n;
    uint8_t pad[3];
    uint32_t n3;
    uint64_t table_offset;
};
static_assert(sizeof(EvalSlot) == 32, "EvalSlot layout");

// Per-system state of the device-resident LSQR recurrence (FP64 scalars).
struct LsqrScalars {
    double alpha, beta;       // current bidiagonalisation scalars
    double rhobar, phibar;
    double step_x, step_w;    // phi/rho and theta/rho of the last rotation
    double nres;              // |Abar^T r_k| = phibar_{k+1} alpha_{k+1} |c_k|
    double nres0;             // alpha_1 beta_1 = |Abar^T b|
    double rnorm;             // phibar = |b - Abar y_k|
    double anorm2;            // sum of alpha^2 + beta^2 (Frobenius estimate of Abar)
    double reduce_a, reduce_b;// scratch for explicit reductions
    double target;            // stop on the device once nres <= target * nres0 (0: never)
    int32_t active;           // 0: converged / frozen by the host
    int32_t iters;
    int32_t stop;             // 0 running; 1 target met this iteration (its step still applies); 2 stopped
    int32_t restart;          // selects the systems that the start / apply kernels (re)initialise
    double beta2;             // |u_raw|^2 of the iteration's gather (beta = sqrt(beta2))
    double rhobar0, phibar0;  // rhobar, phibar and stop as the iteration's rotation found them
    int32_t stop0;
    int32_t pad;
    double pad2;
};
static_assert(sizeof(LsqrScalars) % 16 == 0, "LsqrScalars layout");

// Per-system state of the device-resident LSMR rotations (Fong & Saunders
// 2011, damp = 0; Math/Solver.h:219-287), beside the LsqrScalars of the
// shared Golub-Kahan bidiagonalisation; k_hb, k_x, k_h are the last
// iteration's update coefficients of hbar, x and h.
struct LsmrState {
    double alphabar, zetabar, rho, rhobar, cbar, sbar;
    double k_hb, k_x, k_h, pad;
};

// Per-system state of the device-resident PCG recurrence.
struct PcgScalars {
    double rz, rz_new, pq, alpha, beta, rr;
    int32_t active, iters;
    double rz0;  // rz as the iteration's update found it
};

// Per-system grid bookkeeping shared by the reducing kernels.
struct GridSys {
    const int32_t* __restrict__ blk_first;  // first block of each system
    const int32_t* __restrict__ blk_count;  // blocks of each system
    unsigned* counters;                     // per system
    double* partials;                       // per block (two slots: [2*b], [2*b+1])
};

}  // namespace cad
