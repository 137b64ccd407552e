/*
 * CPU oracle for the pattern-evaluator fit.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain CPU restatement of the
 * reference algorithms (dominichofer/Cassandra) and exists to check the HIP
 * library.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker / the timed CPU baseline.  The
 * product library (cassandra_amd) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  What pins each
 * function (golden vectors, closed forms, brute force) is listed next to it.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Board/BitBoard.cpp:4-84 : symmetry maps on bit b = x + 8y ----------
 * Restated as square permutations (loop over 64 squares), deliberately not
 * with the bit tricks the product uses.  Pinned by BoardTest/BitBoard.cpp:8-23
 * (images of 0xF), by all 64 single-bit images and by the group laws. */
uint64_t or_flip_vertical(uint64_t b);   /* (x,y) -> (x,7-y)      */
uint64_t or_flip_horizontal(uint64_t b); /* (x,y) -> (7-x,y)      */
uint64_t or_flip_diagonal(uint64_t b);   /* (x,y) -> (y,x)        */
uint64_t or_flip_codiagonal(uint64_t b); /* (x,y) -> (7-y,7-x)    */

/* ---- Base/Bit.h:89-98 : PExt / PDep (loop form).  Pinned against x86
 * BMI2 pext/pdep by brute force in the CPU tests. */
uint64_t or_pext(uint64_t src, uint64_t mask);
uint64_t or_pdep(uint64_t src, uint64_t mask);

/* ---- Pattern/Helpers.cpp:44-48 : FastIndex = sum_i 3^i * (P_si + 2 O_si)
 * over the set bits s_0 < s_1 < ... of the mask.  Pinned by the closed form
 * and exhaustive enumeration over small masks. */
uint32_t or_fast_index(uint64_t P, uint64_t O, uint64_t mask);

/* ---- Pattern/Helpers.cpp:81-102 : Configurations(mask), p outer / o inner,
 * skipping p&o != 0.  Writes at most cap (P,O) pairs; returns the count. */
int64_t or_configurations(uint64_t mask, uint64_t* out_PO, int64_t cap);

/* ---- Pattern/Indexer.cpp:18-422 : CreateIndexer + GroupIndexer ---------- */
typedef struct or_group or_group;
or_group* or_group_create(const uint64_t* masks, int n);
void      or_group_free(or_group*);
int       or_group_k(const or_group*);      /* elements per row            */
int64_t   or_group_cols(const or_group*);   /* index_space_size            */
int       or_group_npatterns(const or_group*);
/* class: 0=A 1=V 2=D 3=VH 4=DC 5=VHDC; transformed mask; size; group order */
void      or_group_pattern_info(const or_group*, int i, int* cls, uint64_t* mask,
                                int64_t* index_space_size, int* order);
void      or_group_variations(const or_group*, uint64_t* out_k);
/* rows of k indices (row-major n x k) for n positions given as (P,O) pairs */
void      or_group_indices(const or_group*, const uint64_t* PO, int64_t n,
                           uint32_t* out);
/* single-pattern Indexer::Index (Pattern/Indexer.cpp Index methods) */
uint32_t  or_group_pattern_index(const or_group*, int i, uint64_t P, uint64_t O);

/* ---- Math/MatrixCSR.cpp:21-87 : fixed-k CSR, row-major idx[rows*k] ------ */
/* faithful float A*x: sequential in-row sum from 0.0f (MatrixCSR.cpp:59-62) */
void or_spmv_f32(const uint32_t* idx, int64_t rows, int k, const float* x, float* y);
void or_spmv_f64(const uint32_t* idx, int64_t rows, int k, const double* x, double* y);
/* faithful float x*A: per-thread private column vectors + critical reduce
 * (MatrixCSR.cpp:67-87); nthreads<=0 means OpenMP default */
void or_spmtv_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                  const float* y, float* z, int nthreads);
/* FP64, sequential row order (deterministic) */
void or_spmtv_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                  const double* y, double* z);
/* MatrixCSR.cpp:21-47 : 1/sum_rows mult^2 (inf for unseen columns) */
void or_jacobi_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols, float* d);

/* ---- Math/Solver.h : solvers on A (CSR) ---------------------------------
 * PCG on A^T A with the Jacobi preconditioner, x0 = 0, rhs A^T b: this is
 * FitWeights (PatternFit/PatternFit.cpp:7-39).  f32 is the faithful
 * float restatement (Solver.h:78-117), f64 the double twin.
 * resid_hist (may be NULL) receives Residuum() = r.r after each iteration. */
void or_pcg_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                const float* b, int iterations, float* x, float* resid_hist);
void or_pcg_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                const double* b, int iterations, double* x, double* resid_hist);
/* CG on A^T A, rhs A^T b, x0 = 0 (Solver.h:40-76) */
void or_cg_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
               const double* b, int iterations, double* x, double* resid_hist);
/* LSQR (Solver.h:119-182, Paige & Saunders 1982), x0 = 0.
 * colscale (may be NULL): solve min |A D y - b| and return x = D y, where
 * D = colscale (cols).  phibar_hist (may be NULL) receives Residuum() = phibar
 * after each iteration; nres_hist (may be NULL) receives the recurrence
 * estimate |A^T r_k| = phibar_{k+1} alpha_{k+1} |c_k| (in the scaled space). */
void or_lsqr_f32(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                 const float* b, int iterations, float* x, float* phibar_hist);
void or_lsqr_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                 const double* b, int iterations, const double* colscale,
                 double* x, double* phibar_hist, double* nres_hist);
/* LSMR (Solver.h:219-287, Fong & Saunders 2011), x0 = 0, damp = 0 */
void or_lsmr_f64(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                 const double* b, int iterations, double* x);

/* |A^T (b - A x)|_2 / |A^T b|_2 in FP64: the relative normal-equations
 * residual every fit in this project is stopped on and reported against.
 * Optionally also returns |b - A x|_2 and |A^T b|_2. */
double or_rel_normal_residual(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                              const double* b, const double* x,
                              double* res_norm, double* atb_norm);

/* ---- input side: Board/PossibleMoves.cpp:86-156, Board/Flips.cpp:119-182,
 * Board/Position.cpp:84-102, Game/GameClass.cpp:4-44 ---------------------- */
uint64_t or_possible_moves(uint64_t P, uint64_t O);
uint64_t or_flips(uint64_t P, uint64_t O, int move);  /* loop-based rays */
/* positions reached from the start position after exactly `plies` moves
 * (passes are implicit and do not count; finished games drop out) */
uint64_t or_perft(int plies);
/* BoardTest/PositionGenerator.cpp & Board/PositionGenerator.cpp:22-32:
 * RandomPositionGenerator(seed): n positions from std::mt19937_64 */
void or_random_positions_mt(uint32_t seed, int64_t n, uint64_t* out_PO);
void or_random_positions_mt_dist(uint32_t seed, int64_t n, uint64_t* out_PO);
/* random legal play, per-game splitmix64 stream (the project's documented
 * generator; see cassandra_amd/csrc/board.hpp for the definition) */
void or_random_play(uint64_t seed, int64_t n, int e_lo, int e_hi, uint64_t* out_PO);
/* synthetic targets: x_true ~ sx * IH12, b = clip(2*round((A x_true + sn*IH12)/2)) */
void or_make_targets(const uint32_t* idx, int64_t rows, int k, int64_t cols,
                     uint64_t seed_x, uint64_t seed_n, double sx, double sn,
                     double* x_true, double* b);

/* ---- Pattern/Evaluator.cpp:11-73 : expanded ScoreEstimator ------------- */
/* Expand compact weights (cols) into per-variation tables; tables are
 * written back to back, table i has 3^popcount(variation_i) floats.
 * Returns the total number of floats (call with out=NULL to size). */
int64_t or_expand_weights(const or_group*, const float* w, float* out);
/* ScoreEstimator::Eval over n positions (Evaluator.cpp:48-54) */
void    or_eval_expanded(const or_group*, const float* tables, const uint64_t* PO,
                         int64_t n, float* out);
/* ScoreEstimator::Weights (Evaluator.cpp:56-73): compact from tables */
void    or_compact_weights(const or_group*, const float* tables, float* w);

#ifdef __cplusplus
}
#endif
