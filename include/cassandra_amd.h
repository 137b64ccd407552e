/*
 * cassandra_amd: C-ABI of the MI355X (gfx950) pattern-evaluator fit.
 *
 * The reference (dominichofer/Cassandra) has no FFI on this path; its boundary
 * is its own C++ API (Math/MatrixCSR.h, Math/Solver.h, Pattern/Indexer.h,
 * Pattern/Evaluator.h, PatternFit/PatternFit.cpp).  Each entry point below
 * names the reference interface it replaces.  The C++ drop-in layer
 * (cassandra_amd/cpp) and the Python torch.autograd.Function both bind here.
 *
 * Conventions
 *  - Every function returns CAD_OK (0) or a negative error code; the message of
 *    the last failure on the calling thread is cad_last_error().  No C++
 *    exception crosses this boundary.
 *  - Pointers named d_* are device pointers (caller-owned, on the context's
 *    device); h_* are host pointers.  Handles own their device memory.
 *  - Work is enqueued on the context's stream (its own non-blocking stream by
 *    default, or one set with cad_context_set_stream).  Functions that return
 *    results to the host synchronise that stream.
 *  - Positions are 16 bytes {uint64 P, uint64 O} (Board/Position.h:11-33):
 *    P = side to move; bit b = square (x, y) with b = x + 8y.
 */
#ifndef CASSANDRA_AMD_H
#define CASSANDRA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAD_OK 0
#define CAD_ERR_ARGUMENT (-1)       /* bad argument (reference: std::runtime_error) */
#define CAD_ERR_SIZE_MISMATCH (-2)  /* reference: std::runtime_error("Size mismatch.") */
#define CAD_ERR_PATTERN (-3)        /* reference: indexer precondition failures */
#define CAD_ERR_DEVICE (-4)         /* HIP runtime error */
#define CAD_ERR_NOMEM (-5)          /* device allocation failed */
#define CAD_ERR_BREAKDOWN (-6)      /* solver breakdown (NaN / zero curvature) */

typedef struct cad_context cad_context;
typedef struct cad_patterns cad_patterns;
typedef struct cad_matrix cad_matrix;
typedef struct cad_evaluator cad_evaluator;

/* ---- errors, context ---------------------------------------------------- */
const char* cad_last_error(void);
const char* cad_version(void);
/* Creates a context on `device` with its own default-priority stream. */
int cad_context_create(int device, cad_context** out);
int cad_context_destroy(cad_context* ctx);
/* Enqueue subsequent work on `hip_stream` (a hipStream_t, e.g. torch's
 * current stream; NULL is the null stream).  Solvers capture graphs only on a
 * non-null stream. */
int cad_context_set_stream(cad_context* ctx, void* hip_stream);
/* Go back to the context's own stream. */
int cad_context_reset_stream(cad_context* ctx);
/* The stream work is currently enqueued on (a hipStream_t). */
int cad_context_get_stream(cad_context* ctx, void** out);
int cad_context_synchronize(cad_context* ctx);
/* Device memory helpers for callers without HIP headers (the C++ layer). */
int cad_malloc(cad_context* ctx, size_t bytes, void** d_out);
int cad_free(cad_context* ctx, void* d_ptr);
int cad_memcpy_h2d(cad_context* ctx, void* d_dst, const void* h_src, size_t bytes);
int cad_memcpy_d2h(cad_context* ctx, void* h_dst, const void* d_src, size_t bytes);
int cad_memcpy_d2d(cad_context* ctx, void* d_dst, const void* d_src, size_t bytes);

/* ---- patterns (Pattern/Indexer.h:10-40, Pattern/Indexer.cpp:319-422) ----
 * cad_patterns_create = GroupIndexer(masks): CreateIndexer per mask, columns
 * laid out pattern after pattern.  Fails with CAD_ERR_PATTERN where the
 * reference indexer constructors throw.  ctx may be NULL: a host-only group
 * (info, variations and cad_patterns_indices_host work; device calls fail). */
int cad_patterns_create(cad_context* ctx, const uint64_t* h_masks, int n, cad_patterns** out);
int cad_patterns_destroy(cad_patterns* p);
/* k = Variations().size() (elements per row), cols = index_space_size */
int cad_patterns_info(const cad_patterns* p, int* k, int64_t* cols, int* npatterns);
/* per pattern: class (0 A, 1 V, 2 D, 3 VH, 4 DC, 5 VHDC), transformed mask
 * (Indexer::pattern), index_space_size, symmetry_group_order, column offset */
int cad_patterns_pattern_info(const cad_patterns* p, int i, int* cls, uint64_t* mask,
                              int64_t* size, int* order, int64_t* offset);
/* GroupIndexer::Variations (Pattern/Indexer.cpp:394-403): k masks */
int cad_patterns_variations(const cad_patterns* p, uint64_t* h_out);
/* GroupIndexer::Indices on the host for n positions (row-major n x k). */
int cad_patterns_indices_host(const cad_patterns* p, const uint64_t* h_positions, int64_t n,
                              uint32_t* h_out);

/* ---- index extraction (GroupIndexer::InsertIndices over rows,
 * PatternFit/PatternFit.cpp:19-21) --------------------------------------- */
#define CAD_LAYOUT_ROW_MAJOR 0  /* out[r*k + s]: MatrixCSR::Row(r) layout, int/uint32 */
#define CAD_LAYOUT_SLOT_MAJOR 1 /* out[s*n + r] */
/* d_out receives uint32 column indices (int32 for torch: same bits). */
int cad_index_extract(cad_patterns* p, const uint64_t* d_positions, int64_t n, uint32_t* d_out,
                      int layout);

/* ---- matrix (Math/MatrixCSR.h:11-33) ------------------------------------- */
/* Builds the device matrix of GroupIndexer rows for n device positions. */
int cad_matrix_from_positions(cad_patterns* p, const uint64_t* d_positions, int64_t n,
                              cad_matrix** out);
/* A block-diagonal batch of independent systems, one per empty-count stage
 * (PatternFit/PatternFit.cpp:41-84 fits the stages one FitWeights call at a
 * time; Pattern/Evaluator.cpp:77-88 selects a stage's weights by EmptyCount).
 * Rows [row_begin[s], row_begin[s+1]) of the device positions form system s
 * (host array of nsys + 1 entries, row_begin[0] = 0, every system >= 1 row);
 * system s owns columns [s * cols, (s + 1) * cols) of x, cols being the pattern
 * group's column count.  Every product and solver runs all systems in the same
 * launches; solvers keep per-system scalars and stop each system on its own. */
int cad_matrix_from_positions_systems(cad_patterns* p, const uint64_t* d_positions,
                                      const int64_t* row_begin, int nsys, cad_matrix** out);
/* The same matrix without stored indices (matrix-free): the library keeps a
 * copy of the n device positions (16 bytes a row) and recomputes every row's
 * indices with the pattern programs inside each product, where
 * GroupIndexer::InsertIndices (Pattern/Indexer.cpp:413-422) would have filled
 * MatrixCSR::Row(i) once.  Every product, solver and cad_matrix_download works
 * on it; cad_matrix_info reports index_bytes 0. */
int cad_matrix_from_positions_mf(cad_patterns* p, const uint64_t* d_positions, int64_t n,
                                 cad_matrix** out);
/* No counterpart in the reference (its products have one fixed OpenMP
 * schedule).  Times the candidate schedules of this matrix's products -- rows
 * per thread of the LDS tile gather or the row kernel, scatter tiles per CU
 * and the cap on partial-slice traffic -- with HIP events on the context's
 * stream, keeps the fastest and drops the solvers' captured graphs.  The
 * timings and the choice go to log (may be NULL).  A x does not change (every
 * schedule sums each row in the same order); A^T y stays exact on integer data
 * and equal to rounding otherwise. */
int cad_matrix_autotune(cad_matrix* m, char* log, size_t log_size);
/* No counterpart in the reference.  The schedule of a matrix's products (what
 * cad_matrix_autotune chooses): reading it, or setting one explicitly (the
 * solvers' captured graphs are dropped).  Rows per thread of the FP32 tile
 * gather 2, 4 or 8 (8 for u16 indices only), of the FP64 one 1 or 2, 0 for
 * the row kernel; threads per gather block 1024, or 512 with two blocks per CU
 * (stored indices only); sub-tiles of rpt x block rows sharing each LDS
 * staging 1..3 (2 and 3 only for FP32 with rpt 2 and FP64 with rpt 1, full
 * blocks and stored indices); scatter
 * rows per step 4, or 8 for u16 indices;
 * scatter order 0 (chunks per window) or 1 (chunks shared by the windows,
 * XCD-major); scatter tiles per CU > 0, partial-traffic cap in (0, 1] and
 * balance in [0, 1].
 * Every schedule computes the same products (see cad_matrix_autotune);
 * CAD_ERR_ARGUMENT for a value outside these or a tile gather the matrix
 * cannot use. */
/* No counterpart in the reference.  Average launch time in microseconds of
 * each FP32 product kernel of this matrix with its current schedule, timed
 * with HIP events on the context's stream over reps launches after a warm
 * one (inputs: all ones): us[0] the A x gather, us[1] the A^T y window tiles,
 * us[2] their column reduction, us[3] the max |y| pre-pass of the fixed-point
 * windows (0 where not used), us[4] the first pass of the Jacobi counts
 * (JacobiPreconditionerOfATA).  The first n of them are written. */
int cad_matrix_kernel_times(cad_matrix* m, int reps, double* us, int n);
typedef struct cad_schedule {
    int gather_f32_rpt, gather_f32_block, gather_f32_sub;
    int gather_f64_rpt, gather_f64_block, gather_f64_sub;
    int scatter_rpt, scatter_order;
    double scatter_tiles_per_cu, scatter_partial_cap;
    double scatter_balance; /* order 0: weight of slot loads against atomics in a window's work, in [0, 1] */
} cad_schedule;
int cad_matrix_get_schedule(const cad_matrix* m, cad_schedule* out);
int cad_matrix_set_schedule(cad_matrix* m, const cad_schedule* schedule);
/* nsys (1 for an ordinary matrix), columns per system and, when row_begin is
 * not NULL, the nsys + 1 row boundaries. */
int cad_matrix_systems(const cad_matrix* m, int* nsys, int64_t* cols_per_system, int64_t* row_begin);
/* MatrixCSR(k, rows, cols) with Row(i) filled: row-major host indices. */
int cad_matrix_create(cad_context* ctx, const uint32_t* h_idx, int64_t rows, int k, int64_t cols,
                      cad_matrix** out);
int cad_matrix_destroy(cad_matrix* m);
/* index_bytes: 2 or 4 (device storage of one entry) */
int cad_matrix_info(const cad_matrix* m, int64_t* rows, int* k, int64_t* cols, int* index_bytes);
/* MatrixCSR::Row(i) for all rows, row-major into h_out[rows*k]. */
int cad_matrix_download(const cad_matrix* m, uint32_t* h_out);
/* operator*(const MatrixCSR&, const Vector&) (Math/MatrixCSR.cpp:49-65):
 * FP32 y = A x summed sequentially in row order (bit-exact with the reference). */
int cad_spmv_f32(cad_matrix* m, const float* d_x, float* d_y);
int cad_spmv_f64(cad_matrix* m, const double* d_x, double* d_y);
/* operator*(const Vector&, const MatrixCSR&) (Math/MatrixCSR.cpp:67-87):
 * z = A^T y (scatter-add; order of FP additions not deterministic). */
int cad_spmtv_f32(cad_matrix* m, const float* d_y, float* d_z);
int cad_spmtv_f64(cad_matrix* m, const double* d_y, double* d_z);
/* JacobiPreconditionerOfATA (Math/MatrixCSR.cpp:21-47): 1/diag(A^T A) in float
 * (inf for empty columns). */
int cad_jacobi(cad_matrix* m, float* d_out);

/* ---- solvers (Math/Solver.h:9-287, PatternFit/PatternFit.cpp:7-39) -------- */
#define CAD_SOLVER_LSQR 0
#define CAD_SOLVER_PCG 1   /* PCG on A^T A with Jacobi (FitWeights) */
#define CAD_SOLVER_CG 2    /* CG on A^T A */
#define CAD_SOLVER_LSMR 3

#define CAD_PREC_FP64 0    /* FP64 products and vectors; LSQR stops on the device at
                              tol / 3 and restarts from the FP64 residual if the
                              explicit metric is not met */
#define CAD_PREC_FP32 1    /* FP32 products and vectors (the reference's precision) */
#define CAD_PREC_MIXED 2   /* FP32 products, FP64 scalars/dots/norms/x and FP64
                              refinement residual r = b - A x, A^T r (LSQR only);
                              systems FP32 cannot resolve finish in FP64 */

typedef struct cad_solve_options {
    int solver;           /* CAD_SOLVER_* */
    int precision;        /* CAD_PREC_* */
    int precondition;     /* LSQR: column scaling by 1/|a_c|; PCG: Jacobi (always) */
    int max_iterations;   /* total (inner) iterations */
    double tol;           /* stop when |A^T(b-Ax)| / |A^T b| <= tol (0: run max_iterations) */
    double inner_tol;     /* MIXED: inner LSQR reduction per refinement step (raised to 1e-7, the FP32 floor) */
    int max_refinements;  /* MIXED */
    int check_every;      /* iterations between host convergence checks (graph size) */
    int use_graph;        /* capture check_every iterations in a hipGraph */
    int verbose;
} cad_solve_options;

typedef struct cad_solve_report {
    int iterations;             /* inner iterations run */
    int refinements;            /* outer refinement steps (MIXED) */
    int converged;
    int explicit_checks;
    double rel_normal_residual; /* |A^T(b-Ax)| / |A^T b|, FP64, at return */
    double residual_norm;       /* |b - A x| */
    double seconds;             /* wall time of the solve (device-synchronised) */
} cad_solve_report;

void cad_solve_options_default(cad_solve_options* o);
/* Solves min |A x - b| (LSQR/LSMR) or A^T A x = A^T b (PCG/CG) from x0 = 0.
 * d_b: rows doubles; d_x: cols doubles (output). */
int cad_solve(cad_matrix* m, const double* d_b, double* d_x, const cad_solve_options* opt,
              cad_solve_report* rep);
/* cad_solve with the per-system outcome of a batch (host arrays of nsys
 * entries, either may be NULL): the FP64 metric of each system at return and
 * the iteration at which it converged (or the total).  rep holds the worst
 * system's metric. */
int cad_solve_systems(cad_matrix* m, const double* d_b, double* d_x, const cad_solve_options* opt,
                      cad_solve_report* rep, double* rel_per_system, int* iterations_per_system);
/* |A^T(b - A x)| / |A^T b| in FP64 on the device (the stopping metric). */
int cad_rel_normal_residual(cad_matrix* m, const double* d_b, const double* d_x, double* out,
                            double* residual_norm);

/* ---- differentiable evaluator (Pattern/Evaluator.cpp:48-54) --------------
 * Forward: out[r] = sum_s w[idx[r*k+s]] (row-major int32 indices, as built by
 * cad_index_extract with CAD_LAYOUT_ROW_MAJOR).  Backward: grad_w = A^T g,
 * the scatter-add adjoint (grad_w is overwritten).  The patterns handle gives
 * the column block of every slot. */
int cad_eval_forward_f32(cad_patterns* p, const int32_t* d_idx, int64_t B, const float* d_w,
                         float* d_out);
int cad_eval_forward_f64(cad_patterns* p, const int32_t* d_idx, int64_t B, const double* d_w,
                         double* d_out);
int cad_eval_backward_f32(cad_patterns* p, const int32_t* d_idx, int64_t B, const float* d_g,
                          float* d_grad_w);
int cad_eval_backward_f64(cad_patterns* p, const int32_t* d_idx, int64_t B, const double* d_g,
                          double* d_grad_w);
/* No counterpart in the reference.  cad_eval_backward_f32 for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or
 * allocating routine.  Its partial sums live in a static 32 MiB device array
 * per GPU (held until the process ends, its address taken when the context is
 * created); the pattern group may have at most 4,194,304 columns
 * (CAD_ERR_ARGUMENT otherwise).  Same result as cad_eval_backward_f32 to
 * rounding (FP64 sums, exact on integer g).  Calls that overlap on one GPU
 * share that array and race: serialise them (one stream, or events). */
int cad_eval_backward_f32_noalloc(cad_patterns* p, const int32_t* d_idx, int64_t B, const float* d_g,
                                  float* d_grad_w);
/* No counterpart in the reference (its fit is the least-squares PCG of
 * FitWeights, PatternFit/PatternFit.cpp:7-39); for training the evaluator
 * itself.  One Adam step of the Huber-loss fit on the minibatch rows
 * d_sel[0..B) of the row-major index matrix d_idx [N, k] with targets
 * d_target [N]: pred = the forward above, loss = mean over the batch of
 * huber(pred - target, delta), then the gradient (the backward above, rows
 * read through d_sel) and torch.optim.Adam's update (no weight decay) of
 * d_w with its moments d_m, d_v ([cols] floats, zero before step 1); step
 * counts from 1.  Adds the batch's mean loss to *d_loss (one device double).
 * Device buffers for the per-row gradient, the weight gradient and a
 * contiguous copy of the minibatch's index rows are kept on the pattern group
 * (grown on first use). */
int cad_eval_train_step_f32(cad_patterns* p, const int32_t* d_idx, const float* d_target,
                            const int64_t* d_sel, int64_t B, float delta, float lr, float beta1,
                            float beta2, float eps, int64_t step, float* d_w, float* d_m, float* d_v,
                            double* d_loss);
/* No counterpart in the reference.  An epoch of cad_eval_train_step_f32: the
 * minibatches d_order[b0 .. b0 + batch) for b0 = 0, batch, ... < n (the last
 * one shorter), Adam steps step0, step0 + 1, ...; every step's mean loss is
 * added to *d_loss.  All steps are enqueued by this one call. */
int cad_eval_train_epoch_f32(cad_patterns* p, const int32_t* d_idx, const float* d_target,
                             const int64_t* d_order, int64_t n, int64_t batch, float delta, float lr,
                             float beta1, float beta2, float eps, int64_t step0, float* d_w, float* d_m,
                             float* d_v, double* d_loss);

/* ---- expanded tables (Pattern/Evaluator.cpp:11-29,56-73) ----------------
 * The evaluator's form of the weights: one table of 3^n floats per variation
 * (GroupIndexer::Variations order), table i at offset sum_{j<i} 3^n_j. */
/* Floats of all tables of a pattern group (the sum of 3^n over the variations). */
int cad_patterns_expanded_size(cad_patterns* p, int64_t* n);
/* ScoreEstimator(pattern, weights) (Pattern/Evaluator.cpp:11-29): table i gets
 * w_i[FastIndex(config, variation_i)] = w[Indices(config)[i]] for every
 * configuration of variation i's mask.  d_w: cols floats, d_tables:
 * cad_patterns_expanded_size floats. */
int cad_weights_expand(cad_patterns* p, const float* d_w, float* d_tables);
/* ScoreEstimator::Weights (Pattern/Evaluator.cpp:56-73): per pattern,
 * w[Index(config)] = table 0 at FastIndex(config, pattern) for every
 * configuration of the pattern's mask, the last one in Configurations order
 * winning; synchronises the context's stream. */
int cad_weights_compact(cad_patterns* p, const float* d_tables, float* d_w);

/* ---- resident evaluator (ScoreEstimator::Eval, Pattern/Evaluator.cpp:48-54;
 * the batched form of the leaf evaluations of
 * Search/PrincipalVariation.cpp:224-229) --------------------------------------
 * Holds the expanded tables of one compact weight vector in device memory,
 * so that batches of positions are scored without re-uploading anything.
 * Every score is the float sum over the variations in order from 0, the same
 * bits as ScoreEstimator::Eval.  The pattern group must outlive it. */
int cad_evaluator_create(cad_patterns* p, const float* d_w, cad_evaluator** out);
int cad_evaluator_destroy(cad_evaluator* e);
/* New compact weights (d_w: cols floats), expanded on the device. */
int cad_evaluator_set_weights(cad_evaluator* e, const float* d_w);
/* d_out[r] = Eval(position r) for n device positions (enqueued on the stream). */
int cad_evaluator_eval(cad_evaluator* e, const uint64_t* d_positions, int64_t n, float* d_out);
/* The same for host positions and scores (staged through buffers the
 * evaluator keeps; synchronises the stream). */
int cad_evaluator_eval_host(cad_evaluator* e, const uint64_t* h_positions, int64_t n, float* h_out);

/* ---- synthetic inputs (Game/GameGenerator.cpp:3-28, Game/Player.cpp:3-16) */
/* n positions from uniformly random legal play (empties drawn in [e_lo,e_hi]) */
int cad_random_play(cad_context* ctx, uint64_t seed, int64_t n, int e_lo, int e_hi,
                    uint64_t* d_out);
/* Host versions of the same board code (Board/BitBoard.cpp:4-84,
 * Board/PossibleMoves.cpp:86-156, Board/Flips.cpp:119-182): sym 1 vertical,
 * 2 horizontal, 3 diagonal, 4 codiagonal. */
uint64_t cad_flip(int sym, uint64_t b);
uint64_t cad_possible_moves(uint64_t P, uint64_t O);
uint64_t cad_flips(uint64_t P, uint64_t O, int move);
int cad_random_play_host(uint64_t seed, int64_t n, int e_lo, int e_hi, uint64_t* h_out);
/* x_true ~ sx * N(0,1) (cols) and b = clip(2 round((A x_true + sn N(0,1))/2), +-64) */
int cad_make_targets(cad_matrix* m, uint64_t seed_x, uint64_t seed_n, double sx, double sn,
                     double* d_xtrue, double* d_b);

#ifdef __cplusplus
}
#endif
#endif /* CASSANDRA_AMD_H */
