// Device-resident solvers and products on a cad_matrix.
#pragma once
#include "../../include/cassandra_amd.h"
#include "runtime.hpp"

#include <string>

namespace cad {

// The work lists of an LSQR iteration restricted to some systems of a batch.
// Once most systems of a batch have finished, the iterations of the rest launch
// only their own blocks (a block of a finished system would still be
// scheduled, read its flag and exit).  Each list keeps the system order of
// the full list; systems outside `on` have no blocks.
struct ActiveLists {
    std::vector<uint8_t> on;      // per system: in the lists
    int64_t rows = 0;             // rows of those systems
    int ntiles = 0, nscatter = 0, nreduce = 0, ncols = 0;
    bool row_kernel = false;      // the gather blocks are row blocks of the generic row kernel
    DBuf tiles, tfirst, tcount;   // gather blocks and each system's range of them
    DBuf gpart;                   // the row kernel's block partials
    DBuf scatter;                 // A^T y window tiles, planned for these systems alone
    DBuf part;                    // their partial slices
    int cbits = 0;                // their fixed-point bound (ScatterLists::cbits)
    DBuf reduce, rfirst, rcount;  // column reduction items and each system's range
    DBuf cols;                    // column blocks
    int gen = 0;                  // bumped on every rebuild (captured graphs are keyed on it)
};

// Solver buffers cached on a matrix (one set per solver / precision), with
// the captured iteration graph.
struct Workspace {
    int key = 0;
    DBuf u, r64, v, w, y, x64, xprev, D, z64, vd, z, scal, scal_chk, hb, lsmr;
    DBuf P, p, q, zz, xr, rr, t, bvt, pscal, jac;
    int vd_prec = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_iters = 0, graph_prec = 0;
    int graph_gen = -1;       // ActiveLists generation the graph was captured with (-1: full lists)
    ActiveLists active;
    bool use_active = false;  // iterations launch the active lists
    void* host_scal = nullptr;
    ~Workspace();
};

void matrix_spmv_f32(cad_matrix* m, const float* x, float* y);
void matrix_spmv_f64(cad_matrix* m, const double* x, double* y);
void matrix_spmtv_f32(cad_matrix* m, const float* y, float* z);
void matrix_spmtv_f64(cad_matrix* m, const double* y, double* z);
void matrix_jacobi(cad_matrix* m, float* d);
const uint32_t* jacobi_counts(cad_matrix* m);
// Per-system outcome of a solve on a block-diagonal batch: the FP64 metric
// |A_s^T(b_s - A_s x_s)| / |A_s^T b_s| at return and the iteration count at
// which the system was frozen (converged) or the total.
struct SystemDetail {
    std::vector<double> rel;
    std::vector<int> iters;
};

void solve(cad_matrix* m, const double* b, double* x, const cad_solve_options& o, cad_solve_report& rep,
           SystemDetail* detail = nullptr);
double rel_normal_residual(cad_matrix* m, const double* b, const double* x, double* rnorm);
void matrix_targets(cad_matrix* m, const double* xt, uint64_t seed, double sn, double* b);
// Times the candidate schedules of the FP32 products on this matrix and
// keeps the fastest (rebuilds the plans; drops captured solver graphs).
std::string autotune(cad_matrix* m);
// Average launch times (us) of the FP32 product kernels, HIP events on the
// context's stream: A x, the A^T y window tiles, their reduction, the max |y|
// pre-pass.
void kernel_times(cad_matrix* m, int reps, double* us, int n);
// Sets the schedule (validated; drops the captured graphs, rebuilds the plans).
void set_schedule(cad_matrix* m, const TuneParams& tp);

}  // namespace cad
