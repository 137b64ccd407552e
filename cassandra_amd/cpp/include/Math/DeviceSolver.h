// Device-resident solves of a MatrixCSR system (no counterpart in the
// reference, which iterates on the host): the whole fit stays in HBM, with
// FP64 scalars and, for LSQR, the mixed-precision refined variant.
#pragma once
#include <string>
#include <vector>

#include "MatrixCSR.h"

struct DeviceSolveOptions
{
    enum Method { LSQR = 0, PCG = 1, CG = 2, LSMR = 3 } method = LSQR;
    enum Precision { FP64 = 0, FP32 = 1, Mixed = 2 } precision = Mixed;
    bool precondition = true;     // LSQR: column scaling; PCG: Jacobi
    int max_iterations = 100000;
    double tol = 1e-10;           // |A^T(b - A x)| / |A^T b|; 0 runs max_iterations
    double inner_tol = 1e-5;
    int check_every = 10;
};

struct DeviceSolveReport
{
    int iterations = 0, refinements = 0;
    bool converged = false;
    double rel_normal_residual = 0, residual_norm = 0, seconds = 0;
};

// min |A x - b| (LSQR, LSMR) or A^T A x = A^T b (PCG, CG) from x0 = 0.
std::vector<double> SolveOnDevice(const MatrixCSR&, const std::vector<double>& b, const DeviceSolveOptions& = {},
                                  DeviceSolveReport* = nullptr);
