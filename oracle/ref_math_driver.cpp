// Thin extern "C" driver around the reference's own Math sources
// (Math/Vector.cpp, Math/MatrixCSR.cpp, Math/Solver.h, Math/MatrixProxy.h),
// compiled unmodified from the reference checkout by oracle/Makefile into
// oracle/_ref/libref_math.so.  TEST INFRASTRUCTURE ONLY: used by the tests to
// pin the oracle's CSR products / solvers against the real reference code, and
// by bench.py as the "reference" CPU baseline.  This file is project code; it
// contains no reference source, it only calls the reference's public API.
//
// <cmath> and <tuple> are included first because Math/Solver.h uses
// std::isnormal / std::sqrt / std::tie without including them itself.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <tuple>

#include "Math/MatrixCSR.h"
#include "Math/MatrixProxy.h"
#include "Math/Solver.h"
#include "Math/Vector.h"


extern "C" {

// Opaque handle so that timing loops do not include the copy into MatrixCSR.
void* ref_csr_create(const uint32_t* idx, int64_t rows, int k, int64_t cols)
{
    auto* m = new MatrixCSR(size_t(k), size_t(rows), size_t(cols));
    for (int64_t r = 0; r < rows; r++) {
        auto row = m->Row(size_t(r));
        std::memcpy(row.data(), idx + r * k, sizeof(uint32_t) * size_t(k));
    }
    return m;
}
void ref_csr_free(void* m) { delete static_cast<MatrixCSR*>(m); }

int ref_spmv(void* m, const float* x, int64_t cols, float* y)
{
    try {
        Vector xv(x, x + cols);
        Vector r = *static_cast<MatrixCSR*>(m) * xv;
        std::memcpy(y, r.data(), sizeof(float) * r.size());
        return 0;
    } catch (...) { return -1; }
}

int ref_spmtv(void* m, const float* y, int64_t rows, float* z)
{
    try {
        Vector yv(y, y + rows);
        Vector r = yv * *static_cast<MatrixCSR*>(m);
        std::memcpy(z, r.data(), sizeof(float) * r.size());
        return 0;
    } catch (...) { return -1; }
}

void ref_jacobi(void* m, float* d)
{
    Vector r = JacobiPreconditionerOfATA(*static_cast<MatrixCSR*>(m));
    std::memcpy(d, r.data(), sizeof(float) * r.size());
}

// FitWeights' solve (PatternFit/PatternFit.cpp:53-56) on a prepared matrix.
// The reference writes `PCG solver(transposed(matrix) * matrix, ...)`: the
// TransposedMatrix and TwoMatrix proxies are temporaries that die at the end
// of that statement while PCG (Math/Solver.h:83) and TwoMatrix
// (Math/MatrixProxy.h:168-169) keep references to them, so Iterate() reads
// dangling memory.  Here the proxies are named so that they outlive the solver;
// the arithmetic is the reference's.
void ref_fit_pcg(void* mp, const float* b, int64_t rows, int iterations, float* x,
                 float* resid_hist)
{
    const MatrixCSR& m = *static_cast<MatrixCSR*>(mp);
    Vector score(b, b + rows);
    Vector weights(m.Cols(), 0.0f);
    DiagonalPreconditioner P(JacobiPreconditionerOfATA(m));
    auto At = transposed(m);
    auto AtA = At * m;
    PCG solver(AtA, P, weights, At * score);
    for (int i = 0; i < iterations; i++) {
        solver.Iterate(1);
        if (resid_hist) resid_hist[i] = solver.Residuum();
    }
    Vector r = solver.X();
    std::memcpy(x, r.data(), sizeof(float) * r.size());
}

void ref_lsqr(void* mp, const float* b, int64_t rows, int iterations, float* x,
              float* phibar_hist)
{
    const MatrixCSR& m = *static_cast<MatrixCSR*>(mp);
    Vector bv(b, b + rows);
    LSQR solver(m, Vector(m.Cols(), 0.0f), bv);
    for (int i = 0; i < iterations; i++) {
        solver.Iterate(1);
        if (phibar_hist) phibar_hist[i] = solver.Residuum();
    }
    Vector r = solver.X();
    std::memcpy(x, r.data(), sizeof(float) * r.size());
}

void ref_lsmr(void* mp, const float* b, int64_t rows, int iterations, float* x)
{
    const MatrixCSR& m = *static_cast<MatrixCSR*>(mp);
    Vector bv(b, b + rows);
    LSMR solver(m, Vector(m.Cols(), 0.0f), bv);
    solver.Iterate(iterations);
    Vector r = solver.X();
    std::memcpy(x, r.data(), sizeof(float) * r.size());
}

// Math/Vector.cpp operations used by the MathTest/Vector.cpp known answers.
float ref_dot(const float* a, const float* b, int64_t n)
{
    return dot(Vector(a, a + n), Vector(b, b + n));
}
float ref_norm(const float* a, int64_t n) { return norm(Vector(a, a + n)); }
float ref_sum(const float* a, int64_t n) { return sum(Vector(a, a + n)); }
float ref_l1(const float* a, int64_t n) { return L1_norm(Vector(a, a + n)); }
int ref_dot_size_mismatch_throws(void)
{
    try {
        dot(Vector(2, 1.0f), Vector(3, 1.0f));
    } catch (const std::runtime_error&) {
        return 1;
    }
    return 0;
}

}  // extern "C"
