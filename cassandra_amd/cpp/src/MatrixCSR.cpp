// MatrixCSR on the host (Math/MatrixCSR.cpp:5-19) with its products on the GPU
// (Math/MatrixCSR.cpp:21-87): each product uploads the vector, runs the HIP
// kernel and returns a fresh host Vector, as the reference's value semantics.
#include "Math/MatrixCSR.h"

#include <stdexcept>

#include "Math/DeviceSolver.h"
#include "device.h"

using namespace cad_cpp;

MatrixCSR::MatrixCSR(std::size_t elements_per_row, std::size_t rows, std::size_t cols)
    : elements_per_row(elements_per_row), cols(cols), col_indices(rows * elements_per_row)
{
}

std::span<uint32_t> MatrixCSR::Row(std::size_t row) noexcept
{
    device.reset();  // the caller may write: the device copy is stale
    return std::span<uint32_t>(col_indices.data() + row * elements_per_row, elements_per_row);
}

std::span<const uint32_t> MatrixCSR::Row(std::size_t row) const noexcept
{
    return std::span<const uint32_t>(col_indices.data() + row * elements_per_row, elements_per_row);
}

cad_matrix* MatrixCSR::Device() const
{
    if (!device) {
        cad_matrix* m = nullptr;
        Check(cad_matrix_create(Context(), col_indices.data(), int64_t(Rows()), int(elements_per_row),
                                int64_t(cols), &m));
        device.reset(m, [](cad_matrix* p) { cad_matrix_destroy(p); });
    }
    return device.get();
}

Vector JacobiPreconditionerOfATA(const MatrixCSR& m)
{
    Vector out(m.Cols());
    DevBuf d(sizeof(float) * out.size());
    Check(cad_jacobi(m.Device(), d.as<float>()));
    Check(cad_memcpy_d2h(Context(), out.data(), d.as<float>(), sizeof(float) * out.size()));
    return out;
}

Vector operator*(const MatrixCSR& mat, const Vector& x)
{
    if (x.size() != mat.Cols()) throw std::runtime_error("Size mismatch.");
    Vector y(mat.Rows());
    if (y.empty()) return y;
    DevBuf dx(sizeof(float) * x.size()), dy(sizeof(float) * y.size());
    Check(cad_memcpy_h2d(Context(), dx.as<float>(), x.data(), sizeof(float) * x.size()));
    Check(cad_spmv_f32(mat.Device(), dx.as<float>(), dy.as<float>()));
    Check(cad_memcpy_d2h(Context(), y.data(), dy.as<float>(), sizeof(float) * y.size()));
    return y;
}

Vector operator*(const Vector& x, const MatrixCSR& mat)
{
    if (x.size() != mat.Rows()) throw std::runtime_error("Size mismatch.");
    Vector z(mat.Cols());
    DevBuf dx(sizeof(float) * std::max<std::size_t>(1, x.size())), dz(sizeof(float) * z.size());
    Check(cad_memcpy_h2d(Context(), dx.as<float>(), x.data(), sizeof(float) * x.size()));
    Check(cad_spmtv_f32(mat.Device(), dx.as<float>(), dz.as<float>()));
    Check(cad_memcpy_d2h(Context(), z.data(), dz.as<float>(), sizeof(float) * z.size()));
    return z;
}

std::vector<double> SolveOnDevice(const MatrixCSR& A, const std::vector<double>& b, const DeviceSolveOptions& o,
                                  DeviceSolveReport* report)
{
    if (b.size() != A.Rows()) throw std::runtime_error("Size mismatch.");
    cad_solve_options opt;
    cad_solve_options_default(&opt);
    opt.solver = int(o.method);
    opt.precision = int(o.precision);
    opt.precondition = o.precondition ? 1 : 0;
    opt.max_iterations = o.max_iterations;
    opt.tol = o.tol;
    opt.inner_tol = o.inner_tol;
    opt.check_every = o.check_every;
    DevBuf db(sizeof(double) * std::max<std::size_t>(1, b.size())), dx(sizeof(double) * A.Cols());
    Check(cad_memcpy_h2d(Context(), db.as<double>(), b.data(), sizeof(double) * b.size()));
    cad_solve_report rep;
    Check(cad_solve(A.Device(), db.as<double>(), dx.as<double>(), &opt, &rep));
    std::vector<double> x(A.Cols());
    Check(cad_memcpy_d2h(Context(), x.data(), dx.as<double>(), sizeof(double) * x.size()));
    if (report) {
        report->iterations = rep.iterations;
        report->refinements = rep.refinements;
        report->converged = rep.converged != 0;
        report->rel_normal_residual = rep.rel_normal_residual;
        report->residual_norm = rep.residual_norm;
        report->seconds = rep.seconds;
    }
    return x;
}
