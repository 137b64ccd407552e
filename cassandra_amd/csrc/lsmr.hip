// LSMR (Math/Solver.h:219-287; Fong & Saunders 2011, damp = 0) on the device
// in FP64: vectors and products on the GPU, the scalar recurrence on the host
// (one synchronisation per iteration).  Single system only.
#include <cmath>
#include <vector>

#include "kernels.hpp"
#include "solver.hpp"

namespace cad {
namespace {

constexpr int kBS = 256;

// y = a x + b y
__global__ void k_axpby(int64_t n, double a, const double* x, double b, double* y)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        y[i] = a * x[i] + b * y[i];
}

// per-block partial sums of x^2
__global__ __launch_bounds__(kBS) void k_sumsq(int64_t n, const double* __restrict__ x, double* __restrict__ part)
{
    __shared__ double sh[kBS / 64];
    double acc = 0.0;
    for (int64_t i = blockIdx.x * int64_t(kBS) + threadIdx.x; i < n; i += int64_t(gridDim.x) * kBS)
        acc += x[i] * x[i];
    acc = block_sum<kBS>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

struct Dev {
    hipStream_t s;
    DBuf part;
    std::vector<double> hpart;
    int nb = 1024;
    explicit Dev(hipStream_t st) : s(st), part(sizeof(double) * 1024), hpart(1024) {}

    void axpby(int64_t n, double a, const double* x, double b, double* y)
    {
        int g = int(std::min<int64_t>(4096, (n + kBS - 1) / kBS));
        hipLaunchKernelGGL(k_axpby, dim3(std::max(g, 1)), dim3(kBS), 0, s, n, a, x, b, y);
        CAD_HIP(hipGetLastError());
    }
    double norm(int64_t n, const double* x)
    {
        hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(kBS), 0, s, n, x, part.as<double>());
        CAD_HIP(hipGetLastError());
        CAD_HIP(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
        CAD_HIP(hipStreamSynchronize(s));
        double t = 0;
        for (double v : hpart) t += v;
        return std::sqrt(t);
    }
};

void sym_ortho(double a, double b, double& c, double& s, double& r)
{
    auto sgn = [](double v) { return v < 0 ? -1.0 : (v > 0 ? 1.0 : 0.0); };
    if (b == 0) { c = sgn(a); s = 0; r = std::fabs(a); return; }
    if (a == 0) { c = 0; s = sgn(b); r = std::fabs(b); return; }
    if (std::fabs(b) > std::fabs(a)) {
        const double tau = a / b;
        s = sgn(b) / std::sqrt(1 + tau * tau);
        c = s * tau;
        r = b / s;
    } else {
        const double tau = b / a;
        c = sgn(a) / std::sqrt(1 + tau * tau);
        s = c * tau;
        r = a / c;
    }
}

}  // namespace

void lsmr_solve(cad_matrix* m, const double* b, double* x, const cad_solve_options& o, cad_solve_report& rep)
{
    if (m->nsys != 1) throw Error(CAD_ERR_ARGUMENT, "LSMR supports a single system");
    hipStream_t s = m->ctx->stream;
    const int64_t R = m->rows, C = m->cols;
    DBuf u(sizeof(double) * R), t(sizeof(double) * R), v(sizeof(double) * C), z(sizeof(double) * C),
        h(sizeof(double) * C), hb(sizeof(double) * C);
    Dev d(s);
    CAD_HIP(hipMemcpyAsync(u.p, b, sizeof(double) * R, hipMemcpyDeviceToDevice, s));
    double beta = d.norm(R, u.as<double>());
    if (beta > 0) d.axpby(R, 0.0, u.as<double>(), 1.0 / beta, u.as<double>());
    matrix_spmtv_f64(m, u.as<double>(), v.as<double>());
    double alpha = d.norm(C, v.as<double>());
    if (alpha > 0) d.axpby(C, 0.0, v.as<double>(), 1.0 / alpha, v.as<double>());
    double alphabar = alpha, zetabar = alpha * beta, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
    CAD_HIP(hipMemcpyAsync(h.p, v.p, sizeof(double) * C, hipMemcpyDeviceToDevice, s));
    CAD_HIP(hipMemsetAsync(hb.p, 0, sizeof(double) * C, s));
    CAD_HIP(hipMemsetAsync(x, 0, sizeof(double) * C, s));
    int it = 0;
    for (; it < o.max_iterations && alpha * beta != 0; it++) {
        const double rho_old = rho, rhobar_old = rhobar;
        matrix_spmv_f64(m, v.as<double>(), t.as<double>());
        d.axpby(R, 1.0, t.as<double>(), -alpha, u.as<double>());
        beta = d.norm(R, u.as<double>());
        if (beta > 0) d.axpby(R, 0.0, u.as<double>(), 1.0 / beta, u.as<double>());
        matrix_spmtv_f64(m, u.as<double>(), z.as<double>());
        d.axpby(C, 1.0, z.as<double>(), -beta, v.as<double>());
        alpha = d.norm(C, v.as<double>());
        if (alpha > 0) d.axpby(C, 0.0, v.as<double>(), 1.0 / alpha, v.as<double>());

        double c, sn;
        sym_ortho(alphabar, beta, c, sn, rho);
        const double theta_new = sn * alpha;
        alphabar = c * alpha;
        const double thetabar = sbar * rho;
        sym_ortho(cbar * rho, theta_new, cbar, sbar, rhobar);
        const double zeta = cbar * zetabar;
        zetabar = -sbar * zetabar;
        // hbar = h - (thetabar rho / (rho_old rhobar_old)) hbar ; x += zeta/(rho rhobar) hbar
        d.axpby(C, 1.0, h.as<double>(), -thetabar * rho / (rho_old * rhobar_old), hb.as<double>());
        d.axpby(C, zeta / (rho * rhobar), hb.as<double>(), 1.0, x);
        // h = v - (theta_new / rho) h
        d.axpby(C, 1.0, v.as<double>(), -theta_new / rho, h.as<double>());
    }
    rep.iterations = it;
    rep.rel_normal_residual = rel_normal_residual(m, b, x, &rep.residual_norm);
    rep.converged = o.tol > 0 && rep.rel_normal_residual <= o.tol;
}

}  // namespace cad
