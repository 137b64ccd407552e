// Drop-in for Math/Solver.h (Math/Solver.h:9-287): the iterative solvers with
// the reference's interface.  They work on host Vectors with float scalars as
// the reference does; every product with a MatrixCSR (and through proxies,
// A^T A) runs on the GPU.  For fits that stay on the device end to end, with
// FP64 recurrences and refinement, use SolveOnDevice (Math/DeviceSolver.h).
#pragma once
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <tuple>
#include <utility>

#include "MatrixProxy.h"
#include "Vector.h"

class IterativeSolver
{
public:
    virtual ~IterativeSolver() = default;
    virtual void Iterate(int n) = 0;
    virtual float Residuum() const = 0;
    virtual Vector Error() const = 0;
    virtual Vector X() const = 0;
};

class Preconditioner
{
public:
    virtual ~Preconditioner() = default;
    virtual Vector apply(const Vector&) const = 0;
    virtual Vector revert(const Vector&) const = 0;
};

// Non-normal entries (0, inf, nan, subnormal) are replaced by 1 (Solver.h:29-34).
class DiagonalPreconditioner : public Preconditioner
{
public:
    Vector d;
    explicit DiagonalPreconditioner(Vector diag) : d(std::move(diag))
    {
        for (float& x : d)
            if (!std::isnormal(x)) x = 1.0f;
    }
    Vector apply(const Vector& x) const override { return elementwise_multiplication(d, x); }
    Vector revert(const Vector& x) const override { return elementwise_division(d, x); }
};

// Conjugate gradients on a symmetric positive definite A (Solver.h:40-76).
template <typename Matrix>
class CG final : public IterativeSolver
{
    proxy_storage<Matrix> A;
    Vector x, b, p, r;

public:
    CG(const Matrix& A_, Vector x0, Vector b_) : A(A_), x(std::move(x0)), b(std::move(b_))
    {
        if (A.Cols() != A.Rows() || A.Cols() != x.size() || A.Rows() != b.size())
            throw std::runtime_error("Size mismatch.");
        r = Error();
        p = r;
    }
    void Iterate(int n = 1) override
    {
        for (int k = 0; k < n; k++) {
            const float rr = dot(r, r);
            const Vector Ap = A * p;
            const float alpha = rr / dot(p, Ap);
            x += alpha * p;
            r -= alpha * Ap;
            const float beta = dot(r, r) / rr;
            p = r + beta * p;
        }
    }
    float Residuum() const override { return dot(r, r); }
    Vector Error() const override { return b - A * x; }
    Vector X() const override { return x; }
};

// Preconditioned conjugate gradients (Solver.h:78-117).
template <typename Matrix>
class PCG final : public IterativeSolver
{
    proxy_storage<Matrix> A;
    const Preconditioner& P;
    Vector b, x, z, r, p;

public:
    PCG(const Matrix& A_, const Preconditioner& P_, Vector x0, Vector b_)
        : A(A_), P(P_), b(std::move(b_)), x(std::move(x0))
    {
        if (A.Rows() != A.Cols() || A.Cols() != x.size() || A.Rows() != b.size())
            throw std::runtime_error("Size mismatch.");
        r = Error();
        z = P.apply(r);
        p = z;
    }
    void Iterate(int n = 1) override
    {
        for (int k = 0; k < n; k++) {
            const float rz = dot(r, z);
            const Vector Ap = A * p;
            const float alpha = rz / dot(p, Ap);
            x += alpha * p;
            r -= alpha * Ap;
            z = P.apply(r);
            const float beta = dot(r, z) / rz;
            p = z + beta * p;
        }
    }
    float Residuum() const override { return dot(r, r); }
    Vector Error() const override { return b - A * x; }
    Vector X() const override { return x; }
};

// (|v|, v / |v|), or (0, 0) for the zero vector (Solver.h:119-126).
inline std::pair<float, Vector> decompose(const Vector& v)
{
    const float length = norm(v);
    if (length == 0) return {0.0f, Vector(v.size(), 0.0f)};
    return {length, v / length};
}

// LSQR (Paige & Saunders 1982) from x0 = 0, as Solver.h:128-182 (the start
// ignores x0; b is kept by value).
template <typename Matrix>
class LSQR : public IterativeSolver
{
    proxy_storage<Matrix> A;
    Vector b, x, u, v, w;
    float alpha = 0, phi_bar = 0, rho_bar = 0;

public:
    LSQR(const Matrix& A_, Vector x0, Vector b_) : A(A_), b(std::move(b_)), x(std::move(x0))
    {
        if (A.Cols() != x.size() || A.Rows() != b.size()) throw std::runtime_error("Size mismatch.");
        float beta;
        std::tie(beta, u) = decompose(b);
        std::tie(alpha, v) = decompose(transposed(A) * u);
        w = v;
        phi_bar = beta;
        rho_bar = alpha;
    }
    void Iterate(int n = 1) override
    {
        for (int k = 0; k < n; k++) {
            float beta;
            std::tie(beta, u) = decompose(A * v - alpha * u);
            std::tie(alpha, v) = decompose(transposed(A) * u - beta * v);
            const float rho = std::sqrt(rho_bar * rho_bar + beta * beta);
            const float c = rho_bar / rho, s = beta / rho;
            const float theta = s * alpha;
            rho_bar = -c * alpha;
            const float phi = c * phi_bar;
            phi_bar = s * phi_bar;
            x += (phi / rho) * w;
            w = v - (theta / rho) * w;
        }
    }
    float Residuum() const override { return phi_bar; }
    Vector Error() const override { return b - A * x; }
    Vector X() const override { return x; }
};

inline float signum(float v) { return v < 0 ? -1.0f : (v > 0 ? 1.0f : 0.0f); }

// Symmetric Givens rotation: (c, s, r) with c a + s b = r (Solver.h:193-217).
inline std::tuple<float, float, float> GivensRotation(float a, float b)
{
    if (b == 0) return {signum(a), 0.0f, std::abs(a)};
    if (a == 0) return {0.0f, signum(b), std::abs(b)};
    if (std::abs(b) > std::abs(a)) {
        const float tau = a / b;
        const float s = signum(b) / std::sqrt(1 + tau * tau);
        return {s * tau, s, b / s};
    }
    const float tau = b / a;
    const float c = signum(a) / std::sqrt(1 + tau * tau);
    return {c, c * tau, a / c};
}

// LSMR (Fong & Saunders 2011, damp = 0) as Solver.h:219-287.
template <typename Matrix>
class LSMR : public IterativeSolver
{
    proxy_storage<Matrix> A;
    Vector b, x, v, u, h, h_bar;
    float alpha = 0, alpha_bar = 0, zeta_bar = 0, c_bar = 1, s_bar = 0, rho = 1, rho_bar = 1;

public:
    LSMR(const Matrix& A_, Vector x0, Vector b_) : A(A_), b(std::move(b_)), x(std::move(x0))
    {
        if (A.Cols() != x.size() || A.Rows() != b.size()) throw std::runtime_error("Size mismatch.");
        float beta;
        std::tie(beta, u) = decompose(b);
        std::tie(alpha, v) = decompose(transposed(A) * u);
        alpha_bar = alpha;
        zeta_bar = alpha * beta;
        h = v;
        h_bar = Vector(h.size(), 0.0f);
    }
    void Iterate(int n = 1) override
    {
        for (int k = 0; k < n; k++) {
            const float rho_old = rho, rho_bar_old = rho_bar;
            float beta;
            std::tie(beta, u) = decompose(A * v - alpha * u);
            std::tie(alpha, v) = decompose(transposed(A) * u - beta * v);
            auto [c, s, r] = GivensRotation(alpha_bar, beta);
            rho = r;
            const float theta_new = s * alpha;
            alpha_bar = c * alpha;
            const float theta_bar = s_bar * rho;
            std::tie(c_bar, s_bar, rho_bar) = GivensRotation(c_bar * rho, theta_new);
            const float zeta = c_bar * zeta_bar;
            zeta_bar *= -s_bar;
            h_bar = h - (theta_bar * rho / rho_old / rho_bar_old) * h_bar;
            x += zeta / rho / rho_bar * h_bar;
            h = v - theta_new / rho * h;
        }
    }
    float Residuum() const override { return 0; }
    Vector Error() const override { return b - A * x; }
    Vector X() const override { return x; }
};
