// Math/Vector.cpp:7-167 semantics: element-wise float operations, float
// reductions, std::runtime_error("Size mismatch") on size mismatch.
#include "Math/Vector.h"

#include <cmath>
#include <cstdint>
#include <stdexcept>

namespace {
void same_size(const Vector& a, const Vector& b)
{
    if (a.size() != b.size()) throw std::runtime_error("Size mismatch");
}
}  // namespace

void operator+=(Vector& l, const Vector& o)
{
    same_size(l, o);
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) l[i] += o[i];
}

void operator-=(Vector& l, const Vector& o)
{
    same_size(l, o);
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) l[i] -= o[i];
}

void operator*=(Vector& l, float f)
{
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) l[i] *= f;
}

void operator/=(Vector& l, float f)
{
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) l[i] /= f;
}

Vector operator+(Vector l, const Vector& r) { l += r; return l; }
Vector operator+(const Vector& l, Vector&& r) { r += l; return std::move(r); }
Vector operator-(Vector l, const Vector& r) { l -= r; return l; }

Vector operator-(const Vector& l, Vector&& r)
{
    same_size(l, r);
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) r[i] = l[i] - r[i];
    return std::move(r);
}

Vector operator*(Vector v, float f) { v *= f; return v; }
Vector operator*(float f, Vector v) { v *= f; return v; }
Vector operator/(Vector v, float f) { v /= f; return v; }

Vector elementwise_multiplication(Vector l, const Vector& o)
{
    same_size(l, o);
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) l[i] *= o[i];
    return l;
}

Vector elementwise_division(Vector l, const Vector& o)
{
    same_size(l, o);
    const int64_t n = int64_t(l.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) l[i] /= o[i];
    return l;
}

Vector inv(Vector v)
{
    const int64_t n = int64_t(v.size());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) v[i] = 1.0f / v[i];
    return v;
}

float dot(const Vector& l, const Vector& r)
{
    same_size(l, r);
    const int64_t n = int64_t(l.size());
    float s = 0.0f;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; i++) s += l[i] * r[i];
    return s;
}

float norm(const Vector& x) { return std::sqrt(dot(x, x)); }
float L2_norm(const Vector& x) { return norm(x); }

float L1_norm(const Vector& x)
{
    const int64_t n = int64_t(x.size());
    float s = 0.0f;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; i++) s += std::abs(x[i]);
    return s;
}

float sum(const Vector& x)
{
    const int64_t n = int64_t(x.size());
    float s = 0.0f;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < n; i++) s += x[i];
    return s;
}

std::string to_string(const Vector& v)
{
    std::string s = "(";
    for (std::size_t i = 0; i < v.size(); i++) s += (i ? "," : "") + std::to_string(v[i]);
    return s + ")";
}
