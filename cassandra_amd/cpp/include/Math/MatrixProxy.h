// Drop-in for Math/MatrixProxy.h: lazy transpose and product of two matrices.
// Unlike the reference, a proxy keeps proxies it is built from by value (and
// real matrices by reference), so `transposed(A) * A` stays valid after the
// full expression that created it -- the reference's PCG keeps a reference to
// that temporary (PatternFit/PatternFit.cpp:31, Math/MatrixProxy.h:48-49).
#pragma once
#include <cstddef>
#include <stdexcept>
#include <type_traits>

#include "Vector.h"

template <typename T>
struct TransposedMatrix
{
    const T& m;
    explicit TransposedMatrix(const T& m) noexcept : m(m) {}
    std::size_t Rows() const noexcept { return m.Cols(); }
    std::size_t Cols() const noexcept { return m.Rows(); }
};

template <typename T> struct is_matrix_proxy : std::false_type {};
template <typename T> struct is_matrix_proxy<TransposedMatrix<T>> : std::true_type {};

// Proxies are stored by value, everything else by reference.
template <typename T>
using proxy_storage = std::conditional_t<is_matrix_proxy<T>::value, const T, const T&>;

template <typename T>
Vector operator*(const TransposedMatrix<T>& mat, const Vector& x) { return x * mat.m; }

template <typename T>
Vector operator*(const Vector& x, const TransposedMatrix<T>& mat) { return mat.m * x; }

template <typename T>
auto transposed(const T& m) { return TransposedMatrix<T>(m); }

template <typename T>
const T& transposed(const TransposedMatrix<T>& m) { return m.m; }

template <typename L, typename R>
struct TwoMatrix
{
    proxy_storage<L> l;
    proxy_storage<R> r;

    TwoMatrix(const L& l, const R& r) : l(l), r(r)
    {
        if (l.Cols() != r.Rows())
            throw std::runtime_error("Size mismatch.");
    }
    std::size_t Rows() const noexcept { return l.Rows(); }
    std::size_t Cols() const noexcept { return r.Cols(); }
};

template <typename L, typename R> struct is_matrix_proxy<TwoMatrix<L, R>> : std::true_type {};

template <typename L, typename R>
Vector operator*(const TwoMatrix<L, R>& mat, const Vector& x) { return mat.l * (mat.r * x); }

template <typename L, typename R>
Vector operator*(const Vector& x, const TwoMatrix<L, R>& mat) { return (x * mat.l) * mat.r; }

template <typename L, typename R>
    requires requires(L l, R r) { l.Rows(); l.Cols(); r.Rows(); r.Cols(); }
auto operator*(const L& l, const R& r) { return TwoMatrix<L, R>(l, r); }
