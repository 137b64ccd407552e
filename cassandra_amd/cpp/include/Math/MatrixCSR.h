// Drop-in for Math/MatrixCSR.h (Math/MatrixCSR.h:11-33): fixed number of
// entries per row, implicit value 1, repeated columns allowed.  The indices
// live on the host exactly as in the reference (Row(i) spans); the products run
// on the GPU through libcassandra_amd, which keeps a device copy that is
// re-uploaded after the rows were changed through the mutable Row().
#pragma once
#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <vector>

#include "Vector.h"

struct cad_matrix;

class MatrixCSR
{
    std::size_t elements_per_row;
    std::size_t cols;
    std::vector<uint32_t> col_indices;
    mutable std::shared_ptr<cad_matrix> device;  // device copy (null: stale)

    friend Vector JacobiPreconditionerOfATA(const MatrixCSR&);
    friend Vector operator*(const MatrixCSR&, const Vector&);
    friend Vector operator*(const Vector&, const MatrixCSR&);

public:
    MatrixCSR(std::size_t elements_per_row, std::size_t rows, std::size_t cols);

    std::size_t Rows() const noexcept { return elements_per_row ? col_indices.size() / elements_per_row : 0; }
    std::size_t Cols() const noexcept { return cols; }

    std::span<uint32_t> Row(std::size_t index) noexcept;
    std::span<const uint32_t> Row(std::size_t index) const noexcept;

    // The device matrix (uploaded on first use / after a mutable Row()).
    cad_matrix* Device() const;
};

// Jacobi Preconditioner of A' * A: returns 1 / diag(A' * A) (inf for empty columns)
Vector JacobiPreconditionerOfATA(const MatrixCSR&);

Vector operator*(const MatrixCSR&, const Vector&);
Vector operator*(const Vector&, const MatrixCSR&);
