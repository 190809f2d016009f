// Uniform (pointer, length) access to the containers ComputeUnit accepts, plus helpers that
// turn numeric containers into symbolic ones (src/ContainerSupport.h:7-90 of the reference).
#pragma once

#include <tuple>
#include <vector>

#include "Sparse.hpp"
#include "Symbolic.hpp"

namespace Sym {

template <class Cont>
auto contPtr(Cont& c) { return c.data(); }
template <class Cont>
size_t contSize(const Cont& c) { return c.size(); }
template <class Cont>
auto contPtr(const Cont& c) { return c.data(); }

inline Symbolic* contPtr(Symbolic& c) { return &c; }
inline const Symbolic* contPtr(const Symbolic& c) { return &c; }
inline size_t contSize(const Symbolic&) { return 1; }

template <class T>
T* contPtr(SparseMatrix<T>& c) { return c.valuePtr(); }
template <class T>
const T* contPtr(const SparseMatrix<T>& c) { return c.valuePtr(); }
template <class T>
size_t contSize(const SparseMatrix<T>& c) { return c.nonZeros(); }

// Copy of A's pattern whose k-th stored value is the variable {k, objId}.
template <class T>
SparseMatrix<Symbolic> makeSymbolic(const SparseMatrix<T>& A, const int objId) {
    SparseMatrix<Symbolic> B(A.rows(), A.cols());
    B.setPattern(A.outer(), A.inner());
    for (size_t i = 0; i < B.nonZeros(); ++i) B.valuePtr()[i] = Symbolic((int)i, objId);
    return B;
}

// Dense column-major matrix of variables {k, objId}, k = row + col * rows.
template <class T>
DenseMatrix<Symbolic> makeSymbolic(const DenseMatrix<T>& A, const int objId) {
    DenseMatrix<Symbolic> B(A.rows(), A.cols());
    for (size_t i = 0; i < B.size(); ++i) B.data()[i] = Symbolic((int)i, objId);
    return B;
}

// n x n sparse matrix from (row, col, value) entries, summing entries that meet at one
// position.  With `mirror`, each off-diagonal entry is entered a second time at (col, row);
// the mirrored entries come after all the given ones (so a position sums its given entries
// first).
template <class T>
void toSparseMatrix(SparseMatrix<T>& A, const std::vector<std::tuple<int, int, T>>& entries, const size_t n,
                    const bool mirror = false) {
    std::vector<Triplet<T>> all;
    all.reserve(entries.size() * (mirror ? 2 : 1));
    for (int pass = 0; pass < (mirror ? 2 : 1); ++pass)
        for (const auto& e : entries) {
            const int r = std::get<0>(e), c = std::get<1>(e);
            if (pass == 0) all.emplace_back(r, c, std::get<2>(e));
            else if (r != c) all.emplace_back(c, r, std::get<2>(e));
        }
    A.resize((int)n, (int)n);
    A.setFromTriplets(all.begin(), all.end());
}

}  // namespace Sym
