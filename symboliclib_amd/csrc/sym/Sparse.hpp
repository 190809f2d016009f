// Minimal compressed-sparse-column and dense column-major containers that stand in for the
// Eigen types the reference's front-ends use (Eigen::SparseMatrix, Eigen::Matrix).
//
// Pattern semantics follow Eigen 3.3.7 as used by the reference:
//  * CSC, column-major, inner (row) indices sorted and unique within a column;
//  * setFromTriplets sums duplicate (row, col) entries in insertion order with `+`;
//  * the sparse product is structural: every (i, j) reached by some A(i,k) B(k,j) is stored,
//    whatever its value (no pruning of zeros), and the sparse sum takes the union pattern.
// The scalar type may be double or Sym::Symbolic; with Symbolic the same code records the
// expression of every entry.
#pragma once

#include <algorithm>
#include <cassert>
#include <cstddef>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace Sym {

template <class T>
struct Triplet {
    int r = 0, c = 0;
    T v{};
    Triplet() = default;
    Triplet(int r_, int c_, const T& v_) : r(r_), c(c_), v(v_) {}
    int row() const { return r; }
    int col() const { return c; }
    const T& value() const { return v; }
};

template <class T>
class DenseMatrix {
    int rows_ = 0, cols_ = 0;
    std::vector<T> d_;

public:
    DenseMatrix() = default;
    DenseMatrix(int r, int c) : rows_(r), cols_(c), d_((size_t)r * c) {}
    DenseMatrix(int r, int c, const T* src) : rows_(r), cols_(c), d_(src, src + (size_t)r * c) {}
    void resize(int r, int c) {
        rows_ = r;
        cols_ = c;
        d_.assign((size_t)r * c, T{});
    }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    size_t size() const { return d_.size(); }
    T* data() { return d_.data(); }
    const T* data() const { return d_.data(); }
    T& operator()(int r, int c) { return d_[(size_t)r + (size_t)c * rows_]; }
    const T& operator()(int r, int c) const { return d_[(size_t)r + (size_t)c * rows_]; }
};

template <class T>
class SparseMatrix {
    int rows_ = 0, cols_ = 0;
    std::vector<int> outer_{0};
    std::vector<int> inner_;
    std::vector<T> values_;

public:
    SparseMatrix() = default;
    SparseMatrix(int r, int c) { resize(r, c); }

    // Build from raw CSC arrays (the caller guarantees sorted, unique rows per column).
    SparseMatrix(int r, int c, const int* outer, const int* inner, const T* values) : rows_(r), cols_(c) {
        outer_.assign(outer, outer + c + 1);
        inner_.assign(inner, inner + outer[c]);
        if (values) values_.assign(values, values + outer[c]);
        else values_.assign(outer[c], T{});
    }

    void resize(int r, int c) {
        rows_ = r;
        cols_ = c;
        outer_.assign((size_t)c + 1, 0);
        inner_.clear();
        values_.clear();
    }

    int rows() const { return rows_; }
    int cols() const { return cols_; }
    size_t nonZeros() const { return inner_.size(); }
    size_t size() const { return values_.size(); }

    T* valuePtr() { return values_.data(); }
    const T* valuePtr() const { return values_.data(); }
    T* data() { return values_.data(); }
    const T* data() const { return values_.data(); }
    int* outerIndexPtr() { return outer_.data(); }
    const int* outerIndexPtr() const { return outer_.data(); }
    int* innerIndexPtr() { return inner_.data(); }
    const int* innerIndexPtr() const { return inner_.data(); }
    const std::vector<int>& outer() const { return outer_; }
    const std::vector<int>& inner() const { return inner_; }
    std::vector<T>& values() { return values_; }
    const std::vector<T>& values() const { return values_; }

    // Sum duplicates in insertion order; sort rows within each column.
    template <class It>
    void setFromTriplets(It begin, It end) {
        const size_t m = (size_t)std::distance(begin, end);
        std::vector<const Triplet<T>*> ts;
        ts.reserve(m);
        for (It it = begin; it != end; ++it) {
            if (it->row() < 0 || it->row() >= rows_ || it->col() < 0 || it->col() >= cols_)
                throw std::out_of_range("setFromTriplets: index out of range");
            ts.push_back(&*it);
        }
        std::stable_sort(ts.begin(), ts.end(), [](const Triplet<T>* a, const Triplet<T>* b) {
            return a->col() != b->col() ? a->col() < b->col() : a->row() < b->row();
        });
        outer_.assign((size_t)cols_ + 1, 0);
        inner_.clear();
        values_.clear();
        inner_.reserve(m);
        values_.reserve(m);
        for (size_t k = 0; k < ts.size();) {
            size_t e = k + 1;
            T v = ts[k]->value();
            while (e < ts.size() && ts[e]->col() == ts[k]->col() && ts[e]->row() == ts[k]->row()) {
                v = v + ts[e]->value();
                ++e;
            }
            inner_.push_back(ts[k]->row());
            values_.push_back(v);
            ++outer_[(size_t)ts[k]->col() + 1];
            k = e;
        }
        for (int j = 0; j < cols_; ++j) outer_[(size_t)j + 1] += outer_[j];
    }

    void setFromTriplets(const std::vector<Triplet<T>>& t) { setFromTriplets(t.begin(), t.end()); }

    SparseMatrix<T> transpose() const {
        SparseMatrix<T> R(cols_, rows_);
        std::vector<int> cnt((size_t)rows_ + 1, 0);
        for (int k : inner_) ++cnt[(size_t)k + 1];
        for (int i = 0; i < rows_; ++i) cnt[(size_t)i + 1] += cnt[i];
        R.outer_ = cnt;
        R.inner_.resize(inner_.size());
        R.values_.resize(values_.size());
        for (int j = 0; j < cols_; ++j)
            for (int p = outer_[j]; p < outer_[(size_t)j + 1]; ++p) {
                const int q = cnt[inner_[p]]++;
                R.inner_[q] = j;
                R.values_[q] = values_[p];
            }
        return R;
    }

    template <class U>
    SparseMatrix<U> cast() const {
        SparseMatrix<U> R(rows_, cols_);
        R.setPattern(outer_, inner_);
        for (size_t k = 0; k < values_.size(); ++k) R.values()[k] = U(values_[k]);
        return R;
    }

    // The pattern alone, without values (valuePtr() is then null): output patterns.
    void setStructure(std::vector<int> outer, std::vector<int> inner) {
        outer_ = std::move(outer);
        inner_ = std::move(inner);
        values_.clear();
        values_.shrink_to_fit();
    }
    std::vector<int> releaseInner() { return std::move(inner_); }

    void setPattern(const std::vector<int>& outer, const std::vector<int>& inner) {
        outer_ = outer;
        inner_ = inner;
        values_.assign(inner.size(), T{});
    }

    // Structural product: accumulates A(i,k) * B(k,j) over k in B's column order.
    friend SparseMatrix<T> operator*(const SparseMatrix<T>& A, const SparseMatrix<T>& B) {
        if (A.cols_ != B.rows_) throw std::invalid_argument("sparse product: size mismatch");
        SparseMatrix<T> R(A.rows_, B.cols_);
        std::vector<int> mark((size_t)A.rows_, -1);
        std::vector<T> acc((size_t)A.rows_);
        std::vector<int> rowsInCol;
        for (int j = 0; j < B.cols_; ++j) {
            rowsInCol.clear();
            for (int pb = B.outer_[j]; pb < B.outer_[(size_t)j + 1]; ++pb) {
                const int k = B.inner_[pb];
                const T& bkj = B.values_[pb];
                for (int pa = A.outer_[k]; pa < A.outer_[(size_t)k + 1]; ++pa) {
                    const int i = A.inner_[pa];
                    if (mark[i] != j) {
                        mark[i] = j;
                        acc[i] = A.values_[pa] * bkj;
                        rowsInCol.push_back(i);
                    } else {
                        acc[i] = acc[i] + A.values_[pa] * bkj;
                    }
                }
            }
            std::sort(rowsInCol.begin(), rowsInCol.end());
            for (int i : rowsInCol) {
                R.inner_.push_back(i);
                R.values_.push_back(acc[i]);
                acc[i] = T{};
            }
            R.outer_[(size_t)j + 1] = (int)R.inner_.size();
        }
        return R;
    }

    // Union-pattern sum: A(i,j) + B(i,j) where both are stored, else the stored one.
    friend SparseMatrix<T> operator+(const SparseMatrix<T>& A, const SparseMatrix<T>& B) {
        if (A.rows_ != B.rows_ || A.cols_ != B.cols_) throw std::invalid_argument("sparse sum: size mismatch");
        SparseMatrix<T> R(A.rows_, A.cols_);
        for (int j = 0; j < A.cols_; ++j) {
            int pa = A.outer_[j], ea = A.outer_[(size_t)j + 1];
            int pb = B.outer_[j], eb = B.outer_[(size_t)j + 1];
            while (pa < ea || pb < eb) {
                if (pb >= eb || (pa < ea && A.inner_[pa] < B.inner_[pb])) {
                    R.inner_.push_back(A.inner_[pa]);
                    R.values_.push_back(A.values_[pa++]);
                } else if (pa >= ea || B.inner_[pb] < A.inner_[pa]) {
                    R.inner_.push_back(B.inner_[pb]);
                    R.values_.push_back(B.values_[pb++]);
                } else {
                    R.inner_.push_back(A.inner_[pa]);
                    R.values_.push_back(A.values_[pa++] + B.values_[pb++]);
                }
            }
            R.outer_[(size_t)j + 1] = (int)R.inner_.size();
        }
        return R;
    }
};

}  // namespace Sym
