"""Writes tests/golden/sphere_ata_plus_a.npz.

Input: tests/golden/sphere.mtx, the 642 x 642 matrix the reference's tutorial loads
(examples/Tutorial/main.cpp:36-37, data/sphere.mtx).  Output: the pattern of B = A^T A + A
(Eigen semantics: structural product, union sum) and every value computed exactly in rational
arithmetic from the double inputs and rounded once to double, so the fixture holds the
correctly rounded B independent of any summation order.

Run: python tests/golden/make_golden.py
"""
import os
from fractions import Fraction

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_mtx(path):
    with open(path) as f:
        f.readline()
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        nr, nc, nnz = (int(t) for t in line.split())
        rows, cols, vals = [], [], []
        for _ in range(nnz):
            r, c, v = f.readline().split()
            rows.append(int(r) - 1)
            cols.append(int(c) - 1)
            vals.append(float(v))
    return nc, rows, cols, vals


def main():
    n, rows, cols, vals = load_mtx(os.path.join(HERE, "sphere.mtx"))
    A = {}
    for r, c, v in zip(rows, cols, vals):
        A[(r, c)] = A.get((r, c), Fraction(0)) + Fraction(v)
    colsets = {}
    for (r, c), v in A.items():
        colsets.setdefault(c, {})[r] = v
    rowsets = {}
    for (r, c), v in A.items():
        rowsets.setdefault(r, {})[c] = v
    B = {}
    # (A^T A)(i, j) = sum_k A(k, i) A(k, j); structural: every (i, j) sharing a row k
    for k, row in rowsets.items():
        items = list(row.items())
        for i, aki in items:
            for j, akj in items:
                B[(i, j)] = B.get((i, j), Fraction(0)) + aki * akj
    for (r, c), v in A.items():
        B[(r, c)] = B.get((r, c), Fraction(0)) + v
    keys = sorted(B.keys(), key=lambda rc: (rc[1], rc[0]))
    indptr = np.zeros(n + 1, np.int64)
    for (r, c) in keys:
        indptr[c + 1] += 1
    indptr = np.cumsum(indptr).astype(np.int32)
    indices = np.array([r for (r, c) in keys], np.int32)
    values = np.array([float(B[k]) for k in keys])
    np.savez_compressed(os.path.join(HERE, "sphere_ata_plus_a.npz"), indptr=indptr, indices=indices, values=values)
    print("wrote", len(values), "values")


if __name__ == "__main__":
    main()
