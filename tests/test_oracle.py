"""Pins of the FP64 oracle (CPU only).

Each check is chosen so that a plausible mistake in oracle/oracle.c (a dropped term, a wrong
sign, a transposed index, a wrong vertex order) fails it:
  * A^T A + A: exact rational values on the reference's data/sphere.mtx (golden fixture),
    scipy structural pattern, exact rationals on a random matrix;
  * cotangent Laplacian: closed form on a regular grid (cot 45 = 1, cot 90 = 0), zero row sums,
    symmetry, invariance to a rigid motion;
  * Baraff-Witkin: Hessian = finite difference of the gradient, gradient = finite difference
    of the energy, symmetry, translation invariance;
  * bending: Q e = 0 for constants and for coordinates of a flat quad (discrete quadratic
    bending vanishes on planar deformations).
"""
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from symboliclib_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_sphere_golden_exact():
    n, cp, ri, a = W.load_mtx(os.path.join(GOLDEN, "sphere.mtx"))
    g = np.load(os.path.join(GOLDEN, "sphere_ata_plus_a.npz"))
    bcp, bri, bv, sc = O.ata_plus_a(cp, ri, a)
    assert np.array_equal(bcp, g["indptr"]) and np.array_equal(bri, g["indices"])
    err = O.rel_errors(bv, g["values"], sc)
    assert err.max() <= 1e-14, err.max()


def test_ata_plus_a_structural_pattern_vs_scipy():
    n, cp, ri, _ = W.c1_matrix(subdiv=3)
    ones = sp.csc_matrix((np.ones(len(ri)), ri, cp), shape=(n, n))
    B = (ones.T @ ones + ones).tocsc()  # positive values: no cancellation, pattern = structure
    B.sort_indices()
    bcp, bri, _, _ = O.ata_plus_a(cp, ri, np.ones(len(ri)))
    assert np.array_equal(bcp, B.indptr) and np.array_equal(bri, B.indices)


def test_ata_plus_a_keeps_structural_zeros():
    # A = [[1, 1], [1, -1]] has (A^T A)(0,1) = 1*1 + 1*(-1) = 0: Eigen keeps the entry
    cp = np.array([0, 2, 4], np.int32)
    ri = np.array([0, 1, 0, 1], np.int32)
    bcp, bri, bv, _ = O.ata_plus_a(cp, ri, np.array([1.0, 1.0, 1.0, -1.0]))
    assert list(bcp) == [0, 2, 4] and list(bri) == [0, 1, 0, 1]
    assert np.allclose(bv, [3.0, 1.0, 1.0, 1.0])  # A^T A = [[2, 0], [0, 2]]


def test_ata_plus_a_exact_random():
    rng = np.random.default_rng(7)
    n = 30
    M = sp.random(n, n, density=0.15, random_state=rng, format="csc") + sp.eye(n, format="csc")
    M.sort_indices()
    cp, ri, a = M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64)
    bcp, bri, bv, _ = O.ata_plus_a(cp, ri, a)
    dense = [[Fraction(0)] * n for _ in range(n)]
    for j in range(n):
        for p in range(cp[j], cp[j + 1]):
            dense[ri[p]][j] = Fraction(float(a[p]))
    for j in range(n):
        for p in range(bcp[j], bcp[j + 1]):
            i = bri[p]
            exact = sum((dense[k][i] * dense[k][j] for k in range(n)), Fraction(0)) + dense[i][j]
            assert abs(bv[p] - float(exact)) <= 1e-14 * (1 + abs(float(exact)))


def _dense(cp, ri, v, n):
    return sp.csc_matrix((v, ri, cp), shape=(n, n)).toarray()


def test_cotan_closed_form_regular_grid():
    side = 6
    V, F = W.grid_mesh(side, side)
    cp, ri, vals, _ = O.cotan(V, F)
    L = _dense(cp, ri, vals, len(V))
    assert np.allclose(L, L.T, atol=1e-14)
    assert np.allclose(L.sum(axis=1), 0.0, atol=1e-13)

    def vid(i, j):
        return i + j * side

    # interior axis edge: two 45-degree opposite corners -> 1 + 1; boundary axis edge -> 1;
    # diagonal edge: two right-angle opposite corners -> 0 (entry stored, value zero)
    assert L[vid(2, 2), vid(3, 2)] == pytest.approx(2.0, abs=1e-14)
    assert L[vid(0, 2), vid(0, 3)] == pytest.approx(1.0, abs=1e-14)
    i, j = 2, 2  # (i + j) even: diagonal v00-v11
    assert L[vid(i, j), vid(i + 1, j + 1)] == pytest.approx(0.0, abs=1e-14)
    k = np.searchsorted(ri[cp[vid(i + 1, j + 1)]:cp[vid(i + 1, j + 1) + 1]], vid(i, j))
    assert ri[cp[vid(i + 1, j + 1)] + k] == vid(i, j)
    # interior vertex: four axis neighbours with weight 2
    assert L[vid(2, 2), vid(2, 2)] == pytest.approx(-8.0, abs=1e-13)


def test_cotan_rigid_motion_invariance():
    V, F = W.c2_mesh(side=12)
    _, _, v0, _ = O.cotan(V, F)
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1.0]])
    _, _, v1, sc = O.cotan(V @ Rz.T + np.array([0.3, -2.0, 5.0]), F)
    assert O.rel_errors(v1, v0, sc).max() < 1e-12


def _bw_small(side=6, seed=11):
    return W.cloth(side, seed, fold=0.2, noise=1e-2)


def test_bw_hessian_is_gradient_derivative():
    rest, F, X = _bw_small()
    ks, ksh = W.K_STRETCH, W.K_SHEAR
    cp, ri, vals, _ = O.bw_hessian(rest, F, X, ks, ksh, 0.0)
    n = len(rest)
    H = _dense(cp, ri, vals, 3 * n)
    assert np.allclose(H, H.T, rtol=0, atol=1e-9 * np.abs(H).max())
    rng = np.random.default_rng(3)
    d = rng.normal(size=(n, 3))
    eps = 1e-6
    gp = O.bw_gradient(rest, F, X + eps * d, ks, ksh)
    gm = O.bw_gradient(rest, F, X - eps * d, ks, ksh)
    fd = (gp - gm) / (2 * eps)
    hd = H @ d.T.reshape(-1)
    assert np.linalg.norm(fd - hd) <= 1e-6 * np.linalg.norm(hd)


def test_bw_gradient_is_energy_derivative():
    rest, F, X = _bw_small()
    ks, ksh = W.K_STRETCH, W.K_SHEAR
    g = O.bw_gradient(rest, F, X, ks, ksh)
    rng = np.random.default_rng(5)
    d = rng.normal(size=X.shape)
    eps = 1e-6
    fd = (O.bw_energy(rest, F, X + eps * d, ks, ksh) - O.bw_energy(rest, F, X - eps * d, ks, ksh)) / (2 * eps)
    assert fd == pytest.approx(g @ d.T.reshape(-1), rel=1e-6)


def test_bw_hessian_translation_invariant():
    rest, F, X = _bw_small()
    cp, ri, vals, sc = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
    _, _, vals2, _ = O.bw_hessian(rest, F, X + np.array([1.5, -0.25, 3.0]), W.K_STRETCH, W.K_SHEAR, W.K_BEND)
    assert O.rel_errors(vals2, vals, sc).max() < 1e-9
    # H annihilates rigid translations (row sums of each coordinate block vanish)
    n = len(rest)
    H = _dense(cp, ri, vals, 3 * n)
    t = np.zeros(3 * n)
    t[:n] = 1.0
    assert np.abs(H @ t).max() <= 1e-9 * np.abs(H).max()


def test_bend_Q_annihilates_flat_modes():
    rest, F, _ = _bw_small(side=5)
    E4 = O.interior_edges(F)
    assert len(E4) > 0
    for q in E4[:: max(1, len(E4) // 10)]:
        Q = O.bend_Q(rest, q)
        assert np.allclose(Q, Q.T)
        pts = rest[q]
        for v in (np.ones(4), pts[:, 0], pts[:, 1]):
            assert np.abs(Q @ v).max() <= 1e-10 * max(1.0, np.abs(Q).max())


def test_interior_edge_quads():
    side = 5
    _, F = W.grid_mesh(side, side)
    E4 = O.interior_edges(F)
    # interior edges of a triangulated grid: 3 (s-1)^2 + 2 (s-1) - 4 (s-1) ... counted directly
    edges = {}
    for f in F:
        for k in range(3):
            a, b = sorted((f[k], f[(k + 1) % 3]))
            edges[(a, b)] = edges.get((a, b), 0) + 1
    assert len(E4) == sum(1 for c in edges.values() if c == 2)
    for v0, v1, v2, v3 in E4:
        assert len({v0, v1, v2, v3}) == 4


def test_bw_pattern_counts_full_size():
    # exact counts of the 256 x 256 cloth, stretch + shear and with bending couplings
    rest, F, _ = W.c3_cloth()
    cp, _ = O.bw_pattern(len(rest), F, with_bend=False)
    assert cp[-1] == 4110354
    cp, _ = O.bw_pattern(len(rest), F, with_bend=True)
    assert cp[-1] == 4890648


def test_scales_bound_values():
    # S(v) >= |v| for every value, and S is at most a modest multiple of |v| where nothing cancels
    rest, F, X = _bw_small()
    _, _, vals, sc = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
    assert np.all(sc >= np.abs(vals) * (1 - 1e-12))
    V, F2 = W.c2_mesh(side=10)
    _, _, v2, s2 = O.cotan(V, F2)
    assert np.all(s2 >= np.abs(v2) * (1 - 1e-12))
    n, cp, ri, a = W.c1_matrix(subdiv=2)
    _, _, v1, s1 = O.ata_plus_a(cp, ri, np.abs(a))
    assert np.allclose(s1, v1, rtol=1e-14)  # all terms positive: S = v
