"""FP64 oracle (TEST INFRASTRUCTURE ONLY).

ctypes wrapper over oracle/liboracle.so (oracle/oracle.c).  Only tests/, the smoke check in
__graft_entry__.py and the CPU-baseline leg of bench.py import this module; the product
package symboliclib_amd never does.

Every function returns CSC arrays (indptr, indices) plus values and per-value scales, the
denominator of the relative error used by the parity tests.  Two scale kinds (oracle.c):
  "terms"   the magnitudes of every term inside each per-element contribution (the closed form
            evaluated with |.| of every factor and additive term): the rounding-error scale;
  "contrib" |contribution| summed over the per-element contributions of each output.
For A^T A + A both are the sum of |A(k,i) A(k,j)| and |A(i,j)|.
Reference files restated by the C code are cited in oracle.c.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.oracle_ata_plus_a_pattern.restype = ctypes.c_long
        L.oracle_ata_plus_a_pattern.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, ctypes.c_void_p]
        L.oracle_ata_plus_a_values.argtypes = [ctypes.c_int, _i32p, _i32p, _f64p, _i32p, _i32p, _f64p, ctypes.c_void_p]
        L.oracle_cotan_pattern.restype = ctypes.c_long
        L.oracle_cotan_pattern.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, ctypes.c_void_p]
        L.oracle_cotan_values.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, _f64p, ctypes.c_void_p,
                                          ctypes.c_int]
        L.oracle_bw_pattern.restype = ctypes.c_long
        L.oracle_bw_pattern.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int, _i32p, ctypes.c_void_p]
        L.oracle_bw_energy.restype = ctypes.c_double
        L.oracle_bw_energy.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, _f64p, ctypes.c_double, ctypes.c_double]
        L.oracle_bw_gradient.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, _f64p, ctypes.c_double,
                                         ctypes.c_double, _f64p]
        L.oracle_interior_edges.restype = ctypes.c_long
        L.oracle_interior_edges.argtypes = [ctypes.c_int, _i32p, ctypes.c_void_p]
        L.oracle_bend_Q.argtypes = [ctypes.c_int, _i32p, _f64p, _f64p]
        L.oracle_bw_plan_create.restype = ctypes.c_void_p
        L.oracle_bw_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _f64p, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double]
        L.oracle_bw_plan_nnz.restype = ctypes.c_long
        L.oracle_bw_plan_nnz.argtypes = [ctypes.c_void_p]
        L.oracle_bw_plan_values.argtypes = [ctypes.c_void_p, _f64p, _f64p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_bw_plan_free.argtypes = [ctypes.c_void_p]
        L.oracle_bw_plan_pattern.argtypes = [ctypes.c_void_p, _i32p, _i32p, ctypes.c_void_p]
        L.oracle_bw_plan_check.argtypes = [ctypes.c_void_p, _f64p, _f64p, _f64p, _f64p]
        L.oracle_cotan_plan_create.restype = ctypes.c_void_p
        L.oracle_cotan_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, _i32p]
        L.oracle_cotan_plan_values.argtypes = [ctypes.c_void_p, _f64p, _f64p, ctypes.c_void_p, ctypes.c_int]
        L.oracle_cotan_plan_free.argtypes = [ctypes.c_void_p]
        L.oracle_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


SCALE_KINDS = {"terms": 0, "contrib": 1}


def _colmajor(X):
    """n x 3 array -> column-major flat (x of all vertices, then y, then z)."""
    return _f64(np.asarray(X, dtype=np.float64).T.reshape(-1))


def num_threads():
    return lib().oracle_num_threads()


def ata_plus_a(indptr, indices, values):
    """B = A^T A + A for CSC A (n x n)."""
    L = lib()
    n = len(indptr) - 1
    cp, ri, a = _i32(indptr), _i32(indices), _f64(values)
    bcp = np.zeros(n + 1, np.int32)
    nnz = L.oracle_ata_plus_a_pattern(n, cp, ri, bcp, None)
    bri = np.zeros(nnz, np.int32)
    L.oracle_ata_plus_a_pattern(n, cp, ri, bcp, bri.ctypes.data_as(ctypes.c_void_p))
    bv = np.zeros(nnz)
    sc = np.zeros(nnz)
    L.oracle_ata_plus_a_values(n, cp, ri, a, bcp, bri, bv, sc.ctypes.data_as(ctypes.c_void_p))
    return bcp, bri, bv, sc


def cotan_pattern(nv, F):
    L = lib()
    F = _i32(F).reshape(-1)
    cp = np.zeros(nv + 1, np.int32)
    nnz = L.oracle_cotan_pattern(nv, len(F) // 3, F, cp, None)
    ri = np.zeros(nnz, np.int32)
    L.oracle_cotan_pattern(nv, len(F) // 3, F, cp, ri.ctypes.data_as(ctypes.c_void_p))
    return cp, ri


def cotan(V, F, scale="terms"):
    """Cotangent Laplacian of mesh (V: n x 3, F: m x 3)."""
    L = lib()
    nv = len(V)
    cp, ri = cotan_pattern(nv, F)
    vals = np.zeros(len(ri))
    sc = np.zeros(len(ri))
    Ff = _i32(F).reshape(-1)
    L.oracle_cotan_values(nv, len(Ff) // 3, Ff, _colmajor(V), vals, sc.ctypes.data_as(ctypes.c_void_p),
                          SCALE_KINDS[scale])
    return cp, ri, vals, sc


def bw_pattern(nv, F, with_bend=True):
    L = lib()
    F = _i32(F).reshape(-1)
    cp = np.zeros(3 * nv + 1, np.int32)
    nnz = L.oracle_bw_pattern(nv, len(F) // 3, F, int(with_bend), cp, None)
    ri = np.zeros(nnz, np.int32)
    L.oracle_bw_pattern(nv, len(F) // 3, F, int(with_bend), cp, ri.ctypes.data_as(ctypes.c_void_p))
    return cp, ri


class BWPlan:
    """Prepared Baraff-Witkin Hessian assembly; values() may be called for many X."""

    def __init__(self, rest, F, k_stretch, k_shear, k_bend):
        self._F = _i32(F).reshape(-1)
        self._R = _colmajor(rest)
        self.nv = len(rest)
        self.nf = len(self._F) // 3
        self._p = lib().oracle_bw_plan_create(self.nv, self.nf, self._F, self._R, k_stretch, k_shear, k_bend)
        self.nnz = lib().oracle_bw_plan_nnz(self._p)

    def values(self, X, with_scale=True, scale="terms"):
        vals = np.zeros(self.nnz)
        sc = np.zeros(self.nnz) if with_scale else None
        lib().oracle_bw_plan_values(self._p, _colmajor(X), vals,
                                    sc.ctypes.data_as(ctypes.c_void_p) if with_scale else None, SCALE_KINDS[scale])
        return vals, sc

    def pattern(self):
        """(indptr, indices, bend): bend[s] is True when output s has a bending constant."""
        cp = np.zeros(3 * self.nv + 1, np.int32)
        ri = np.zeros(self.nnz, np.int32)
        bend = np.zeros(self.nnz, np.uint8)
        lib().oracle_bw_plan_pattern(self._p, cp, ri, bend.ctypes.data_as(ctypes.c_void_p))
        return cp, ri, bend.astype(bool)

    def check(self, X):
        """(values, scale "terms", scale "contrib")."""
        vals, st, sk = np.zeros(self.nnz), np.zeros(self.nnz), np.zeros(self.nnz)
        lib().oracle_bw_plan_check(self._p, _colmajor(X), vals, st, sk)
        return vals, st, sk

    def close(self):
        if self._p:
            lib().oracle_bw_plan_free(self._p)
            self._p = None

    def __del__(self):
        self.close()


class CotanPlan:
    def __init__(self, nv, F):
        self._F = _i32(F).reshape(-1)
        self.nv = nv
        self._p = lib().oracle_cotan_plan_create(nv, len(self._F) // 3, self._F)

    def values(self, V, nnz, with_scale=False, scale="terms"):
        vals = np.zeros(nnz)
        sc = np.zeros(nnz) if with_scale else None
        lib().oracle_cotan_plan_values(self._p, _colmajor(V), vals,
                                       sc.ctypes.data_as(ctypes.c_void_p) if with_scale else None,
                                       SCALE_KINDS[scale])
        return vals, sc

    def close(self):
        if self._p:
            lib().oracle_cotan_plan_free(self._p)
            self._p = None

    def __del__(self):
        self.close()


def bw_hessian(rest, F, X, k_stretch, k_shear, k_bend, scale="terms"):
    """Full symmetric Hessian of the stretch + shear energy plus k_bend * Q per interior edge."""
    cp, ri = bw_pattern(len(rest), F, with_bend=k_bend != 0.0)
    plan = BWPlan(rest, F, k_stretch, k_shear, k_bend)
    vals, sc = plan.values(X, scale=scale)
    plan.close()
    assert len(vals) == len(ri)
    return cp, ri, vals, sc


def bw_hessian_check(rest, F, X, k_stretch, k_shear, k_bend):
    """Everything a cloth parity check needs: (indptr, indices, values, scale "terms", scale
    "contrib", bend) where bend[s] is True when output s carries a bending constant."""
    plan = BWPlan(rest, F, k_stretch, k_shear, k_bend)
    cp, ri, bend = plan.pattern()
    vals, st, sk = plan.check(X)
    plan.close()
    return cp, ri, vals, st, sk, bend


def bw_energy(rest, F, X, k_stretch, k_shear):
    Ff = _i32(F).reshape(-1)
    return lib().oracle_bw_energy(len(rest), len(Ff) // 3, Ff, _colmajor(rest), _colmajor(X), k_stretch, k_shear)


def bw_gradient(rest, F, X, k_stretch, k_shear):
    """Gradient (length 3n, index vertex + coord * n) of the stretch + shear energy."""
    Ff = _i32(F).reshape(-1)
    g = np.zeros(3 * len(rest))
    lib().oracle_bw_gradient(len(rest), len(Ff) // 3, Ff, _colmajor(rest), _colmajor(X), k_stretch, k_shear, g)
    return g


def interior_edges(F):
    Ff = _i32(F).reshape(-1)
    ne = lib().oracle_interior_edges(len(Ff) // 3, Ff, None)
    E4 = np.zeros(4 * ne, np.int32)
    lib().oracle_interior_edges(len(Ff) // 3, Ff, E4.ctypes.data_as(ctypes.c_void_p))
    return E4.reshape(-1, 4)


def bend_Q(rest, quad):
    Q = np.zeros(16)
    lib().oracle_bend_Q(len(rest), _i32(quad), _colmajor(rest), Q)
    return Q.reshape(4, 4)


_LIB_LD = None
_ldp = np.ctypeslib.ndpointer(dtype=np.longdouble, flags="C_CONTIGUOUS")


def lib_extended():
    """The same closed forms in 80-bit x87 extended precision (oracle/liboracle_ld.so)."""
    global _LIB_LD
    if _LIB_LD is None:
        path = os.path.join(_HERE, "liboracle_ld.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle_ld.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        LD = ctypes.c_longdouble
        L.oracle_bw_plan_create.restype = ctypes.c_void_p
        L.oracle_bw_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _ldp, LD, LD, LD]
        L.oracle_bw_plan_nnz.restype = ctypes.c_long
        L.oracle_bw_plan_nnz.argtypes = [ctypes.c_void_p]
        L.oracle_bw_plan_check.argtypes = [ctypes.c_void_p, _ldp, _ldp, _ldp, _ldp]
        L.oracle_bw_plan_free.argtypes = [ctypes.c_void_p]
        L.oracle_cotan_values.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _ldp, _ldp, ctypes.c_void_p, ctypes.c_int]
        _LIB_LD = L
    return _LIB_LD


def bw_hessian_extended(rest, F, X, k_stretch, k_shear, k_bend):
    """Values of bw_hessian evaluated in 80-bit extended precision (long double array, same
    pattern and order as bw_hessian)."""
    L = lib_extended()
    Ff = _i32(F).reshape(-1)
    R = np.ascontiguousarray(np.asarray(rest, np.float64).T.reshape(-1), dtype=np.longdouble)
    Xl = np.ascontiguousarray(np.asarray(X, np.float64).T.reshape(-1), dtype=np.longdouble)
    p = L.oracle_bw_plan_create(len(rest), len(Ff) // 3, Ff, R, k_stretch, k_shear, k_bend)
    n = L.oracle_bw_plan_nnz(p)
    vals, st, sk = (np.zeros(n, np.longdouble) for _ in range(3))
    L.oracle_bw_plan_check(p, Xl, vals, st, sk)
    L.oracle_bw_plan_free(p)
    return vals


def cotan_extended(V, F):
    """Values of cotan evaluated in 80-bit extended precision (long double array)."""
    L = lib_extended()
    nv = len(V)
    nnz = cotan_pattern(nv, F)[0][-1]
    Ff = _i32(F).reshape(-1)
    vals = np.zeros(nnz, np.longdouble)
    Vl = np.ascontiguousarray(np.asarray(V, np.float64).T.reshape(-1), dtype=np.longdouble)
    L.oracle_cotan_values(nv, len(Ff) // 3, Ff, Vl, vals, None, 0)
    return vals


def implicit_euler_system(indptr, indices, K, D, m, v, f, h):
    """The implicit-Euler system of a cloth step (examples/BaraffWitkin/implementation/
    cloth.cpp:97-112, Cloth::generateMatrixAndRHS): A = M - h (D + h K) on K's CSC pattern with
    M = diag(m), and b = M v + h (f - D2 v), D2 the symmetric matrix of D's lower triangle
    (cloth.cpp:107, selfadjointView<Lower>: a stored entry D(i, j) with i >= j stands for
    D2(i, j) and D2(j, i); entries above the diagonal are not read).  D = None: no damping.
    Returns (A, b, scale_A, scale_b), the scales being the sums of the absolute terms."""
    indptr = np.asarray(indptr, np.int64)
    indices = np.asarray(indices, np.int64)
    n = len(indptr) - 1
    cols = np.repeat(np.arange(n), np.diff(indptr))
    K = np.asarray(K, np.float64)
    D = np.zeros_like(K) if D is None else np.asarray(D, np.float64)
    m, v, f = (np.asarray(x, np.float64) for x in (m, v, f))
    diag = indices == cols
    a = -h * (D + h * K)
    A = np.where(diag, m[cols] + a, a)
    scale_A = np.where(diag, np.abs(m[cols]), 0.0) + abs(h) * (np.abs(D) + abs(h) * np.abs(K))
    lower = indices >= cols
    strict = indices > cols
    D2v = np.zeros(n)
    D2abs = np.zeros(n)
    np.add.at(D2v, indices[lower], D[lower] * v[cols[lower]])
    np.add.at(D2v, cols[strict], D[strict] * v[indices[strict]])
    np.add.at(D2abs, indices[lower], np.abs(D[lower] * v[cols[lower]]))
    np.add.at(D2abs, cols[strict], np.abs(D[strict] * v[indices[strict]]))
    b = m * v + h * (f - D2v)
    scale_b = np.abs(m * v) + abs(h) * (np.abs(f) + D2abs)
    return A, b, scale_A, scale_b


def csc_symv(indptr, indices, A, x):
    """y = A x for a symmetric CSC A, and the scale sum_j |A(i, j) x(j)| of each y(i)."""
    indptr = np.asarray(indptr, np.int64)
    indices = np.asarray(indices, np.int64)
    n = len(indptr) - 1
    cols = np.repeat(np.arange(n), np.diff(indptr))
    y = np.zeros(n)
    s = np.zeros(n)
    np.add.at(y, indices, A * x[cols])
    np.add.at(s, indices, np.abs(A * x[cols]))
    return y, s


def rel_errors(values, ref_values, scale):
    """Per-value |v - v_ref| / scale (scale 0 -> absolute error)."""
    scale = np.where(scale > 0, scale, 1.0)
    return np.abs(np.asarray(values) - np.asarray(ref_values)) / scale
