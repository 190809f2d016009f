"""The step after the Hessian on the GPU (include/symhip.h: symhip_implicit_euler_system,
symhip_csc_symv): the implicit-Euler system A = M - h (D + h K), b = M v + h (f - D2 v) of a
cloth step (examples/BaraffWitkin/implementation/cloth.cpp:97-112) assembled from the stiffness
K where the fused program left it in device memory, and the symmetric product y = A x a solver
of the system runs.  Compared with the oracle's restatement (oracle.implicit_euler_system,
oracle.csc_symv) on the same inputs: values within 1e-12 of the sum of the absolute terms."""
import ctypes

import numpy as np
import pytest

import oracle as O
from symboliclib_amd import _native
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

from hip_runtime import DeviceBuffer, Stream

pytestmark = pytest.mark.gpu
TOL = 1e-12
H = 0.05  # the time step of examples/BaraffWitkin/main.cpp:53


def _symmetric_damping(outer, inner, seed):
    """Values on a symmetric pattern with D(i, j) = D(j, i): a hash of the unordered pair."""
    cols = np.repeat(np.arange(len(outer) - 1), np.diff(outer))
    lo, hi = np.minimum(inner, cols).astype(np.uint64), np.maximum(inner, cols).astype(np.uint64)
    z = (lo * np.uint64(0x9E3779B97F4A7C15) ^ hi * np.uint64(0xC2B2AE3D27D4EB4F) ^ np.uint64(seed)) >> np.uint64(11)
    return -(z.astype(np.float64) / float(1 << 53)) * 1e-2


def _check(name, got, want, scale):
    err = O.rel_errors(got, want, scale)
    print(f"{name}: n={len(want)} max_rel_err={err.max():.3e} median={np.median(err):.3e}", flush=True)
    assert np.isfinite(got).all() and err.max() <= TOL, f"{name}: max relative error {err.max():.3e}"


@pytest.mark.parametrize("side,damping", [(3, False), (24, True), (24, False), (128, True)])
def test_implicit_euler_system_from_the_fused_hessian(require_device, side, damping):
    L = _native.lib()
    rest, F, X = W.cloth(side, seed=40 + side)
    rng = np.random.default_rng(side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.compile().execute(W.to_colmajor(X))
        outer, inner = (np.asarray(a, np.int32) for a in u.output_pattern())
        n = len(outer) - 1
        K = u.get_results()
        D = _symmetric_damping(outer, inner, seed=side) if damping else None
        m = np.full(n, 1.0 / len(rest))  # total mass 1 over the vertices, on every dof
        v = 0.1 * rng.uniform(-1, 1, n)
        f = rng.standard_normal(n)
        x = rng.standard_normal(n)
        bufs = {k: DeviceBuffer(a) for k, a in dict(outer=outer, inner=inner, m=m, v=v, f=f, x=x,
                                                     A=np.zeros(len(inner)), b=np.zeros(n), y=np.zeros(n)).items()}
        if damping:
            bufs["D"] = DeviceBuffer(D)
        stream = Stream()
        try:
            print(f"implicit-Euler system: cloth {side}x{side}, n={n}, nnz={len(inner)}, damping={damping}", flush=True)
            rc = L.symhip_implicit_euler_system(n, bufs["outer"].ptr, bufs["inner"].ptr, u.output_device_ptr(0),
                                                bufs["D"].ptr if damping else None, bufs["m"].ptr, bufs["v"].ptr,
                                                bufs["f"].ptr, H, bufs["A"].ptr, bufs["b"].ptr, stream.handle)
            _native.check(rc, "symhip_implicit_euler_system")
            _native.check(L.symhip_csc_symv(n, bufs["outer"].ptr, bufs["inner"].ptr, bufs["A"].ptr, bufs["x"].ptr,
                                            bufs["y"].ptr, stream.handle), "symhip_csc_symv")
            stream.synchronize()
            A_ref, b_ref, sa, sb = O.implicit_euler_system(outer, inner, K, D, m, v, f, H)
            _check("A", bufs["A"].download(), A_ref, sa)
            _check("b", bufs["b"].download(), b_ref, sb)
            y_ref, sy = O.csc_symv(outer, inner, A_ref, x)
            _check("y = A x", bufs["y"].download(), y_ref, sy)
        finally:
            for b in bufs.values():
                b.free()
            stream.destroy()


def test_system_entry_points_refuse_bad_arguments(require_device):
    L = _native.lib()
    assert L.symhip_implicit_euler_system(4, None, None, None, None, None, None, None, H, None, None, None) != 0
    assert L.symhip_csc_symv(-1, None, None, None, None, None, None) != 0
    assert L.symhip_csc_symv(0, None, None, None, None, None, None) == 0
