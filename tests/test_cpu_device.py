"""The CPU device (a Device without UseHIP: the reference's VecWidth/NumThreads path, symcpu.h)
against the FP64 oracle, on the CPU: the same compiled programs (decomposition, kernels,
simplification, tables) the GPU waves form runs, emitted as host C++ with OpenMP and compiled
with the reference's flags.  This pins the symbolic front- and middle-end's values without a
GPU.  Same bar as the GPU parity tests: patterns bit-exact, every value within 1e-12 of the
oracle's term scale.  Structure counts of small programs (kernels, waves, table sizes) are
pinned too, so a change to decomposition or grouping shows."""
import os

import numpy as np
import pytest

import oracle as O
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-12
THREADS = 4


def _check(name, u, cp, ri, rv, sc, values, k=0):
    outer, inner = u.output_pattern(k)
    assert np.array_equal(outer, cp) and np.array_equal(inner, ri), f"{name}: pattern differs"
    err = O.rel_errors(values, rv, sc)
    assert np.isfinite(values).all()
    assert err.max() <= TOL, f"{name}: max relative error {err.max():.3e}"
    return err


def test_sphere_mtx_golden_on_cpu():
    n, cp, ri, a = W.load_mtx(os.path.join(ROOT, "tests", "golden", "sphere.mtx"))
    g = np.load(os.path.join(ROOT, "tests", "golden", "sphere_ata_plus_a.npz"))
    with Unit.ata_plus_a(cp, ri, cpu_threads=THREADS) as u:
        u.compile().execute(a)
        _, _, _, sc = O.ata_plus_a(cp, ri, a)
        _check("sphere.mtx", u, g["indptr"], g["indices"], g["values"], sc, u.get_results())


@pytest.mark.parametrize("subdiv", [0, 2])
def test_ata_plus_a_on_cpu(subdiv):
    n, cp, ri, a = W.c1_matrix(subdiv=subdiv)
    with Unit.ata_plus_a(cp, ri, cpu_threads=THREADS) as u:
        u.compile()
        for seed in (1, 2):
            a = W.random_values(len(ri), seed=seed)
            u.execute(a)
            _check(f"icosphere({subdiv}) seed {seed}", u, *O.ata_plus_a(cp, ri, a), u.get_results())


@pytest.mark.parametrize("side", [4, 16])
def test_cotan_on_cpu(side):
    V, F = W.c2_mesh(side=side)
    with Unit.cotan(len(V), F, cpu_threads=THREADS) as u:
        u.compile().execute(W.to_colmajor(V))
        _check(f"cotan {side}", u, *O.cotan(V, F), u.get_results())


@pytest.mark.parametrize("side", [2, 5, 64])
def test_cotan_direct_assembly_on_cpu(side):
    """The element-matrix assembly of the Laplacian (one 3 x 3 symmetric element matrix per
    face, a node pattern over the vertices): same pattern and values as the per-nonzero trees."""
    V, F = W.c2_mesh(side=side)
    with Unit.cotan(len(V), F, cpu_threads=THREADS, direct_assembly=True) as u:
        info = u.info()
        assert info["num_kernels"] >= 2 and info["num_waves"] == 2, info
        u.compile().execute(W.to_colmajor(V))
        _check(f"cotan direct {side}", u, *O.cotan(V, F), u.get_results())


@pytest.mark.parametrize("side,share,direct,kb", [(4, False, False, W.K_BEND), (8, True, False, W.K_BEND),
                                                  (8, True, True, W.K_BEND), (9, True, False, 0.0),
                                                  (12, True, True, 0.0)])
def test_bw_hessian_on_cpu(side, share, direct, kb):
    rest, F, X = W.cloth(side, seed=30 + side, fold=0.3, noise=1e-2)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, kb, share_templates=share, direct_assembly=direct,
                         cpu_threads=THREADS) as u:
        u.compile().execute(W.to_colmajor(X))
        cp, ri, rv, st, sk, bend = O.bw_hessian_check(rest, F, X, W.K_STRETCH, W.K_SHEAR, kb)
        _check(f"cloth {side} share={share} direct={direct} k_bend={kb}", u, cp, ri, rv, st, u.get_results())


def test_bw_hessian_batch_on_cpu():
    cloths = [W.cloth(s, seed=70 + s) for s in (5, 9, 3)]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND,
                               cpu_threads=THREADS) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            cp, ri, rv, st, sk, bend = O.bw_hessian_check(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
            _check(f"batch cloth {k}", u, cp, ri, rv, st, u.get_results(k), k)


def test_bend_stiffness_known_answer_on_cpu():
    # the quad whose bending stiffness is computed by hand: Q = -1.5 K0 K0^T, K0 = (2, 2, -2, -2)
    R = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 0], [0, 1, 0]], float)
    F = np.array([[0, 1, 2], [1, 0, 3]], np.int32)
    with Unit.bw_hessian(R, F, 0.0, 0.0, 0.5, cpu_threads=1) as u:
        u.compile().execute(W.to_colmajor(R))
        outer, inner = u.output_pattern()
        vals = u.get_results()
    H = np.zeros((12, 12))
    for j in range(12):
        H[inner[outer[j]:outer[j + 1]], j] = vals[outer[j]:outer[j + 1]]
    K0 = np.array([2.0, 2.0, -2.0, -2.0])
    assert np.array_equal(H, np.kron(np.eye(3), 0.5 * -1.5 * np.outer(K0, K0)))


def test_threads_do_not_change_results():
    rest, F, X = W.cloth(10, seed=5)
    got = []
    for t in (1, 3):
        with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, cpu_threads=t) as u:
            u.compile().execute(W.to_colmajor(X))
            got.append(u.get_results())
    assert np.array_equal(got[0], got[1])


def test_cpu_device_rejects_gpu_only_forms():
    rest, F, X = W.cloth(4, seed=1)
    for kw in ({"fused": True}, {"tiled": True}):
        with pytest.raises(Exception):
            Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, cpu_threads=2, **kw)


def _counts(u):
    i = u.info()
    return {k: i[k] for k in ("num_kernels", "num_waves", "num_intermediates", "num_index", "num_const",
                              "num_outpos", "num_results")}


# kernels, waves and table sizes of small programs (decomposition threshold 6, templates
# simplified, shared element templates)
STRUCTURE = {
    "ata_icosphere1": dict(num_kernels=5, num_waves=1, num_intermediates=0, num_index=3552, num_const=0,
                           num_outpos=702, num_results=702),
    "ata_sphere_mtx": dict(num_kernels=5, num_waves=1, num_intermediates=0, num_index=58752, num_const=0,
                           num_outpos=12102, num_results=12102),
    "cotan_6": dict(num_kernels=8, num_waves=3, num_intermediates=200, num_index=1350, num_const=0,
                    num_outpos=206, num_results=206),
    "cloth_5": dict(num_kernels=17, num_waves=2, num_intermediates=1440, num_index=2656, num_const=1560,
                    num_outpos=1425, num_results=1425),
    "cloth_5_noshare": dict(num_kernels=17, num_waves=2, num_intermediates=2376, num_index=2656, num_const=1560,
                            num_outpos=1425, num_results=1425),
    "cloth_5_nobend": dict(num_kernels=7, num_waves=2, num_intermediates=1440, num_index=2656, num_const=0,
                           num_outpos=1233, num_results=1233),
}


def test_program_structure_of_small_programs():
    got = {}
    n, cp, ri, a = W.c1_matrix(subdiv=1)
    got["ata_icosphere1"] = _counts(Unit.ata_plus_a(cp, ri))
    n, cp, ri, a = W.load_mtx(os.path.join(ROOT, "tests", "golden", "sphere.mtx"))
    got["ata_sphere_mtx"] = _counts(Unit.ata_plus_a(cp, ri))
    V, F = W.c2_mesh(side=6)
    got["cotan_6"] = _counts(Unit.cotan(len(V), F))
    r, F, X = W.c3_cloth(side=5)
    got["cloth_5"] = _counts(Unit.bw_hessian(r, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND))
    got["cloth_5_noshare"] = _counts(Unit.bw_hessian(r, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, share_templates=False))
    got["cloth_5_nobend"] = _counts(Unit.bw_hessian(r, F, W.K_STRETCH, W.K_SHEAR, 0.0))
    assert got == STRUCTURE
