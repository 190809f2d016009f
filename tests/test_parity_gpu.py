"""GPU parity: programs compiled by the new HIP emitter and run through the C ABI against the
FP64 oracle.

Bar: output patterns (CSC outer/inner arrays) bit-exact; every value within relative error
1e-12 of the oracle, the error of a value v being |v - v_oracle| / S(v) where S(v) is the sum of
the absolute magnitudes of the terms v is made of: the oracle's closed form evaluated with the
absolute value of every factor and of every additive term (products A(k,i) A(k,j) and A(i,j);
|e0_k e1_k| / |e0 x e1| per corner weight; for the cloth, the stretch and shear terms of every
per-face Hessian entry from the edge vectors on, plus |k_bend Q| per bending constant).  The
maximum and median errors are printed for every case.

Cloth cases also print the error against the per-element contribution scale (the sum of
|per-face Hessian entry| and |bending constant| over an output's contributions) and, separately,
the largest error on outputs with and without a bending constant.  Every fused and batched
cloth case runs with bending on and off: with k_bend = 0 every output is a pure
stretch + shear automatic-differentiation Hessian entry, checked at the same bar.
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-12


def _check(name, u, ref, values, k=0):
    cp, ri, rv, sc = ref
    outer, inner = u.output_pattern(k)
    assert np.array_equal(outer, cp), f"{name}: outer index array differs"
    assert np.array_equal(inner, ri), f"{name}: inner index array differs"
    err = O.rel_errors(values, rv, sc)
    print(f"\n{name}: nnz={len(rv)} max_rel_err={err.max():.3e} median_rel_err={np.median(err):.3e}", flush=True)
    assert np.isfinite(values).all(), f"{name}: non-finite values"
    assert err.max() <= TOL, f"{name}: max relative error {err.max():.3e} at {int(err.argmax())}"


def _cloth_errors(name, u, rest, F, X, ks, ksh, kb, values, k=0):
    """Pattern and value check of one cloth Hessian; returns the errors on both scales."""
    cp, ri, rv, st, sk, bend = O.bw_hessian_check(rest, F, X, ks, ksh, kb)
    outer, inner = u.output_pattern(k)
    assert np.array_equal(outer, cp), f"{name}: outer index array differs"
    assert np.array_equal(inner, ri), f"{name}: inner index array differs"
    assert np.isfinite(values).all(), f"{name}: non-finite values"
    et = O.rel_errors(values, rv, st)
    ek = O.rel_errors(values, rv, sk)
    return et, ek, bend


def _cloth_summary(et, ek, bend):
    s = (f"max_rel_err={et.max():.3e} median={np.median(et):.3e} | contribution scale max={ek.max():.3e} "
         f"median={np.median(ek):.3e}")
    if bend.any():
        s += f" | with bend max={et[bend].max():.3e}"
        if (~bend).any():
            s += f", without bend max={et[~bend].max():.3e} (contribution scale {ek[~bend].max():.3e})"
    return s


def _check_cloth(name, u, rest, F, X, kb, values, k=0, ks=None, ksh=None):
    ks = W.K_STRETCH if ks is None else ks
    ksh = W.K_SHEAR if ksh is None else ksh
    et, ek, bend = _cloth_errors(name, u, rest, F, X, ks, ksh, kb, values, k)
    print(f"\n{name} k_bend={kb}: nnz={len(et)} {_cloth_summary(et, ek, bend)}", flush=True)
    assert et.max() <= TOL, f"{name}: max relative error {et.max():.3e} at {int(et.argmax())}"
    return et, ek


BENDS = [pytest.param(W.K_BEND, id="bend"), pytest.param(0.0, id="nobend")]


def test_sphere_mtx_golden(require_device):
    n, cp, ri, a = W.load_mtx(os.path.join(ROOT, "tests", "golden", "sphere.mtx"))
    g = np.load(os.path.join(ROOT, "tests", "golden", "sphere_ata_plus_a.npz"))
    with Unit.ata_plus_a(cp, ri, threads_per_block=128) as u:
        u.compile().execute(a)
        _, _, _, sc = O.ata_plus_a(cp, ri, a)
        _check("sphere.mtx A^T A + A", u, (g["indptr"], g["indices"], g["values"], sc), u.get_results())


FORMS = [pytest.param(True, id="tiled"), pytest.param(False, id="waves")]


@pytest.mark.parametrize("tiled", FORMS)
@pytest.mark.parametrize("subdiv", [1, 5])
def test_ata_plus_a(require_device, subdiv, tiled):
    n, cp, ri, a = W.c1_matrix(subdiv=subdiv)
    with Unit.ata_plus_a(cp, ri, tiled=tiled) as u:
        u.compile()
        u.execute(a)
        _check(f"A^T A + A icosphere({subdiv})", u, O.ata_plus_a(cp, ri, a), u.get_results())
        # a second value set through the same compiled program
        a2 = W.random_values(len(ri), seed=99)
        u.execute(a2)
        _check(f"A^T A + A icosphere({subdiv}) second values", u, O.ata_plus_a(cp, ri, a2), u.get_results())


@pytest.mark.parametrize("side,tiled", [(4, True), (4, False), (64, True), (64, False), (512, True), (512, False)])
def test_cotan(require_device, side, tiled):
    V, F = W.c2_mesh(side=side)
    with Unit.cotan(len(V), F, tiled=tiled) as u:
        assert u.info()["tiled"] == int(tiled), u.info()["tiling"]
        u.compile().execute(W.to_colmajor(V))
        _check(f"cotan grid {side}x{side}", u, O.cotan(V, F), u.get_results())


@pytest.mark.parametrize("side", [3, 64, 512])
def test_cotan_direct_assembly(require_device, side):
    V, F = W.c2_mesh(side=side)
    with Unit.cotan(len(V), F, direct_assembly=True) as u:
        u.compile().execute(W.to_colmajor(V))
        _check(f"cotan direct {side}x{side}", u, O.cotan(V, F), u.get_results())


@pytest.mark.parametrize("side,lds,groups,schedule", [(3, 53760, 0, "queue"), (64, 40448, 1, "static"),
                                                      (64, 81408, 2, "largest"), (512, 53760, 0, "queue"),
                                                      (512, 81408, 1, "static"), (512, 81408, 3, "strided")])
def test_cotan_fused(require_device, side, lds, groups, schedule):
    """The cotangent Laplacian through the fused element tiles (element matrices in LDS, sums
    per owned vertex column), replayed from a graph a second time with new positions."""
    V, F = W.c2_mesh(side=side)
    with Unit.cotan(len(V), F, fused=True, tile_lds_bytes=lds, output_groups=groups, schedule=schedule) as u:
        assert u.info()["tiled"] == 2, u.info()["tiling"]
        u.compile().execute(W.to_colmajor(V))
        _check(f"cotan fused {side}x{side} lds={lds} groups={groups} {schedule}", u, O.cotan(V, F), u.get_results())
        V2, _ = W.c2_mesh(side=side, seed=77)
        u.execute(W.to_colmajor(V2))
        u.run_graph()
        _check(f"cotan fused {side}x{side} second positions", u, O.cotan(V2, F), u.get_results())


@pytest.mark.parametrize("side,share,tiled", [(4, False, True), (4, True, True), (4, True, False), (32, True, True),
                                              (32, True, False), (256, True, True)])
def test_bw_hessian(require_device, side, share, tiled):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, share_templates=share, tiled=tiled) as u:
        assert u.info()["tiled"] == int(tiled), u.info()["tiling"]
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"Baraff-Witkin {side}x{side} share={share}", u, rest, F, X, W.K_BEND, u.get_results())


def test_bw_hessian_stretch_shear_only(require_device):
    rest, F, X = W.cloth(9, seed=4, fold=0.3, noise=1e-2)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, 0.0) as u:
        u.compile().execute(W.to_colmajor(X))
        _check_cloth("Baraff-Witkin 9x9", u, rest, F, X, 0.0, u.get_results())


# Bending stiffness alone on the two-triangle quads whose Q is computed by hand: v0 = (0,0,0),
# v1 = (1,1,0) the hinge, v2 = (1,0,0), v3 = (0,1,0): every cotangent 1, K0 = (2, 2, -2, -2),
# A0 = A1 = 1, Q = -1.5 K0 K0^T; and v0 = (0,0,0), v1 = (2,0,0), v2 = (0,1,0), v3 = (1,-2,0):
# K0 = (2.5, 0.5, -2, -1), A0 = 2, A1 = 4, Q = -0.5 K0 K0^T.  With k_stretch = k_shear = 0 the
# Hessian is k_bend (I_3 kron Q) on the four vertices, exactly.
BEND_QUADS = [
    (np.array([[0, 0, 0], [1, 1, 0], [1, 0, 0], [0, 1, 0]], float), -1.5, np.array([2.0, 2.0, -2.0, -2.0])),
    (np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0], [1, -2, 0]], float), -0.5, np.array([2.5, 0.5, -2.0, -1.0])),
]


def _bend_expected(coef, K0, kb):
    Q = coef * np.outer(K0, K0)
    return np.kron(np.eye(3), kb * Q)  # row / column = vertex + 4 * coordinate


@pytest.mark.parametrize("form", ["symbolic", "direct", "fused", "batch_fused"])
def test_bend_stiffness_known_answer(require_device, form):
    F = np.array([[0, 1, 2], [1, 0, 3]], np.int32)
    kb = 0.5
    if form == "batch_fused":
        with Unit.bw_hessian_batch([(R, F) for R, _, _ in BEND_QUADS], 0.0, 0.0, kb, fused=True) as u:
            u.compile().execute(*[W.to_colmajor(R) for R, _, _ in BEND_QUADS])
            got = [(u.output_pattern(k), u.get_results(k)) for k in range(len(BEND_QUADS))]
    else:
        got = []
        for R, _, _ in BEND_QUADS:
            opt = {"symbolic": {}, "direct": {"direct_assembly": True}, "fused": {"fused": True}}[form]
            with Unit.bw_hessian(R, F, 0.0, 0.0, kb, **opt) as u:
                u.compile().execute(W.to_colmajor(R))
                got.append((u.output_pattern(), u.get_results()))
    for (R, coef, K0), ((outer, inner), vals) in zip(BEND_QUADS, got):
        H = np.zeros((12, 12))
        for j in range(12):
            H[inner[outer[j]:outer[j + 1]], j] = vals[outer[j]:outer[j + 1]]
        expect = _bend_expected(coef, K0, kb)
        print(f"\nbend known answer {form} coef={coef}: max |H - expected| = {np.abs(H - expect).max():.3e}")
        assert np.allclose(H, expect, rtol=1e-15, atol=1e-15), f"{form}: bending Hessian differs from the known answer"


@pytest.mark.parametrize("tiled", FORMS)
@pytest.mark.parametrize("tpb", [64, 1024])
def test_threads_per_block(require_device, tpb, tiled):
    V, F = W.c2_mesh(side=20)
    with Unit.cotan(len(V), F, threads_per_block=tpb, tiled=tiled) as u:
        u.compile().execute(W.to_colmajor(V))
        _check(f"cotan tpb={tpb}", u, O.cotan(V, F), u.get_results())


def test_single_entry_matrix(require_device):
    cp = np.array([0, 1], np.int32)
    ri = np.array([0], np.int32)
    with Unit.ata_plus_a(cp, ri) as u:
        u.compile().execute(np.array([3.0]))
        assert u.get_results()[0] == 12.0  # 3 * 3 + 3


@pytest.mark.parametrize("lds", [16 * 1024, 160 * 1024])
def test_bw_hessian_tile_budgets(require_device, lds):
    rest, F, X = W.c3_cloth(side=24)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, tiled=True, tile_lds_bytes=lds) as u:
        assert u.info()["tiled"] == 1, u.info()["tiling"]
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"Baraff-Witkin 24x24 tile LDS {lds}", u, rest, F, X, W.K_BEND, u.get_results())


def test_matrix_with_empty_columns(require_device):
    # 4 x 4 with columns 1 and 3 empty
    cp = np.array([0, 2, 2, 3, 3], np.int32)
    ri = np.array([0, 2, 2], np.int32)
    a = np.array([1.5, -2.0, 0.25])
    with Unit.ata_plus_a(cp, ri) as u:
        u.compile().execute(a)
        _check("empty columns", u, O.ata_plus_a(cp, ri, a), u.get_results())


@pytest.mark.parametrize("side,tiled", [(16, False), (16, True), (256, False)])
def test_bw_hessian_direct_assembly(require_device, side, tiled):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, direct_assembly=True, tiled=tiled) as u:
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"Baraff-Witkin {side}x{side} direct assembly tiled={tiled}", u, rest, F, X, W.K_BEND,
                     u.get_results())


@pytest.mark.parametrize("tiled", FORMS)
def test_bw_hessian_batch(require_device, tiled):
    sides = [5, 17, 32, 9, 40]
    cloths = [W.cloth(s, seed=100 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, tiled=tiled) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            _check_cloth(f"batch cloth {k} ({sides[k]}x{sides[k]}) tiled={tiled}", u, rest, F, X, W.K_BEND,
                         u.get_results(k), k)


FUSED_CASES = [
    # (side, schedule, lds bytes, output groups, waves per EU, split)
    (4, "queue", 0, 0, 0, False),
    (16, "static", 0, 0, 0, False),
    (16, "queue", 8192, 1, 0, False),
    (33, "largest", 16384, 9, 5, False),
    (64, "interleaved", 49152, 3, 4, False),
    (256, "queue", 0, 9, 5, False),
    (4, "static", 0, 0, 0, True),
    (33, "queue", 16384, 3, 0, True),
    (64, "strided", 49152, 5, 4, True),
    (256, "queue", 0, 5, 0, True),
    (33, "strided", 0, 2, 0, False),
    (256, "strided", 81408, 3, 0, False),
]


@pytest.mark.parametrize("kb", BENDS)
@pytest.mark.parametrize("side,schedule,lds,groups,wpe,split", FUSED_CASES)
def test_bw_hessian_fused(require_device, side, schedule, lds, groups, wpe, split, kb):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, kb, fused=True, schedule=schedule, tile_lds_bytes=lds,
                         output_groups=groups, waves_per_eu=wpe, split=split) as u:
        assert u.info()["tiled"] == 2, u.info()["tiling"]
        u.compile().execute(W.to_colmajor(X))
        assert u.num_launches() == (2 if split else 1)
        name = f"Baraff-Witkin {side}x{side} fused {schedule} lds={lds} groups={groups} wpe={wpe} split={split}"
        _check_cloth(name, u, rest, F, X, kb, u.get_results())
        # a second evaluation of the same launch (persistent work-queue counters start from zero again)
        X2 = W.cloth(side, seed=7)[2]
        u.execute(W.to_colmajor(X2))
        _check_cloth(f"{name} second positions", u, rest, F, X2, kb, u.get_results())


# One contribution list per output (output_sums: the form for Hessians that are not made of
# node blocks), in one and two launches, staged directly and through registers.
OUTPUT_SUM_CASES = [(16, "static", 0, 0, False, False), (64, "queue", 49152, 3, True, False),
                    (33, "strided", 81408, 2, False, True), (256, "static", 81408, 2, False, False)]


@pytest.mark.parametrize("kb", BENDS)
@pytest.mark.parametrize("side,schedule,lds,groups,split,staging", OUTPUT_SUM_CASES)
def test_bw_hessian_fused_output_sums(require_device, side, schedule, lds, groups, split, staging, kb):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, kb, fused=True, schedule=schedule, tile_lds_bytes=lds,
                         output_groups=groups, split=split, register_staging=staging, output_sums=True) as u:
        assert "block records" not in u.source()
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"Baraff-Witkin {side}x{side} fused output sums {schedule} lds={lds} split={split} "
                     f"staging={staging}", u, rest, F, X, kb, u.get_results())


# Staging through registers instead of direct global-to-LDS loads (the fallback path): element
# tiles in one and two launches, several schedules and tile budgets.
STAGING_CASES = [(16, "static", 0, 0, False), (33, "queue", 16384, 3, False), (64, "strided", 49152, 5, True),
                 (48, "static", 81408, 2, False)]


@pytest.mark.parametrize("kb", BENDS)
@pytest.mark.parametrize("side,schedule,lds,groups,split", STAGING_CASES)
def test_bw_hessian_fused_register_staging(require_device, side, schedule, lds, groups, split, kb):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, kb, fused=True, schedule=schedule, tile_lds_bytes=lds,
                         output_groups=groups, split=split, register_staging=True) as u:
        assert "__builtin_amdgcn_global_load_lds" not in u.source()
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"Baraff-Witkin {side}x{side} fused register staging {schedule} lds={lds} split={split}", u,
                     rest, F, X, kb, u.get_results())


@pytest.mark.parametrize("tpb", [64, 128])
def test_bw_hessian_fused_strided_small_blocks(require_device, tpb):
    # blocks of one or two waves: the node numbers of a tile take several rounds of 32 nodes
    # per wave (loaded ahead of the header, or in the staging loop when that needs more than
    # four rounds)
    rest, F, X = W.c3_cloth(side=48)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, schedule="strided",
                         tile_lds_bytes=81408, output_groups=2, threads_per_block=tpb) as u:
        assert max(u.tile_stats()[:, 3]) > 0
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"Baraff-Witkin 48x48 fused strided tpb={tpb}", u, rest, F, X, W.K_BEND, u.get_results())


@pytest.mark.parametrize("kb", BENDS)
@pytest.mark.parametrize("schedule,split,staging", [("static", False, False), ("queue", False, False),
                                                    ("largest", False, False), ("interleaved", False, False),
                                                    ("static", True, False), ("queue", True, False),
                                                    ("queue", False, True), ("static", True, True)])
def test_bw_hessian_batch_fused(require_device, schedule, split, staging, kb):
    sides = [5, 17, 32, 9, 40, 3]
    cloths = [W.cloth(s, seed=200 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, kb, fused=True,
                               schedule=schedule, split=split, register_staging=staging) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            _check_cloth(f"fused batch {schedule} split={split} staging={staging} cloth {k} ({sides[k]}x{sides[k]})",
                         u, rest, F, X, kb, u.get_results(k), k)


@pytest.mark.parametrize("schedule", ["static", "queue"])
def test_bw_hessian_fused_tiny_cloths(require_device, schedule):
    # one-tile instances: two triangles (a single interior edge) and a 3x3 sheet, batched
    sides = [2, 3, 2]
    cloths = [W.cloth(s, seed=400 + k) for k, s in enumerate(sides)]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               schedule=schedule, output_groups=2) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            _check_cloth(f"tiny fused {schedule} cloth {k} ({sides[k]}x{sides[k]})", u, rest, F, X, W.K_BEND,
                         u.get_results(k), k)


@pytest.mark.parametrize("sums", ["blocks", "outputs"])
@pytest.mark.parametrize("n,schedule", [(10, "static"), (18, "static"), (18, "queue")])
def test_bw_hessian_fused_long_contribution_lists(require_device, n, schedule, sums):
    # the center of an n-triangle fan sums n element contributions.  One list per output: 5 slot
    # pairs (n = 10, the most the unrolled sums take) or 9 (n = 18, the per-lane loop); block
    # sums: the center's diagonal block reads n entries.  Then fans batched with a cloth.
    outputs = sums == "outputs"
    rest, F, X = W.wheel(n, seed=5)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, schedule=schedule,
                         output_groups=2, output_sums=outputs) as u:
        if outputs:
            assert ("q < k; q += 2u" in u.source()) == (n > 10)
        else:
            assert "block records" in u.source()
        u.compile().execute(W.to_colmajor(X))
        _check_cloth(f"fan of {n} fused {schedule} {sums}", u, rest, F, X, W.K_BEND, u.get_results())
    meshes = [W.wheel(18, seed=6), W.cloth(9, seed=8), W.wheel(7, seed=9)]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in meshes], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               schedule=schedule, output_groups=2, output_sums=outputs) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in meshes])
        for k, (rest, F, X) in enumerate(meshes):
            _check_cloth(f"fans and a cloth fused {schedule}, mesh {k}", u, rest, F, X, W.K_BEND, u.get_results(k), k)


def test_fused_retune_between_one_and_two_launches(require_device):
    sides = [20, 7, 33]
    cloths = [W.cloth(s, seed=250 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        one = [u.get_results(k) for k in range(len(sides))]
        for k, (rest, F, X) in enumerate(cloths):
            _check_cloth(f"one launch cloth {k}", u, rest, F, X, W.K_BEND, one[k], k)
        for step, (split, schedule) in enumerate([(True, "queue"), (True, "static"), (False, "strided")]):
            u.retune(split=split, schedule=schedule)
            assert u.num_launches() == (2 if split else 1)
            # new positions, so every output must be rewritten by the retuned launches
            Xs = [W.cloth(s, seed=260 + 10 * step + s)[2] for s in sides]
            u.execute(*[W.to_colmajor(X) for X in Xs])
            for k, (rest, F, _) in enumerate(cloths):
                _check_cloth(f"retuned split={split} {schedule} cloth {k}", u, rest, F, Xs[k], W.K_BEND,
                             u.get_results(k), k)


@pytest.mark.parametrize("split", [False, True])
def test_fused_graph_replays_match_eager(require_device, split):
    sides = [12, 30, 7]
    cloths = [W.cloth(s, seed=300 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               split=split) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        eager = [u.get_results(k) for k in range(len(sides))]
        for k, (rest, F, X) in enumerate(cloths):
            _check_cloth(f"graph test eager cloth {k}", u, rest, F, X, W.K_BEND, eager[k], k)
        for rep in range(4):
            u.run_graph()  # null stream; get_results synchronises the device
            for k in range(len(sides)):
                assert np.array_equal(u.get_results(k), eager[k]), f"graph replay {rep} differs on cloth {k}"
            u.run()  # eager launch between replays
            for k in range(len(sides)):
                assert np.array_equal(u.get_results(k), eager[k]), f"eager launch after replay {rep} differs"


@pytest.mark.parametrize("subdiv,schedule,lds,staging", [(1, "static", 0, False), (5, "queue", 0, False),
                                                         (5, "strided", 65536, False), (2, "static", 0, True),
                                                         (5, "queue", 0, True)])
def test_ata_plus_a_product_tiles(require_device, subdiv, schedule, lds, staging):
    """Product tiles staged by direct loads to LDS and, as the fallback, through registers."""
    n, cp, ri, a = W.c1_matrix(subdiv=subdiv)
    with Unit.ata_plus_a(cp, ri, fused=True, schedule=schedule, tile_lds_bytes=lds, register_staging=staging) as u:
        assert u.info()["tiled"] == 2
        assert ("global_load_lds" in u.source()) == (not staging)
        u.compile().execute(a)
        _check(f"A^T A + A icosphere({subdiv}) product tiles {schedule} staging={'registers' if staging else 'direct'}",
               u, O.ata_plus_a(cp, ri, a), u.get_results())
        a2 = W.random_values(len(ri), seed=77)
        u.execute(a2)
        _check(f"A^T A + A icosphere({subdiv}) product tiles second values", u, O.ata_plus_a(cp, ri, a2),
               u.get_results())


def test_sphere_mtx_golden_product_tiles(require_device):
    n, cp, ri, a = W.load_mtx(os.path.join(ROOT, "tests", "golden", "sphere.mtx"))
    g = np.load(os.path.join(ROOT, "tests", "golden", "sphere_ata_plus_a.npz"))
    with Unit.ata_plus_a(cp, ri, fused=True) as u:
        u.compile().execute(a)
        _, _, _, sc = O.ata_plus_a(cp, ri, a)
        _check("sphere.mtx A^T A + A product tiles", u, (g["indptr"], g["indices"], g["values"], sc), u.get_results())


@pytest.mark.parametrize("rank,lds", [(0, 32768), (20, 65536)])
def test_synthetic_pattern_product_tiles(require_device, rank, lds):
    # power-law rows, band plus random fill: the cost-model patterns (smallest sizes)
    idx = int(np.argsort(W.c5_sizes())[rank])
    n, cp, ri, a = W.c5_pattern(idx)
    with Unit.ata_plus_a(cp, ri, fused=True, tile_lds_bytes=lds, schedule="queue") as u:
        u.compile().execute(a)
        _check(f"synthetic pattern {idx} (n={n}) product tiles", u, O.ata_plus_a(cp, ri, a), u.get_results())


def _tuned_c4_options():
    with open(os.path.join(ROOT, "symboliclib_amd", "tuning", "c4_selected.json")) as f:
        t = json.load(f)["selected"]
    return dict(fused=True, tile_lds_bytes=t["lds"], output_groups=t["groups"], waves_per_eu=t["wpe"],
                threads_per_block=t["tpb"], schedule=t["schedule"], split=bool(t.get("split", 0)))


@pytest.mark.parametrize("kb", BENDS)
def test_c4_full_batch(require_device, kb):
    # the benchmark's batch: 256 cloths, sides log-uniform 32..256, in the tuned fused form,
    # replayed from a HIP graph; every cloth checked against the oracle
    cloths = W.c4_cloths()
    sides = W.c4_sides()
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, kb,
                               **_tuned_c4_options()) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        u.run_graph()
        got = [u.get_results(k) for k in range(len(cloths))]

        def one(k):
            rest, F, X = cloths[k]
            et, ek, bend = _cloth_errors(f"C4 cloth {k}", u, rest, F, X, W.K_STRETCH, W.K_SHEAR, kb, got[k], k)
            j = int(et.argmax())
            return k, et, ek, bend, j

        with ThreadPoolExecutor(8) as ex:
            res = list(ex.map(one, range(len(cloths))))
    worst = sorted(res, key=lambda r: -r[1].max())
    allt = np.concatenate([r[1] for r in res])
    allk = np.concatenate([r[2] for r in res])
    print(f"\nC4 full batch k_bend={kb}: {len(cloths)} cloths, nnz={len(allt)}, max_rel_err={allt.max():.3e} "
          f"median={np.median(allt):.3e} | contribution scale max={allk.max():.3e} median={np.median(allk):.3e}")
    for k, et, ek, bend, j in worst[:5]:
        n = len(cloths[k][0])
        print(f"  cloth {k} side {sides[k]}: {_cloth_summary(et, ek, bend)}; worst output {j} "
              f"(bend={bool(bend[j])}, contribution-scale error {ek[j]:.3e})")
    per_cloth_max = np.array([r[1].max() for r in res])
    print(f"  per-cloth max: median {np.median(per_cloth_max):.3e}, max {per_cloth_max.max():.3e}")
    # the GPU's own error, against the same closed forms in 80-bit extended precision, on the
    # cloths with the largest errors and on the smallest and largest cloth
    sample = sorted({r[0] for r in worst[:4]} | {int(np.argmin(sides)), int(np.argmax(sides))})
    for k in sample:
        rest, F, X = cloths[k]
        _, _, _, st, _, _ = O.bw_hessian_check(rest, F, X, W.K_STRETCH, W.K_SHEAR, kb)
        vl = O.bw_hessian_extended(rest, F, X, W.K_STRETCH, W.K_SHEAR, kb)
        ex = np.abs((got[k] - vl).astype(np.float64)) / st
        print(f"  cloth {k} side {sides[k]} against 80-bit evaluation: max {ex.max():.3e} median {np.median(ex):.3e}")
        assert ex.max() <= TOL
    assert allt.max() <= TOL


@pytest.mark.parametrize("rank", [32, 63])
def test_c5_large_patterns_product_tiles(require_device, rank):
    # the cost-model patterns at their middle and largest size (n = 101,561 and 996,213)
    idx = int(np.argsort(W.c5_sizes())[rank])
    n, cp, ri, a = W.c5_pattern(idx)
    with Unit.ata_plus_a(cp, ri, fused=True, tile_lds_bytes=65536, schedule="queue") as u:
        u.compile().execute(a)
        vals = u.get_results()
        _check(f"synthetic pattern {idx} (n={n}, nnz(A)={len(ri)}) product tiles", u, O.ata_plus_a(cp, ri, a), vals)


def test_device_io_on_a_side_stream(require_device):
    # inputs from device memory, the launch and the result copy all on one non-blocking stream
    from hip_runtime import DeviceArray, Stream

    rest, F, X = W.c3_cloth(side=40)
    X2 = W.cloth(40, seed=17)[2]
    side = Stream()
    with Unit.bw_hessian_batch([(rest, F)], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.compile()
        n = u.output_shape(0)[2]
        xin, out = DeviceArray(3 * len(rest)), DeviceArray(n)
        for pos in (X, X2):
            xin.upload(W.to_colmajor(pos))
            out.upload(np.full(n, np.nan))
            u.execute_device(xin.ptr, stream=side.handle)
            u.run(side.handle)
            u.get_results_device(0, out.ptr, stream=side.handle)
            side.synchronize()
            _check_cloth("device I/O on a side stream", u, rest, F, pos, W.K_BEND, out.download())
        # host inputs ordered on the stream, then a graph replay on it
        u.execute_async(W.to_colmajor(X), stream=side.handle)
        u.run_graph(side.handle)
        side.synchronize()
        _check_cloth("stream-ordered host inputs, graph replay", u, rest, F, X, W.K_BEND, u.get_results())
        xin.free()
        out.free()
    side.destroy()


def test_execute_after_run_on_a_nonblocking_stream(require_device):
    # execute() copies host inputs while launches of the program may still run on a
    # non-blocking stream: the copy waits for them, and the next launch sees the new inputs
    from hip_runtime import Stream

    rest, F, X = W.c3_cloth(side=128)
    X2 = W.cloth(128, seed=23)[2]
    side = Stream()
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.compile().execute(W.to_colmajor(X))
        for _ in range(20):
            u.run(side.handle)
        u.execute(W.to_colmajor(X2))
        u.run(side.handle)
        _check_cloth("execute after queued launches", u, rest, F, X2, W.K_BEND, u.get_results())
        u.retune(schedule="strided")
        u.execute(W.to_colmajor(X))
        for _ in range(5):
            u.run(side.handle)
        u.retune(schedule="static")
        u.run(side.handle)
        _check_cloth("retune after queued launches", u, rest, F, X, W.K_BEND, u.get_results())
    side.destroy()
