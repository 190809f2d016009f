"""GPU parity: programs compiled by the new HIP emitter and run through the C ABI against the
FP64 oracle.

Bar: output patterns (CSC outer/inner arrays) bit-exact; every value within relative error
1e-12 of the oracle, the error of a value v being |v - v_oracle| / S(v) where S(v) is the sum of
the absolute magnitudes of the terms v is made of: the oracle's closed form evaluated with the
absolute value of every factor and of every additive term (products A(k,i) A(k,j) and A(i,j);
|e0_k e1_k| / |e0 x e1| per corner weight; for the cloth, the stretch and shear terms of every
per-face Hessian entry from the edge vectors on, plus |k_bend Q| per bending constant).  The
maximum and median errors are printed for every case.
"""
import os

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


@pytest.mark.parametrize("side,share,tiled", [(4, False, True), (4, True, True), (4, True, False), (32, True, True),
                                              (32, True, False), (256, True, True)])
def test_bw_hessian(require_device, side, share, tiled):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, share_templates=share, tiled=tiled) as u:
        assert u.info()["tiled"] == int(tiled), u.info()["tiling"]
        u.compile().execute(W.to_colmajor(X))
        ref = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
        _check(f"Baraff-Witkin {side}x{side} share={share}", u, ref, u.get_results())


def test_bw_hessian_stretch_shear_only(require_device):
    rest, F, X = W.cloth(9, seed=4, fold=0.3, noise=1e-2)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, 0.0) as u:
        u.compile().execute(W.to_colmajor(X))
        _check("Baraff-Witkin 9x9 no bend", u, O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, 0.0), u.get_results())


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
        ref = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
        _check(f"Baraff-Witkin 24x24 tile LDS {lds}", u, ref, u.get_results())


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
        ref = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
        _check(f"Baraff-Witkin {side}x{side} direct assembly tiled={tiled}", u, ref, u.get_results())


@pytest.mark.parametrize("tiled", FORMS)
def test_bw_hessian_batch(require_device, tiled):
    sides = [5, 17, 32, 9, 40]
    cloths = [W.cloth(s, seed=100 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, tiled=tiled) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            ref = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
            _check(f"batch cloth {k} ({sides[k]}x{sides[k]}) tiled={tiled}", u, ref, u.get_results(k), k)


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


@pytest.mark.parametrize("side,schedule,lds,groups,wpe,split", FUSED_CASES)
def test_bw_hessian_fused(require_device, side, schedule, lds, groups, wpe, split):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, schedule=schedule, tile_lds_bytes=lds,
                         output_groups=groups, waves_per_eu=wpe, split=split) as u:
        assert u.info()["tiled"] == 2, u.info()["tiling"]
        u.compile().execute(W.to_colmajor(X))
        assert u.num_launches() == (2 if split else 1)
        ref = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
        _check(f"Baraff-Witkin {side}x{side} fused {schedule} lds={lds} groups={groups} wpe={wpe} split={split}", u,
               ref, u.get_results())
        # a second evaluation of the same launch (persistent work-queue counters start from zero again)
        X2 = W.cloth(side, seed=7)[2]
        u.execute(W.to_colmajor(X2))
        _check(f"Baraff-Witkin {side}x{side} fused second positions", u,
               O.bw_hessian(rest, F, X2, W.K_STRETCH, W.K_SHEAR, W.K_BEND), u.get_results())


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
        _check(f"Baraff-Witkin 48x48 fused strided tpb={tpb}", u,
               O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND), u.get_results())


@pytest.mark.parametrize("schedule,split", [("static", False), ("queue", False), ("largest", False),
                                            ("interleaved", False), ("static", True), ("queue", True)])
def test_bw_hessian_batch_fused(require_device, schedule, split):
    sides = [5, 17, 32, 9, 40, 3]
    cloths = [W.cloth(s, seed=200 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               schedule=schedule, split=split) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            ref = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
            _check(f"fused batch {schedule} split={split} cloth {k} ({sides[k]}x{sides[k]})", u, ref, u.get_results(k), k)


@pytest.mark.parametrize("schedule", ["static", "queue"])
def test_bw_hessian_fused_tiny_cloths(require_device, schedule):
    # one-tile instances: two triangles (a single interior edge) and a 3x3 sheet, batched
    sides = [2, 3, 2]
    cloths = [W.cloth(s, seed=400 + k) for k, s in enumerate(sides)]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               schedule=schedule, output_groups=2) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        for k, (rest, F, X) in enumerate(cloths):
            _check(f"tiny fused {schedule} cloth {k} ({sides[k]}x{sides[k]})", u,
                   O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND), u.get_results(k), k)


@pytest.mark.parametrize("n,schedule", [(10, "static"), (18, "static"), (18, "queue")])
def test_bw_hessian_fused_long_contribution_lists(require_device, n, schedule):
    # the center of an n-triangle fan sums n element contributions: 5 slot pairs (n = 10, the
    # most the unrolled sums take) or 9 (n = 18, the per-lane loop); then fans batched with a
    # cloth (the loop)
    rest, F, X = W.wheel(n, seed=5)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, schedule=schedule,
                         output_groups=2) as u:
        assert ("q < k; q += 2u" in u.source()) == (n > 10)
        u.compile().execute(W.to_colmajor(X))
        _check(f"fan of {n} fused {schedule}", u, O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND),
               u.get_results())
    meshes = [W.wheel(18, seed=6), W.cloth(9, seed=8), W.wheel(7, seed=9)]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in meshes], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               schedule=schedule, output_groups=2) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in meshes])
        for k, (rest, F, X) in enumerate(meshes):
            _check(f"fans and a cloth fused {schedule}, mesh {k}", u,
                   O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND), u.get_results(k), k)


def test_fused_retune_between_one_and_two_launches(require_device):
    sides = [20, 7, 33]
    cloths = [W.cloth(s, seed=250 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        one = [u.get_results(k) for k in range(len(sides))]
        for k, (rest, F, X) in enumerate(cloths):
            _check(f"one launch cloth {k}", u, O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND), one[k], k)
        for step, (split, schedule) in enumerate([(True, "queue"), (True, "static"), (False, "strided")]):
            u.retune(split=split, schedule=schedule)
            assert u.num_launches() == (2 if split else 1)
            # new positions, so every output must be rewritten by the retuned launches
            Xs = [W.cloth(s, seed=260 + 10 * step + s)[2] for s in sides]
            u.execute(*[W.to_colmajor(X) for X in Xs])
            for k, (rest, F, _) in enumerate(cloths):
                _check(f"retuned split={split} {schedule} cloth {k}", u,
                       O.bw_hessian(rest, F, Xs[k], W.K_STRETCH, W.K_SHEAR, W.K_BEND), u.get_results(k), k)


@pytest.mark.parametrize("split", [False, True])
def test_fused_graph_replays_match_eager(require_device, split):
    sides = [12, 30, 7]
    cloths = [W.cloth(s, seed=300 + s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               split=split) as u:
        u.compile().execute(*[W.to_colmajor(X) for _, _, X in cloths])
        eager = [u.get_results(k) for k in range(len(sides))]
        for k, (rest, F, X) in enumerate(cloths):
            _check(f"graph test eager cloth {k}", u, O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND), eager[k], k)
        for rep in range(4):
            u.run_graph()  # null stream; get_results synchronises the device
            for k in range(len(sides)):
                assert np.array_equal(u.get_results(k), eager[k]), f"graph replay {rep} differs on cloth {k}"
            u.run()  # eager launch between replays
            for k in range(len(sides)):
                assert np.array_equal(u.get_results(k), eager[k]), f"eager launch after replay {rep} differs"


@pytest.mark.parametrize("subdiv,schedule,lds", [(1, "static", 0), (5, "queue", 0), (5, "strided", 65536)])
def test_ata_plus_a_product_tiles(require_device, subdiv, schedule, lds):
    n, cp, ri, a = W.c1_matrix(subdiv=subdiv)
    with Unit.ata_plus_a(cp, ri, fused=True, schedule=schedule, tile_lds_bytes=lds) as u:
        assert u.info()["tiled"] == 2
        u.compile().execute(a)
        _check(f"A^T A + A icosphere({subdiv}) product tiles {schedule}", u, O.ata_plus_a(cp, ri, a), u.get_results())
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
