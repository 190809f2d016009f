"""Host-side checks of the symbolic front-end, compiler middle-end, emitter and C ABI (no GPU).

Output patterns are compared bit-exactly with the oracle; emitted programs are compiled by
hipRTC for gfx950 (no device needed); error paths must return codes with messages.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from symboliclib_amd import _native
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = _native.lib()
    declared = []
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        declared += re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(sym\w+)\s*\(", text, flags=re.M)
    assert len(declared) >= 30
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    assert set(declared) <= set(_native.SIGNATURES), set(declared) - set(_native.SIGNATURES)


def test_tutorial_known_answers():
    # examples/Tutorial/main.cpp:27-33: 3.82843, 2.41421, 2.12132 = 1 + 2 sqrt 2, 1 + sqrt 2, 3 / sqrt 2
    out = (ctypes.c_double * 3)()
    _native.check(_native.lib().symlib_tutorial_known_answers(out), "known answers")
    assert out[0] == pytest.approx(1 + 2 * 2 ** 0.5, rel=1e-15)
    assert out[1] == pytest.approx(1 + 2 ** 0.5, rel=1e-15)
    assert out[2] == pytest.approx(3 / 2 ** 0.5, rel=1e-15)


@pytest.mark.parametrize("subdiv", [0, 1, 3])
def test_ata_plus_a_pattern(subdiv):
    n, cp, ri, a = W.c1_matrix(subdiv=subdiv)
    with Unit.ata_plus_a(cp, ri) as u:
        outer, inner = u.output_pattern()
        bcp, bri, _, _ = O.ata_plus_a(cp, ri, a)
        assert np.array_equal(outer, bcp) and np.array_equal(inner, bri)


def test_sphere_mtx_pattern():
    n, cp, ri, a = W.load_mtx(os.path.join(ROOT, "tests", "golden", "sphere.mtx"))
    g = np.load(os.path.join(ROOT, "tests", "golden", "sphere_ata_plus_a.npz"))
    with Unit.ata_plus_a(cp, ri) as u:
        outer, inner = u.output_pattern()
        assert np.array_equal(outer, g["indptr"]) and np.array_equal(inner, g["indices"])


@pytest.mark.parametrize("side", [3, 8, 33])
def test_cotan_pattern(side):
    V, F = W.c2_mesh(side=side)
    with Unit.cotan(len(V), F) as u:
        outer, inner = u.output_pattern()
        cp, ri = O.cotan_pattern(len(V), F)
        assert np.array_equal(outer, cp) and np.array_equal(inner, ri)
        info = u.info()
        # per-face area and per-corner weight become shared intermediates: three waves
        assert info["num_waves"] == 3
        assert info["num_intermediates"] == len(F) + 3 * len(F)


@pytest.mark.parametrize("side,share", [(3, True), (5, False), (12, True)])
def test_bw_pattern(side, share):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, share_templates=share) as u:
        outer, inner = u.output_pattern()
        cp, ri = O.bw_pattern(len(rest), F)
        assert np.array_equal(outer, cp) and np.array_equal(inner, ri)
        info = u.info()
        if share:
            # one 45-entry upper-triangle block per face
            assert info["num_intermediates"] == 45 * len(F)


def test_emitted_sources_compile_for_gfx950(tmp_path):
    L = _native.lib()
    n, cp, ri, a = W.c1_matrix(subdiv=1)
    V, F2 = W.c2_mesh(side=5)
    rest, F3, X = W.c3_cloth(side=4)
    units = [Unit.ata_plus_a(cp, ri), Unit.cotan(len(V), F2), Unit.bw_hessian(rest, F3, 1e3, 1e2, 1e-2)]
    for k, u in enumerate(units):
        src = u.source()
        out = str(tmp_path / f"prog{k}.co")
        rc = L.symhip_compile_to_file(src.encode(), None, b"gfx950", out.encode())
        assert rc == 0, L.symhip_last_error().decode()
        assert os.path.getsize(out) > 1000
        u.close()


def test_errors_are_reported_not_thrown():
    L = _native.lib()
    h = ctypes.c_void_p()
    bad_faces = np.array([[0, 1, 7]], np.int32)
    rc = L.symlib_unit_cotan(ctypes.byref(h), 3, 1, bad_faces.ctypes.data, None)
    assert rc != 0 and b"out of range" in L.symlib_last_error()
    cp = np.array([0, 2, 3], np.int32)
    ri = np.array([1, 0, 1], np.int32)  # unsorted rows in column 0
    rc = L.symlib_unit_ata_plus_a(ctypes.byref(h), 2, cp.ctypes.data, ri.ctypes.data, None)
    assert rc != 0 and b"sorted" in L.symlib_last_error()
    # compile errors come back with the compiler log
    rc = L.symhip_compile_to_file(b"this is not C++", None, b"gfx950", b"/tmp/never_written.co")
    assert rc == -3 and len(L.symhip_last_error()) > 0
    # runtime entry points reject null handles
    assert L.symhip_run(None, None) == -1
    assert L.symhip_set_arg(None, 0, None, 0) == -1


def test_execute_before_compile_fails():
    n, cp, ri, a = W.c1_matrix(subdiv=0)
    with Unit.ata_plus_a(cp, ri) as u:
        with pytest.raises(_native.SymlibError, match="before compile"):
            u.execute(a)


def test_template_sharing_matches_direct_tables():
    rest, F, X = W.c3_cloth(side=4)
    with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, share_templates=True) as a, \
            Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, share_templates=False) as b:
        ia, ib = a.info(), b.info()
        assert ia["num_results"] == ib["num_results"]
        assert ia["num_outpos"] == ib["num_outpos"]


@pytest.mark.parametrize("side", [8, 24])
def test_tiled_form_compiles(tmp_path, side):
    L = _native.lib()
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, tiled=True) as u:
        info = u.info()
        assert info["tiled"] == 1, info["tiling"]
        fields = dict(kv.split("=") for kv in info["tiling"].split())
        assert int(fields["lds"]) <= 64 * 1024
        assert 1.0 <= float(fields["recompute"]) < 4.0
        src = u.source()
        assert "symlib_tiled" in src and "__syncthreads" in src
        out = str(tmp_path / "tiled.co")
        assert L.symhip_compile_to_file(src.encode(), None, b"gfx950", out.encode()) == 0, L.symhip_last_error()


def test_waves_form_on_request():
    V, F = W.c2_mesh(side=6)
    with Unit.cotan(len(V), F, tiled=False) as u:
        info = u.info()
        assert info["tiled"] == 0
        assert "symlib_wave2" in u.source()


def test_tiling_with_tiny_budget_falls_back():
    # an output that needs more intermediates than one tile holds cannot be tiled
    rest, F, X = W.c3_cloth(side=4)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, tiled=True, tile_lds_bytes=256) as u:
        info = u.info()
        assert info["tiled"] == 0 and info["tiling"].startswith("untiled")


def _fields(info):
    return dict(kv.split("=") for kv in info["tiling"].split()[1:])


@pytest.mark.parametrize("side,lds", [(3, 0), (16, 0), (40, 16384), (40, 65536)])
def test_fused_pattern_and_tiles(side, lds):
    rest, F, X = W.c3_cloth(side=side)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, tile_lds_bytes=lds) as u:
        info = u.info()
        assert info["tiled"] == 2 and info["tiling"].startswith("fused"), info["tiling"]
        outer, inner = u.output_pattern()
        cp, ri = O.bw_pattern(len(rest), F)
        assert np.array_equal(outer, cp) and np.array_equal(inner, ri)
        f = _fields(info)
        assert int(f["tiles"]) >= 1
        assert int(f["element_lds"]) < int(f["lds"]) <= (lds or 53760)
        # every face evaluated at least once; compact patches keep re-evaluation bounded
        assert 1.0 <= float(f["recompute"]) < 2.5
        assert info["num_results"] == len(ri)


def test_fused_batch_patterns():
    sides = [3, 11, 6]
    cloths = [W.cloth(s, seed=s) for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], 1e3, 1e2, 1e-2, fused=True, schedule="interleaved") as u:
        for k, (rest, F, _) in enumerate(cloths):
            outer, inner = u.output_pattern(k)
            cp, ri = O.bw_pattern(len(rest), F)
            assert np.array_equal(outer, cp) and np.array_equal(inner, ri)


@pytest.mark.parametrize("groups,wpe,schedule,split", [(1, 0, "static", False), (9, 5, "queue", False),
                                                      (3, 0, "queue", True), (5, 4, "static", True)])
def test_fused_source_compiles_for_gfx950(tmp_path, groups, wpe, schedule, split):
    L = _native.lib()
    rest, F, X = W.c3_cloth(side=6)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, output_groups=groups,
                         waves_per_eu=wpe, schedule=schedule, split=split) as u:
        src = u.source()
        assert "symlib_elem0_0" in src
        kernels = ["symlib_fused_elem", "symlib_fused_sum"] if split else ["symlib_fused("]
        assert all(k in src for k in kernels)
        assert ("symlib_fused_elem" in src) == split
        assert ("atomicAdd" in src) == (schedule != "static")
        # the retuned source of the other form
        other = u.fused_source(output_groups=groups, waves_per_eu=wpe, schedule=schedule, split=not split)
        assert ("symlib_fused_elem" in other) == (not split)
        out = str(tmp_path / "fused.co")
        assert L.symhip_compile_to_file(src.encode(), None, b"gfx950", out.encode()) == 0, L.symhip_last_error()


def test_fused_form_errors():
    n, cp, ri, a = W.c1_matrix(subdiv=2)
    with pytest.raises(_native.SymlibError, match="more LDS"):
        Unit.ata_plus_a(cp, ri, fused=True, tile_lds_bytes=512)
    V, F2 = W.c2_mesh(side=4)
    with pytest.raises(_native.SymlibError, match="more LDS"):
        Unit.cotan(len(V), F2, fused=True, tile_lds_bytes=128)
    rest, F, X = W.c3_cloth(side=4)
    with pytest.raises(_native.SymlibError, match="more LDS"):
        Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, tile_lds_bytes=256)
    # budgets beyond one CU's 160 KiB of LDS could never launch
    with pytest.raises(_native.SymlibError, match="160 KiB"):
        Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, tile_lds_bytes=200000)
    with pytest.raises(_native.SymlibError, match="160 KiB"):
        Unit.ata_plus_a(cp, ri, fused=True, tile_lds_bytes=200000)


def test_fused_sums_unrolled_or_looped():
    # one list per output (output_sums): contribution lists of at most 5 slot pairs: every lane
    # reads the longest length at once; the center of an 18-triangle fan sums 18 contributions
    # (9 pairs), so lists loop per lane
    rest, F, _ = W.c3_cloth(side=16)
    with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, output_groups=2, output_sums=True) as u:
        src = u.source()
        assert "np2" in src and "q < k; q += 2u" not in src
    rest, F, _ = W.wheel(18, seed=5)
    with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, output_groups=2, output_sums=True) as u:
        src = u.source()
        assert "q < k; q += 2u" in src and "np2" not in src
        cp, ri = O.bw_pattern(len(rest), F)
        outer, inner = u.output_pattern()
        assert np.array_equal(outer, cp) and np.array_equal(inner, ri)
    # the bound: 10 contributions (5 pairs) unrolled, 12 (6 pairs) looped
    for n, looped in [(10, False), (12, True)]:
        rest, F, _ = W.wheel(n, seed=5)
        with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, output_groups=2, output_sums=True) as u:
            assert ("q < k; q += 2u" in u.source()) == looped


def test_fused_block_sums_chosen_for_node_blocks():
    # cloth Hessians are made of 3 x 3 node blocks: the fused form sums whole blocks (one list per
    # block, the outputs through LDS); the 18-triangle fan's center block sums 18 entries
    for rest, F in [W.c3_cloth(side=16)[:2], W.wheel(18, seed=5)[:2]]:
        with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, output_groups=2) as u:
            src = u.source()
            assert "block records" in src and "np2" not in src and "q < k; q += 2u" not in src
            cp, ri = O.bw_pattern(len(rest), F)
            outer, inner = u.output_pattern()
            assert np.array_equal(outer, cp) and np.array_equal(inner, ri)
            # every face adds its 9 x 9 Hessian once per owning tile: 81 contributions per face
            assert u.fused_stats()["contributions"] == 81 * len(F)
    # product tiles (A^T A + A) and general constants keep one list per output
    n, cp, ri, a = W.c1_matrix(subdiv=2)
    with Unit.ata_plus_a(cp, ri, fused=True) as u:
        assert "block records" not in u.source()


def test_fused_values_within_byte_offsets():
    # one list per output: slots are u16 LDS byte offsets, so with a whole CU's LDS tiles stop
    # growing when their element values and the zero slot would pass 64 KiB
    rest, F, _ = W.c3_cloth(side=64)
    with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, tile_lds_bytes=160 * 1024, output_sums=True) as u:
        values = u.tile_stats()[:, 4]
        assert 8 * (values.max() + 1) <= 65536
        assert values.max() > 8000  # the cap, not the LDS budget, bounds the largest tiles
    # block sums index values in doubles (15 bits): the LDS budget bounds the tiles
    with Unit.bw_hessian(rest, F, 1e3, 1e2, 1e-2, fused=True, tile_lds_bytes=160 * 1024) as u:
        assert "block records" in u.source()
        values = u.tile_stats()[:, 4]
        assert 8 * values.max() > 65536 and values.max() + 9 <= 0x7fff


@pytest.mark.parametrize("subdiv", [0, 2, 4])
def test_product_tiles_pattern(subdiv):
    n, cp, ri, a = W.c1_matrix(subdiv=subdiv)
    with Unit.ata_plus_a(cp, ri, fused=True, tile_lds_bytes=16384) as u:
        info = u.info()
        assert info["tiled"] == 2
        outer, inner = u.output_pattern()
        bcp, bri, _, _ = O.ata_plus_a(cp, ri, a)
        assert np.array_equal(outer, bcp) and np.array_equal(inner, bri)
        f = _fields(info)
        assert int(f["lds"]) <= 16384
        ts = u.tile_stats()
        # every output owned by exactly one tile; every tile stages what its products read
        assert int(ts[:, 0].sum()) == len(bri)
        assert (ts[:, 3] > 0).all()


def test_synthetic_patterns_are_symmetric_with_power_law_rows():
    n, cp, ri, a = W.c5_pattern(int(np.argsort(W.c5_sizes())[0]))
    import scipy.sparse as sp

    A = sp.csc_matrix((np.ones(len(ri)), ri, cp), shape=(n, n))
    assert (A != A.T).nnz == 0 and A.diagonal().min() == 1
    deg = np.diff(cp)
    assert 4 <= np.median(deg) <= 64 and deg.max() <= 2 * 64 + 1
    with Unit.ata_plus_a(cp, ri, fused=True) as u:
        outer, inner = u.output_pattern()
        B = (A.T @ A + A).tocsc()
        B.sort_indices()
        assert np.array_equal(outer, B.indptr) and np.array_equal(inner, B.indices)


@pytest.mark.parametrize("lds", [12288, 16384, 24576, 40448, 81408])
def test_block_sum_tiles_within_the_lds_budget(lds):
    # tiles grow on an estimate of their list padding; a tile whose exact LDS passes the budget
    # is grown again smaller, so every tile fits (fans: long, odd diagonal lists)
    meshes = [W.c3_cloth(side=48)[:2], W.wheel(18, seed=5)[:2], W.wheel(7, seed=3)[:2], W.cloth(23, seed=9)[:2]]
    with Unit.bw_hessian_batch(meshes, 1e3, 1e2, 1e-2, fused=True, tile_lds_bytes=lds) as u:
        f = _fields(u.info())
        assert "block records" in u.source()
        assert int(f["lds"]) <= lds, f
        for k, (rest, F) in enumerate(meshes):
            cp, ri = O.bw_pattern(len(rest), F)
            outer, inner = u.output_pattern(k)
            assert np.array_equal(outer, cp) and np.array_equal(inner, ri)
