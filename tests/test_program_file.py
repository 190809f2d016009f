"""Programs on disk (include/symlib.h: symlib_unit_save / symlib_unit_load; reference: the table
files of src/ComputeUnit.cpp:364-366 read back by src/StringResources.h:233-269).

A saved and reloaded unit is the same program: same emitted source, launches, tables, tile
pools and output patterns; a reloaded fused unit emits the same source for every retune (its
templates came back); results are bit-identical (CPU device here, the GPU in the gpu test).
Damaged files are refused with a message."""
import os

import numpy as np
import pytest

import oracle as O
from symboliclib_amd import _native
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _same_program(a, b):
    ia, ib = a.info(), b.info()
    for k in ("num_inputs", "num_intermediates", "num_index", "num_const", "num_outpos", "num_results", "num_kernels",
              "num_waves", "source_bytes", "tiled", "tile_pool_entries", "tiling"):
        assert ia[k] == ib[k], (k, ia[k], ib[k])
    assert a.source() == b.source()
    assert a.input_lengths == b.input_lengths
    assert _native.lib().symlib_unit_num_outputs(a._h) == _native.lib().symlib_unit_num_outputs(b._h)
    for k in range(_native.lib().symlib_unit_num_outputs(a._h)):
        oa, na = a.output_pattern(k)
        ob, nb = b.output_pattern(k)
        assert np.array_equal(oa, ob) and np.array_equal(na, nb)


def _cases():
    n, cp, ri, a = W.load_mtx(os.path.join(ROOT, "tests", "golden", "sphere.mtx"))
    V, F = W.c2_mesh(side=9)
    rest, Fc, X = W.c3_cloth(side=7)
    cl = [W.cloth(s, seed=30 + s) for s in (6, 9)]
    return {
        "ata_waves": lambda: Unit.ata_plus_a(cp, ri),
        "ata_tiled": lambda: Unit.ata_plus_a(cp, ri, tiled=True),
        "ata_product_tiles": lambda: Unit.ata_plus_a(cp, ri, fused=True),
        "cotan_trees": lambda: Unit.cotan(len(V), F),
        "cotan_fused": lambda: Unit.cotan(len(V), F, fused=True, output_groups=2),
        "cloth_direct": lambda: Unit.bw_hessian(rest, Fc, W.K_STRETCH, W.K_SHEAR, W.K_BEND, direct_assembly=True),
        "cloth_batch_fused": lambda: Unit.bw_hessian_batch([(r, f) for r, f, _ in cl], W.K_STRETCH, W.K_SHEAR,
                                                           W.K_BEND, fused=True, schedule="largest"),
    }


@pytest.mark.parametrize("name", list(_cases()))
def test_saved_program_is_the_same_program(tmp_path, name):
    with _cases()[name]() as u:
        path = str(tmp_path / f"{name}.symprog")
        u.save(path)
        with Unit.load(path) as v:
            _same_program(u, v)
            if u.info()["tiled"] == 2 and name != "ata_product_tiles":
                # templates and tile order came back: every retune emits the same source
                for opt in (dict(output_groups=3, waves_per_eu=4), dict(schedule="static", threads_per_block=512)):
                    assert u.fused_source(**opt) == v.fused_source(**opt)
                assert np.array_equal(u.tile_stats(), v.tile_stats())


@pytest.mark.parametrize("direct", [False, True])
def test_cpu_device_program_reloads_with_identical_results(tmp_path, direct):
    rest, F, X = W.c3_cloth(side=6)
    path = str(tmp_path / "cloth_cpu.symprog")
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, cpu_threads=2, direct_assembly=direct) as u:
        u.compile().execute(W.to_colmajor(X))
        ref = u.get_results()
        u.save(path)
    with Unit.load(path, cpu_threads=2) as v:
        v.compile().execute(W.to_colmajor(X))
        assert np.array_equal(v.get_results(), ref)
    # the device kind is part of the program
    with pytest.raises(_native.SymlibError, match="CPU device"):
        Unit.load(path)


def test_damaged_program_files_are_refused(tmp_path):
    rest, F, _ = W.c3_cloth(side=5)
    path = str(tmp_path / "p.symprog")
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.save(path)
    data = open(path, "rb").read()
    for name, blob, msg in (("truncated", data[: len(data) // 2], "truncated|longer than the file"),
                            ("magic", b"NOTAPROG" + data[8:], "not a program file"),
                            ("tail", data[:-8] + b"XXXXXXXX", "end marker")):
        bad = str(tmp_path / f"{name}.symprog")
        open(bad, "wb").write(blob)
        with pytest.raises(_native.SymlibError, match=msg):
            Unit.load(bad)
    with pytest.raises(_native.SymlibError, match="cannot read"):
        Unit.load(str(tmp_path / "missing.symprog"))


def test_cached_builds_once(tmp_path):
    rest, F, _ = W.c3_cloth(side=5)
    calls = []

    def build():
        calls.append(1)
        return Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True)

    key = b"cloth 5x5 " + rest.tobytes() + F.tobytes()
    with Unit.cached(str(tmp_path), key, build) as a, Unit.cached(str(tmp_path), key, build) as b:
        assert len(calls) == 1
        _same_program(a, b)


@pytest.mark.gpu
def test_reloaded_fused_batch_gives_identical_results(require_device, tmp_path):
    cl = [W.cloth(s, seed=50 + s) for s in (40, 33, 57, 21)]
    inputs = [W.to_colmajor(X) for _, _, X in cl]
    path = str(tmp_path / "batch.symprog")
    with Unit.bw_hessian_batch([(r, f) for r, f, _ in cl], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True) as u:
        u.compile().execute(*inputs)
        ref = [u.get_results(k) for k in range(len(cl))]
        u.save(path)
    with Unit.load(path) as v:
        v.compile().execute(*inputs)
        v.run_graph()
        for k in range(len(cl)):
            assert np.array_equal(v.get_results(k), ref[k]), f"cloth {k}"
        v.retune(output_groups=3, schedule="largest")
        v.run()
        for k, (rest, F, X) in enumerate(cl):
            cp, ri, rv, sc = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)
            err = O.rel_errors(v.get_results(k), rv, sc)
            assert err.max() <= 1e-12, f"cloth {k} after retune: {err.max():.3e}"
