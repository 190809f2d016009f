"""symhip_run_noalloc (the entry point that allocates nothing) against symhip_run.

Both enqueue the same compiled kernels; the no-allocation call takes the work words of
persistent schedules from the library's static scratch array instead of the handle's buffer.
Bar: results bit-identical to symhip_run's on the same inputs, on two calls in a row whose
second program is smaller (the array keeps the first call's contents), with inputs changed
between calls so that every call has to rewrite every output.
"""
import ctypes

import numpy as np
import pytest

from symboliclib_amd import _native
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

pytestmark = pytest.mark.gpu


def _results(u, n):
    return [u.get_results(k) for k in range(n)]


def _compare(sides, schedule, seed, profile=False):
    L = _native.lib()
    cloths = [W.cloth(s, seed=seed + s) for s in sides]
    other = [W.cloth(s, seed=seed + 1000 + s)[2] for s in sides]
    with Unit.bw_hessian_batch([(r, F) for r, F, _ in cloths], W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True,
                               schedule=schedule, profile_tiles=profile) as u:
        u.compile()
        rt = u.runtime()
        first = [W.to_colmajor(X) for _, _, X in cloths]
        second = [W.to_colmajor(X) for X in other]
        u.execute(*first)
        _native.check(L.symhip_run(rt, None), "symhip_run")
        ref_first = _results(u, len(sides))
        u.execute(*second)
        _native.check(L.symhip_run(rt, None), "symhip_run")
        ref_second = _results(u, len(sides))
        for inputs, ref, label in ((first, ref_first, "first"), (second, ref_second, "second"),
                                   (first, ref_first, "first again")):
            u.execute(*inputs)
            _native.check(L.symhip_run_noalloc(rt, None), "symhip_run_noalloc")
            got = _results(u, len(sides))
            for k in range(len(sides)):
                assert np.array_equal(got[k], ref[k]), f"{label} inputs, cloth {k}: symhip_run_noalloc differs"
        print(f"\nsymhip_run_noalloc {schedule} sides={sides}: bit-identical to symhip_run", flush=True)


def test_noalloc_matches_run_then_smaller_program(require_device):
    _compare([40, 33, 25, 17], "queue", 500)   # persistent work queue: the counters live in the scratch array
    _compare([9, 12], "queue", 600)            # a smaller program next, on the same array
    _compare([7, 5], "static", 700)            # a launch that uses no work words
    _compare([20, 11], "largest", 800, profile=True)  # profiled: stamps after the counters


def test_noalloc_rejects_programs_beyond_the_scratch_array(require_device):
    # a profiled launch needs 4 + 8 words per tile: small tiles of a 256 x 256 cloth pass the
    # 16,384 words of the array
    rest, F, X = W.c3_cloth(side=256)
    with Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, fused=True, profile_tiles=True,
                         tile_lds_bytes=40448) as u:
        assert 4 + 8 * u.fused_stats()["num_tiles"] > 16384
        u.compile().execute(W.to_colmajor(X))
        rc = _native.lib().symhip_run_noalloc(u.runtime(), None)
        msg = _native.lib().symhip_last_error().decode()
        assert rc == -1 and "static scratch array" in msg, (rc, msg)


def test_noalloc_before_compile_is_an_error(require_device):
    assert _native.lib().symhip_run_noalloc(None, None) == -1
