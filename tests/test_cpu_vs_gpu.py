"""The GPU against the reference's CPU path restated (the CPU device: Device(VecWidth(4),
NumThreads(n)), the same program as host C++ with OpenMP, compiled with the reference's flags,
-ffast-math included), on the same inputs: patterns identical, every value within 1e-12 of the
oracle's term scale of the other (the CPU side may reassociate under -ffast-math, so the
comparison is not bit-for-bit)."""
import numpy as np
import pytest

import oracle as O
from symboliclib_amd import workloads as W
from symboliclib_amd.unit import Unit

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _pair(build, inputs, ref_scale, name):
    with build(cpu_threads=4) as c, build() as g:
        c.compile().execute(*inputs)
        g.compile().execute(*inputs)
        for k in range(len(c.input_lengths) if name.startswith("batch") else 1):
            oc, ic = c.output_pattern(k)
            og, ig = g.output_pattern(k)
            assert np.array_equal(oc, og) and np.array_equal(ic, ig), f"{name} {k}: patterns differ"
            err = O.rel_errors(g.get_results(k), c.get_results(k), ref_scale[k])
            print(f"{name} output {k}: GPU vs CPU device max {err.max():.3e} median {np.median(err):.3e}", flush=True)
            assert err.max() <= TOL, f"{name} {k}: {err.max():.3e}"


def test_ata_plus_a_gpu_matches_cpu_device(require_device):
    n, cp, ri, a = W.c1_matrix(subdiv=3)
    sc = O.ata_plus_a(cp, ri, a)[3]
    _pair(lambda **o: Unit.ata_plus_a(cp, ri, **o), [a], [sc], "A^T A + A icosphere(3)")


@pytest.mark.parametrize("direct", [False, True])
def test_cotan_gpu_matches_cpu_device(require_device, direct):
    V, F = W.c2_mesh(side=48)
    sc = O.cotan(V, F)[3]
    _pair(lambda **o: Unit.cotan(len(V), F, direct_assembly=direct, **o), [W.to_colmajor(V)], [sc],
          f"cotan 48x48 direct={direct}")


@pytest.mark.parametrize("direct", [False, True])
def test_bw_hessian_gpu_matches_cpu_device(require_device, direct):
    rest, F, X = W.c3_cloth(side=24)
    sc = O.bw_hessian(rest, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)[3]
    _pair(lambda **o: Unit.bw_hessian(rest, F, W.K_STRETCH, W.K_SHEAR, W.K_BEND, direct_assembly=direct, **o),
          [W.to_colmajor(X)], [sc], f"cloth 24x24 direct={direct}")


def test_bw_batch_gpu_matches_cpu_device(require_device):
    cl = [W.cloth(s, seed=60 + s) for s in (9, 17, 12)]
    scs = [O.bw_hessian(r, F, X, W.K_STRETCH, W.K_SHEAR, W.K_BEND)[3] for r, F, X in cl]
    _pair(lambda **o: Unit.bw_hessian_batch([(r, F) for r, F, _ in cl], W.K_STRETCH, W.K_SHEAR, W.K_BEND, **o),
          [W.to_colmajor(X) for _, _, X in cl], scs, "batch of 3 cloths")
