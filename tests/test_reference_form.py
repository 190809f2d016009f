"""Programs in the reference's own kernel form run through the C ABI (include/symhip.h:
symhip_set_launch_offsets, symhip_set_result_gather).

The reference emits every group kernel as k<i>(const RealT* x, const RealT* c, RealT* y,
const unsigned* p), launches it as k<i>(x, c + constantsOffset[i], x + outputOffset[i],
p + indexOffset[i]) with grid 1 + (n - 1) / T (src/StringResources.h:196,
src/ComputeUnit.cpp:412-421) and gathers results[j] = x[outIds[j]] afterwards
(src/StringResources.h:218-221).  The program below is written in that form by hand: two
kernels, the second reading the intermediates of the first, tables laid out as the
reference lays them out (SoA per kernel, one offset per kernel into each table), two output
containers gathered from x.  Values are checked against numpy within the 1e-12 term-scale
bar (hipRTC may contract a * c + b into one fma)."""
import ctypes
import os

import numpy as np
import pytest

from symboliclib_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
N_IN, N0, N1, N_OUT1 = 1000, 1000, 700, 300
T = 128

# the kernels as the reference emits them (after its header)
KERNELS = f"""
KERNEL void k0(const RealT* restrict x, const RealT* restrict c, RealT* y, const unsigned int* restrict p) {{
	const unsigned int i = hipThreadIdx_x + hipBlockDim_x * hipBlockIdx_x;

	if(i < {N0}) {{
		y[i] = x[p[i]] * c[i] + x[p[{N0} + i]];
		y[{N0} + i] = x[p[i]] - c[{N0} + i];
	}}
}}

KERNEL void k1(const RealT* restrict x, const RealT* restrict c, RealT* y, const unsigned int* restrict p) {{
	const unsigned int i = hipThreadIdx_x + hipBlockDim_x * hipBlockIdx_x;

	if(i < {N1}) {{
		const RealT r0 = x[p[i]];
		y[i] = sqrt(r0 * r0 + 1.0) * c[i];
	}}
}}
"""


def _source():
    """The runtime's prelude (KERNEL, RealT, INDEX, restrict, the HIP index names) + the kernels."""
    return _native.lib().symhip_reference_prelude().decode() + KERNELS


def _program(seed=0):
    """Tables of the two-kernel program in the reference's layout."""
    rng = np.random.default_rng(seed)
    a_start, b_start, slots0, slots1 = 0, N_IN, 2 * N_IN, 2 * N_IN + 2 * N0
    num_values = slots1 + N1
    perm = rng.permutation(N_IN)
    reads1 = rng.integers(0, 2 * N0, N1)
    index = np.concatenate([np.arange(N0), b_start + perm, slots0 + reads1]).astype(np.uint32)
    consts = rng.uniform(-2, 2, 2 * N0 + N1)
    out0 = slots1 + rng.permutation(N1)  # container 0: k1's values, permuted
    out1 = slots0 + rng.choice(2 * N0, N_OUT1, replace=False)  # container 1: some of k0's
    outpos = np.concatenate([out0, out1]).astype(np.uint32)
    offsets = [(0, slots0, 0), (2 * N0, slots1, 2 * N0)]  # (constants, outputs, index) per kernel
    return dict(num_values=num_values, args=[(a_start, N_IN), (b_start, N_IN)], outputs=[(0, N1), (N1, N_OUT1)],
                index=index, consts=consts, outpos=outpos, offsets=offsets, perm=perm, reads1=reads1, out0=out0,
                out1=out1)


def _expected(P, a, b):
    c = P["consts"]
    y0 = a * c[:N0] + b[P["perm"]]
    y1 = a - c[N0:2 * N0]
    s0 = np.abs(a * c[:N0]) + np.abs(b[P["perm"]])
    s1 = np.abs(a) + np.abs(c[N0:2 * N0])
    k0 = np.concatenate([y0, y1])
    k0s = np.concatenate([s0, s1])
    r = k0[P["reads1"]]
    v1 = np.sqrt(r * r + 1.0) * c[2 * N0:]
    s_1 = np.sqrt(r * r + 1.0) * np.abs(c[2 * N0:]) + np.abs(c[2 * N0:]) * k0s[P["reads1"]]
    x = np.zeros(P["num_values"])
    sx = np.zeros(P["num_values"])
    x[2 * N_IN:2 * N_IN + 2 * N0], sx[2 * N_IN:2 * N_IN + 2 * N0] = k0, k0s
    x[2 * N_IN + 2 * N0:], sx[2 * N_IN + 2 * N0:] = v1, s_1
    return [(x[P["out0"]], sx[P["out0"]]), (x[P["out1"]], sx[P["out1"]])]


def test_reference_form_source_compiles_for_gfx950(tmp_path):
    """hipRTC accepts kernels written in the reference's form (HIP index macros included)."""
    out = str(tmp_path / "refform.co")
    rc = _native.lib().symhip_compile_to_file(_source().encode(), None, b"gfx950", out.encode())
    _native.check(rc, "symhip_compile_to_file(reference form)")
    assert os.path.getsize(out) > 0


def _u64(v):
    return (ctypes.c_uint64 * len(v))(*v)


@pytest.mark.gpu
def test_reference_form_program_runs_through_the_runtime():
    L = _native.lib()
    P = _program()
    n_launch = [N0, N1]
    launches = (_native.Launch * 2)(*[_native.Launch(f"k{i}".encode(), 1 + (n - 1) // T, T, 0)
                                      for i, n in enumerate(n_launch)])
    keep = [P["index"], P["consts"], P["outpos"]]
    d = _native.ProgramDesc()
    d.num_values = P["num_values"]
    d.num_args = 2
    arg_start, arg_len = _u64([s for s, _ in P["args"]]), _u64([n for _, n in P["args"]])
    out_start, out_len = _u64([s for s, _ in P["outputs"]]), _u64([n for _, n in P["outputs"]])
    d.arg_start, d.arg_length = arg_start, arg_len
    d.num_outputs = 2
    d.out_start, d.out_length = out_start, out_len
    d.num_index, d.index_data = len(keep[0]), keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    d.num_const, d.const_data = len(keep[1]), keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    d.num_outpos, d.outpos_data = len(keep[2]), keep[2].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    d.num_launches, d.launches = 2, launches
    d.num_work_words = 0
    h = ctypes.c_void_p()
    print("reference-form program: create, compile, offsets, gather", flush=True)
    _native.check(L.symhip_create(ctypes.byref(h), ctypes.byref(d)), "symhip_create")
    try:
        _native.check(L.symhip_compile(h, _source().encode(), None), "symhip_compile")
        for i, (co, yo, po) in enumerate(P["offsets"]):
            _native.check(L.symhip_set_launch_offsets(h, i, co, yo, po), "symhip_set_launch_offsets")
        # offsets beyond a table and a gather without one position per result are refused
        assert L.symhip_set_launch_offsets(h, 0, len(P["consts"]) + 1, 0, 0) != 0
        assert L.symhip_set_launch_offsets(h, 2, 0, 0, 0) != 0
        _native.check(L.symhip_set_result_gather(h, 1), "symhip_set_result_gather")
        rng = np.random.default_rng(7)
        for call, run in enumerate((L.symhip_run, L.symhip_run_graph, L.symhip_run_graph, L.symhip_run_noalloc)):
            a, b = rng.uniform(-1, 1, N_IN), rng.uniform(-1, 1, N_IN)
            for i, v in enumerate((a, b)):
                _native.check(L.symhip_set_arg(h, i, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(v)),
                              "symhip_set_arg")
            _native.check(run(h, None), f"run {call}")
            for k, (want, scale) in enumerate(_expected(P, a, b)):
                got = np.empty(len(want))
                _native.check(L.symhip_get_result(h, k, got.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                  len(got)), "symhip_get_result")
                err = np.abs(got - want) / np.maximum(scale, np.finfo(float).tiny)
                assert err.max() <= 1e-12, f"call {call} output {k}: max relative error {err.max():.3e}"
    finally:
        L.symhip_destroy(h)
