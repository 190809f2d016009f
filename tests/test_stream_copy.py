"""The STREAM copy kernel that bench.py uses as the measured HBM ceiling (symhip_stream_copy):
both forms copy every element exactly, odd lengths included (the last element goes through
the tail kernel), and misaligned buffers are refused."""
import ctypes

import numpy as np
import pytest

from symboliclib_amd import _native

from hip_runtime import DeviceArray, Stream

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 513, 1 << 20, (1 << 20) + 7])
@pytest.mark.parametrize("blocks", [0, 3, 256])
def test_stream_copy_copies_every_element(require_device, n, blocks):
    L = _native.lib()
    src, dst = DeviceArray(n), DeviceArray(n)
    stream = Stream()
    try:
        a = np.random.default_rng(n + blocks).standard_normal(n)
        src.upload(a)
        dst.upload(np.zeros(n))
        print(f"stream copy n={n} blocks={blocks}", flush=True)
        _native.check(L.symhip_stream_copy(dst.ptr, src.ptr, n, blocks, stream.handle), "symhip_stream_copy")
        stream.synchronize()
        assert np.array_equal(dst.download(), a)
    finally:
        src.free()
        dst.free()
        stream.destroy()


def test_stream_copy_refuses_misaligned_buffers(require_device):
    L = _native.lib()
    src, dst = DeviceArray(16), DeviceArray(16)
    try:
        rc = L.symhip_stream_copy(ctypes.c_void_p(dst.ptr.value + 8), src.ptr, 8, 0, None)
        assert rc != 0 and b"aligned" in L.symhip_last_error()
    finally:
        src.free()
        dst.free()
