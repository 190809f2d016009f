"""Python mirror of ComputeUnit for the three workloads, over the C ABI of include/symlib.h.

    unit = Unit.ata_plus_a(indptr, indices)          # construct: symbolic front-end + ComputeUnit
    unit.compile()                                   # hipRTC for the GPU, tables to HBM
    unit.execute(a_values)                           # inputs in constructor order
    values = unit.get_results(0)                     # outputs in constructor order

These calls map one to one onto ComputeUnit<double>(Device(UseHIP(), ThreadsPerBlock(t)),
in, out).compile().execute(in).getResults(out) of the reference (src/ComputeUnit.hpp:192-222).
"""
import ctypes
import os

import numpy as np

from . import _native
from .workloads import to_colmajor


SCHEDULES = {"static": 0, "queue": 1, "largest": 2, "interleaved": 3, "strided": 4}


def _opts(threads_per_block=256, decomposition_threshold=-1, simplify=True, share_templates=True, tiled=False,
          tile_lds_bytes=0, direct_assembly=False, fused=False, schedule="queue", output_groups=0, waves_per_eu=0,
          profile_tiles=False, split=False, register_staging=False, output_sums=False, cpu_threads=0):
    o = _native.Options()
    _native.lib().symlib_default_options(ctypes.byref(o))
    o.threads_per_block = int(threads_per_block)
    o.decomposition_threshold = int(decomposition_threshold)
    o.simplify = int(bool(simplify))
    o.share_templates = int(bool(share_templates))
    o.tiled = int(bool(tiled))
    o.tile_lds_bytes = int(tile_lds_bytes)
    o.direct_assembly = int(bool(direct_assembly))
    o.fused = int(bool(fused))
    o.schedule = SCHEDULES[schedule] if isinstance(schedule, str) else int(schedule)
    o.output_groups = int(output_groups)
    o.waves_per_eu = int(waves_per_eu)
    o.profile_tiles = int(bool(profile_tiles))
    o.split = int(bool(split))
    o.register_staging = int(bool(register_staging))
    o.output_sums = int(bool(output_sums))
    o.cpu_threads = int(cpu_threads)
    return o


class Unit:
    def __init__(self, handle, input_lengths):
        self._h = handle
        self.input_lengths = list(input_lengths)
        self._keep = []

    # ---- programs on disk
    def save(self, path):
        """Write the built program to `path` (symlib_unit_save)."""
        _native.check(_native.lib().symlib_unit_save(self._h, os.fsencode(path)), "symlib_unit_save")
        return self

    @classmethod
    def load(cls, path, cpu_threads=0):
        """The unit saved at `path`, without front-end or build (symlib_unit_load); cpu_threads > 0
        for a program saved from the CPU device."""
        L = _native.lib()
        h = ctypes.c_void_p()
        o = _opts(cpu_threads=cpu_threads)
        _native.check(L.symlib_unit_load(ctypes.byref(h), os.fsencode(path), ctypes.byref(o)), "symlib_unit_load")
        lengths = []
        for i in range(L.symlib_unit_num_inputs(h)):
            n = ctypes.c_uint64()
            _native.check(L.symlib_unit_input_length(h, i, ctypes.byref(n)), "symlib_unit_input_length")
            lengths.append(n.value)
        return cls(h, lengths)

    @classmethod
    def cached(cls, directory, key, build):
        """Unit of the program identified by `key` (bytes: a digest of everything the program
        is built from) from `directory` when saved there before, else build() saved there."""
        import hashlib

        os.makedirs(directory, exist_ok=True)
        path = os.path.join(directory, hashlib.sha256(key).hexdigest()[:32] + ".symprog")
        if os.path.exists(path):
            return cls.load(path)
        u = build()
        tmp = f"{path}.{os.getpid()}.tmp"
        u.save(tmp)
        os.replace(tmp, path)
        return u

    # ---- constructors (symbolic front-ends of the reference examples)
    @classmethod
    def ata_plus_a(cls, indptr, indices, **opt):
        L = _native.lib()
        cp = np.ascontiguousarray(indptr, np.int32)
        ri = np.ascontiguousarray(indices, np.int32)
        h = ctypes.c_void_p()
        o = _opts(**opt)
        _native.check(L.symlib_unit_ata_plus_a(ctypes.byref(h), len(cp) - 1, cp.ctypes.data, ri.ctypes.data,
                                               ctypes.byref(o)), "symlib_unit_ata_plus_a")
        return cls(h, [len(ri)])

    @classmethod
    def cotan(cls, num_vertices, faces, **opt):
        L = _native.lib()
        F = np.ascontiguousarray(faces, np.int32).reshape(-1)
        h = ctypes.c_void_p()
        o = _opts(**opt)
        _native.check(L.symlib_unit_cotan(ctypes.byref(h), num_vertices, len(F) // 3, F.ctypes.data, ctypes.byref(o)),
                      "symlib_unit_cotan")
        return cls(h, [3 * num_vertices])

    @classmethod
    def bw_hessian(cls, rest, faces, k_stretch, k_shear, k_bend, **opt):
        L = _native.lib()
        F = np.ascontiguousarray(faces, np.int32).reshape(-1)
        R = to_colmajor(rest)
        h = ctypes.c_void_p()
        o = _opts(**opt)
        _native.check(L.symlib_unit_bw_hessian(ctypes.byref(h), len(rest), len(F) // 3, F.ctypes.data, R.ctypes.data,
                                               k_stretch, k_shear, k_bend, ctypes.byref(o)), "symlib_unit_bw_hessian")
        return cls(h, [3 * len(rest)])

    @classmethod
    def bw_hessian_batch(cls, cloths, k_stretch, k_shear, k_bend, **opt):
        """cloths: list of (rest, faces); one program, cloth k = input k and output k."""
        L = _native.lib()
        Fs = [np.ascontiguousarray(F, np.int32).reshape(-1) for _, F in cloths]
        Rs = [to_colmajor(r) for r, _ in cloths]
        nv = np.array([len(r) for r, _ in cloths], np.int32)
        nf = np.array([len(F) // 3 for F in Fs], np.int32)
        fp = (ctypes.c_void_p * len(Fs))(*[F.ctypes.data for F in Fs])
        rp = (ctypes.c_void_p * len(Rs))(*[R.ctypes.data for R in Rs])
        h = ctypes.c_void_p()
        o = _opts(**opt)
        _native.check(L.symlib_unit_bw_hessian_batch(ctypes.byref(h), len(cloths), nv.ctypes.data, nf.ctypes.data,
                                                     ctypes.cast(fp, ctypes.c_void_p), ctypes.cast(rp, ctypes.c_void_p),
                                                     k_stretch, k_shear, k_bend, ctypes.byref(o)),
                      "symlib_unit_bw_hessian_batch")
        u = cls(h, [3 * int(v) for v in nv])
        u._keep = [Fs, Rs]
        return u

    # ---- ComputeUnit surface
    def compile(self):
        _native.check(_native.lib().symlib_unit_compile(self._h), "symlib_unit_compile")
        return self

    def execute(self, *inputs):
        arrs = [np.ascontiguousarray(x, np.float64).reshape(-1) for x in inputs]
        ptrs = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
        lens = (ctypes.c_size_t * len(arrs))(*[a.size for a in arrs])
        _native.check(_native.lib().symlib_unit_execute(self._h, len(arrs), ptrs, lens), "symlib_unit_execute")
        return self

    def execute_async(self, *inputs, stream=None):
        """execute() ordered on `stream` (symhip_set_arg_async): the arrays must stay unchanged
        until the stream has passed the copies."""
        rt = self.runtime()
        if not rt:
            raise _native.SymlibError("execute_async: compile() first")
        arrs = [np.ascontiguousarray(x, np.float64).reshape(-1) for x in inputs]
        for i, a in enumerate(arrs):
            _native.check(_native.lib().symhip_set_arg_async(rt, i, a.ctypes.data_as(_native._DP), a.size, stream),
                          "symhip_set_arg_async")
        self._keep_async = arrs
        return self

    def execute_device(self, *addresses, stream=None):
        """Inputs already in device memory: input container i is copied from device address
        addresses[i] (its length in doubles) on `stream`."""
        rt = self.runtime()
        if not rt:
            raise _native.SymlibError("execute_device: compile() first")
        for i, a in enumerate(addresses):
            _native.check(_native.lib().symhip_set_arg_device(rt, i, a, self.input_lengths[i], stream),
                          "symhip_set_arg_device")
        return self

    def get_results_device(self, k, address, stream=None):
        """Copy output container k to device address `address` on `stream` (no synchronisation)."""
        n = self.output_shape(k)[2]
        _native.check(_native.lib().symhip_get_result_device(self.runtime(), k, address, n, stream),
                      "symhip_get_result_device")
        return self

    def output_device_ptr(self, k=0):
        """Device address of output container k (valid until close; compile() first)."""
        rt = self.runtime()
        if not rt:
            raise _native.SymlibError("output_device_ptr: compile() first")
        return _native.lib().symhip_output_device_ptr(rt, k)

    def run(self, stream=None):
        _native.check(_native.lib().symlib_unit_run(self._h, stream), "symlib_unit_run")
        return self

    def run_graph(self, stream=None):
        """run() replayed from a HIP graph captured on the first call."""
        _native.check(_native.lib().symlib_unit_run_graph(self._h, stream), "symlib_unit_run_graph")
        return self

    def get_results(self, k=0):
        n = self.output_shape(k)[2]
        out = np.empty(n, np.float64)
        _native.check(_native.lib().symlib_unit_get_results(self._h, k, out.ctypes.data, n), "symlib_unit_get_results")
        return out

    # ---- introspection
    def output_shape(self, k=0):
        r, c, nnz = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        _native.check(_native.lib().symlib_unit_output_shape(self._h, k, ctypes.byref(r), ctypes.byref(c),
                                                             ctypes.byref(nnz)), "symlib_unit_output_shape")
        return r.value, c.value, nnz.value

    def output_pattern(self, k=0):
        r, c, nnz = self.output_shape(k)
        outer = np.empty(c + 1, np.int32)
        inner = np.empty(nnz, np.int32)
        _native.check(_native.lib().symlib_unit_output_pattern(self._h, k, outer.ctypes.data, inner.ctypes.data),
                      "symlib_unit_output_pattern")
        return outer, inner

    def info(self):
        i = _native.UnitInfo()
        _native.check(_native.lib().symlib_unit_get_info(self._h, ctypes.byref(i)), "symlib_unit_get_info")
        d = {f: getattr(i, f) for f, _ in i._fields_}
        d["tiling"] = d["tiling"].decode()
        return d

    def source(self):
        return _native.lib().symlib_unit_source(self._h).decode()

    def runtime(self):
        return _native.lib().symlib_unit_runtime(self._h)

    def num_launches(self):
        rt = self.runtime()
        return _native.lib().symhip_num_launches(rt) if rt else 0

    # ---- fused form: tuning and profiling
    def retune(self, **opt):
        """New output groups / waves per EU / block size / schedule / profiling, same tiles."""
        o = _opts(**opt)
        _native.check(_native.lib().symlib_unit_retune(self._h, ctypes.byref(o)), "symlib_unit_retune")
        return self

    def fused_source(self, **opt):
        o = _opts(**opt)
        src = _native.lib().symlib_unit_fused_source(self._h, ctypes.byref(o))
        if src is None:
            _native.check(1, "symlib_unit_fused_source")
        return src.decode()

    def fused_stats(self):
        st = _native.FusedStats()
        _native.check(_native.lib().symlib_unit_fused_stats(self._h, ctypes.byref(st)), "symlib_unit_fused_stats")
        return {f: getattr(st, f) for f, _ in st._fields_}

    def tile_stats(self):
        n = int(self.fused_stats()["num_tiles"])
        out = np.empty((n, 8), np.float32)
        _native.check(_native.lib().symlib_unit_tile_stats(self._h, out.ctypes.data, out.size), "symlib_unit_tile_stats")
        return out

    def tile_profile(self):
        n = int(self.fused_stats()["num_tiles"])
        out = np.empty((n, 4), np.uint64)
        _native.check(_native.lib().symlib_unit_tile_profile(self._h, out.ctypes.data, out.size),
                      "symlib_unit_tile_profile")
        return out

    def launch_info(self, idx=0):
        v, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _native.check(_native.lib().symhip_launch_info(self.runtime(), idx, ctypes.byref(v), ctypes.byref(b),
                                                       ctypes.byref(c)), "symhip_launch_info")
        return {"vgprs": v.value, "blocks_per_cu": b.value, "cus": c.value}

    def launch_grid(self, idx=0):
        rt = self.runtime()
        return _native.lib().symhip_launch_grid(rt, idx) if rt else 0

    def launch_one(self, idx, stream=None):
        _native.check(_native.lib().symhip_launch_one(self.runtime(), idx, stream), "symhip_launch_one")

    def close(self):
        if self._h:
            _native.lib().symlib_unit_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def device_available():
    return bool(_native.lib().symhip_device_available())
