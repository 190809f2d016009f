"""ctypes bindings of the native library (symboliclib_amd/lib/libsymlib.so).

The library is loaded from this package's lib/ folder only; when it is missing the import
fails with an error instead of falling back to anything else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsymlib.so")


class SymlibError(RuntimeError):
    pass


class Options(ctypes.Structure):
    _fields_ = [("threads_per_block", ctypes.c_int), ("decomposition_threshold", ctypes.c_int),
                ("simplify", ctypes.c_int), ("share_templates", ctypes.c_int), ("tiled", ctypes.c_int),
                ("tile_lds_bytes", ctypes.c_int), ("direct_assembly", ctypes.c_int), ("fused", ctypes.c_int),
                ("schedule", ctypes.c_int), ("output_groups", ctypes.c_int), ("waves_per_eu", ctypes.c_int),
                ("profile_tiles", ctypes.c_int), ("split", ctypes.c_int),
                ("register_staging", ctypes.c_int), ("output_sums", ctypes.c_int), ("cpu_threads", ctypes.c_int)]


class UnitInfo(ctypes.Structure):
    _fields_ = [("num_inputs", ctypes.c_uint64), ("num_intermediates", ctypes.c_uint64),
                ("num_index", ctypes.c_uint64), ("num_const", ctypes.c_uint64), ("num_outpos", ctypes.c_uint64),
                ("num_results", ctypes.c_uint64), ("num_kernels", ctypes.c_uint32), ("num_waves", ctypes.c_uint32),
                ("source_bytes", ctypes.c_uint64), ("frontend_seconds", ctypes.c_double),
                ("build_seconds", ctypes.c_double), ("compile_seconds", ctypes.c_double), ("tiled", ctypes.c_uint32),
                ("tile_pool_entries", ctypes.c_uint64), ("tiling", ctypes.c_char * 128)]


class FusedStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in ("num_tiles", "num_outputs", "num_elements", "element_evaluations",
                                                "contributions", "lds_bytes", "element_lds_bytes", "adds", "muls",
                                                "divs", "sqrts", "statements")] + \
               [(f, ctypes.c_int32) for f in ("max_live", "output_groups", "waves_per_eu", "threads_per_block",
                                               "schedule", "profile_tiles")]


class Launch(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("grid", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("shared", ctypes.c_uint32)]


class ProgramDesc(ctypes.Structure):
    _fields_ = [("num_values", ctypes.c_uint64), ("num_args", ctypes.c_uint32),
                ("arg_start", ctypes.POINTER(ctypes.c_uint64)), ("arg_length", ctypes.POINTER(ctypes.c_uint64)),
                ("num_outputs", ctypes.c_uint32), ("out_start", ctypes.POINTER(ctypes.c_uint64)),
                ("out_length", ctypes.POINTER(ctypes.c_uint64)), ("num_index", ctypes.c_uint64),
                ("index_data", ctypes.POINTER(ctypes.c_uint32)), ("num_const", ctypes.c_uint64),
                ("const_data", ctypes.POINTER(ctypes.c_double)), ("num_outpos", ctypes.c_uint64),
                ("outpos_data", ctypes.POINTER(ctypes.c_uint32)), ("num_tiles_data", ctypes.c_uint64),
                ("tiles_data", ctypes.POINTER(ctypes.c_uint32)), ("num_launches", ctypes.c_uint32),
                ("launches", ctypes.POINTER(Launch)), ("num_work_words", ctypes.c_uint64)]


# symbol -> (restype, argtypes); every exported C function of include/symhip.h and include/symlib.h
_VP, _I, _U32, _U64, _SZ, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_double
_IP = ctypes.POINTER(ctypes.c_int)
_DP = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    # symhip.h
    "symhip_device_available": (_I, []),
    "symhip_create": (_I, [ctypes.POINTER(_VP), ctypes.POINTER(ProgramDesc)]),
    "symhip_compile": (_I, [_VP, ctypes.c_char_p, ctypes.c_char_p]),
    "symhip_compile_to_file": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "symhip_cache_compile": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p]),
    "symhip_update_tiles": (_I, [_VP, _U64, _U64, _VP]),
    "symhip_set_launch": (_I, [_VP, _U32, ctypes.POINTER(Launch)]),
    "symhip_set_launches": (_I, [_VP, _U32, ctypes.POINTER(Launch)]),
    "symhip_set_launch_offsets": (_I, [_VP, _U32, _U64, _U64, _U64]),
    "symhip_set_result_gather": (_I, [_VP, _I]),
    "symhip_reference_prelude": (ctypes.c_char_p, []),
    "symhip_set_work_words": (_I, [_VP, _U64]),
    "symhip_read_work": (_I, [_VP, _U64, _U64, _VP]),
    "symhip_launch_info": (_I, [_VP, _U32, _IP, _IP, _IP]),
    "symhip_wall_clock_khz": (_I, [_IP]),
    "symhip_stream_copy": (_I, [_VP, _VP, _SZ, _U32, _VP]),
    "symhip_implicit_euler_system": (_I, [_I, _VP, _VP, _VP, _VP, _VP, _VP, _VP, _D, _VP, _VP, _VP]),
    "symhip_csc_symv": (_I, [_I, _VP, _VP, _VP, _VP, _VP, _VP]),
    "symhip_set_arg": (_I, [_VP, _I, _DP, _SZ]),
    "symhip_set_arg_async": (_I, [_VP, _I, _DP, _SZ, _VP]),
    "symhip_set_arg_device": (_I, [_VP, _I, _VP, _SZ, _VP]),
    "symhip_run": (_I, [_VP, _VP]),
    "symhip_run_graph": (_I, [_VP, _VP]),
    "symhip_run_noalloc": (_I, [_VP, _VP]),
    "symhip_launch_one": (_I, [_VP, _U32, _VP]),
    "symhip_num_launches": (_U32, [_VP]),
    "symhip_launch_grid": (_U32, [_VP, _U32]),
    "symhip_get_result": (_I, [_VP, _I, _DP, _SZ]),
    "symhip_get_result_device": (_I, [_VP, _I, _VP, _SZ, _VP]),
    "symhip_output_device_ptr": (_VP, [_VP, _I]),
    "symhip_values_device_ptr": (_VP, [_VP]),
    "symhip_destroy": (None, [_VP]),
    "symhip_last_error": (ctypes.c_char_p, []),
    # symlib.h
    "symlib_default_options": (None, [ctypes.POINTER(Options)]),
    "symlib_unit_save": (_I, [_VP, ctypes.c_char_p]),
    "symlib_unit_num_inputs": (_I, [_VP]),
    "symlib_unit_input_length": (_I, [_VP, _I, ctypes.POINTER(_U64)]),
    "symlib_unit_load": (_I, [ctypes.POINTER(_VP), ctypes.c_char_p, ctypes.POINTER(Options)]),
    "symlib_unit_ata_plus_a": (_I, [ctypes.POINTER(_VP), _I, _VP, _VP, ctypes.POINTER(Options)]),
    "symlib_unit_cotan": (_I, [ctypes.POINTER(_VP), _I, _I, _VP, ctypes.POINTER(Options)]),
    "symlib_unit_bw_hessian": (_I, [ctypes.POINTER(_VP), _I, _I, _VP, _VP, _D, _D, _D, ctypes.POINTER(Options)]),
    "symlib_unit_bw_hessian_batch": (_I, [ctypes.POINTER(_VP), _I, _VP, _VP, _VP, _VP, _D, _D, _D,
                                          ctypes.POINTER(Options)]),
    "symlib_unit_compile": (_I, [_VP]),
    "symlib_unit_execute": (_I, [_VP, _I, ctypes.POINTER(_VP), ctypes.POINTER(_SZ)]),
    "symlib_unit_run": (_I, [_VP, _VP]),
    "symlib_unit_run_graph": (_I, [_VP, _VP]),
    "symlib_unit_get_results": (_I, [_VP, _I, _VP, _SZ]),
    "symlib_unit_num_outputs": (_I, [_VP]),
    "symlib_unit_output_shape": (_I, [_VP, _I, _IP, _IP, ctypes.POINTER(_U64)]),
    "symlib_unit_output_pattern": (_I, [_VP, _I, _VP, _VP]),
    "symlib_unit_get_info": (_I, [_VP, ctypes.POINTER(UnitInfo)]),
    "symlib_unit_retune": (_I, [_VP, ctypes.POINTER(Options)]),
    "symlib_unit_fused_source": (ctypes.c_char_p, [_VP, ctypes.POINTER(Options)]),
    "symlib_unit_fused_stats": (_I, [_VP, ctypes.POINTER(FusedStats)]),
    "symlib_unit_tile_stats": (_I, [_VP, _VP, _SZ]),
    "symlib_unit_tile_profile": (_I, [_VP, _VP, _SZ]),
    "symlib_unit_source": (ctypes.c_char_p, [_VP]),
    "symlib_unit_runtime": (_VP, [_VP]),
    "symlib_unit_destroy": (None, [_VP]),
    "symlib_tutorial_known_answers": (_I, [_DP]),
    # symcost.h
    "symcost_create": (_I, [ctypes.POINTER(_VP), _I, _I, _U64]),
    "symcost_train": (_I, [_VP, _VP, _VP, _SZ, _I, _I, ctypes.c_float, _U64, _VP]),
    "symcost_predict": (_I, [_VP, _VP, _SZ, _VP]),
    "symcost_num_params": (_I, [_VP, ctypes.POINTER(_SZ)]),
    "symcost_get_params": (_I, [_VP, _VP, _SZ]),
    "symcost_set_params": (_I, [_VP, _VP, _SZ]),
    "symcost_get_scaling": (_I, [_VP, _VP, _VP, _VP]),
    "symcost_set_scaling": (_I, [_VP, _VP, _VP, _VP]),
    "symcost_step": (_I, [_VP, _VP, _VP, _SZ, ctypes.c_float]),
    "symcost_get_gradient": (_I, [_VP, _VP, _SZ]),
    "symcost_save": (_I, [_VP, ctypes.c_char_p]),
    "symcost_load": (_I, [ctypes.POINTER(_VP), ctypes.c_char_p]),
    "symcost_destroy": (None, [_VP]),
    "symcost_last_error": (ctypes.c_char_p, []),
    "symlib_last_error": (ctypes.c_char_p, []),
    # symcpu.h
    "symcpu_create": (_I, [ctypes.POINTER(_VP), ctypes.POINTER(ProgramDesc)]),
    "symcpu_compile": (_I, [_VP, ctypes.c_char_p, ctypes.c_char_p]),
    "symcpu_set_arg": (_I, [_VP, _I, _DP, _SZ]),
    "symcpu_run": (_I, [_VP, _I]),
    "symcpu_get_result": (_I, [_VP, _I, _DP, _SZ]),
    "symcpu_destroy": (None, [_VP]),
    "symcpu_last_error": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """The loaded native library; raises SymlibError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SymlibError(f"native library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        L = lib()
        msg = ((L.symlib_last_error() or b"").decode() or (L.symhip_last_error() or b"").decode()
               or (L.symcost_last_error() or b"").decode() or (L.symcpu_last_error() or b"").decode())
        raise SymlibError(f"{what} failed ({rc}): {msg}")
