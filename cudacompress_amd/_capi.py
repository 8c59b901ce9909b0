"""ctypes binding of the C ABI in include/cudacompress_amd.h.

Loads cudacompress_amd/lib/libcudacompress_amd.so built in-tree; there is no fallback: if the
library is missing the import fails with an error naming the build command.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcudacompress_amd.so")

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_float = ctypes.c_float
c_vp = ctypes.c_void_p
P_u32 = ctypes.POINTER(c_u32)
P_u64 = ctypes.POINTER(c_u64)
PP_u32 = ctypes.POINTER(P_u32)
PP_u64 = ctypes.POINTER(P_u64)

# name -> (restype, argtypes)
SIGNATURES = {
    "cc_status_string": (ctypes.c_char_p, [c_int]),
    "cc_last_error": (ctypes.c_char_p, []),
    "cc_create_instance": (c_int, [c_int, c_u32, c_u32, c_u32, c_u32, ctypes.POINTER(c_vp)]),
    "cc_destroy_instance": (None, [c_vp]),
    "cc_get_instance_device": (c_int, [c_vp]),
    "cc_get_instance_stream_count_max": (c_u32, [c_vp]),
    "cc_get_instance_elem_count_per_stream_max": (c_u32, [c_vp]),
    "cc_get_instance_coding_block_size": (c_u32, [c_vp]),
    "cc_get_instance_log2_huffman_distinct_symbol_count_max": (c_u32, [c_vp]),
    "cc_get_instance_use_long_symbols": (c_int, [c_vp]),
    "cc_set_instance_stream": (None, [c_vp, c_vp]),
    "cc_get_instance_stream": (c_vp, [c_vp]),
    "cc_set_instance_dialect": (c_int, [c_vp, c_int]),
    "cc_set_instance_brick_groups": (c_int, [c_vp, c_u32]),
    "cc_encode_rlhuff": (c_int, [c_vp, c_int, c_vp, c_u32, c_u32, c_vp, PP_u32, PP_u64, PP_u64]),
    "cc_encode_huff": (c_int, [c_vp, c_int, c_vp, c_u32, c_u32, c_vp, PP_u32, PP_u64, PP_u64]),
    "cc_encode_device": (c_int, [c_vp, c_int, c_int, c_vp, c_u32, c_u32, c_vp, c_u64, P_u64]),
    "cc_decode_device": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_vp, c_u32, c_u32, P_u64]),
    "cc_decode_device_noalloc": (c_int, [c_vp, c_int, c_int, c_vp, c_u64, c_vp, c_u32, c_u32, P_u64]),
    "cc_decode_rlhuff": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, c_u32, c_vp, c_vp, c_u32, c_u32]),
    "cc_decode_huff": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp, c_u32, c_vp, c_vp, c_u32, c_u32]),
    "cc_dwt_float_forward": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_dwt_double_forward": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_dwt_float_inverse": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_dwt_float_inverse_from_symbols": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_float, c_int, c_int, c_int,
                                                   c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_dwt_int_forward": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_dwt_int_inverse": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_symbolize": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u32, c_vp]),
    "cc_unsymbolize": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u32, c_vp]),
    "cc_quantize_to_symbols": (c_int, [c_vp, c_int, c_vp, c_u32, c_u32, c_u32, c_float, c_u32, c_u32, c_int, c_vp]),
    "cc_quantize_to_symbols_roundtrip": (c_int, [c_vp, c_int, c_vp, c_u32, c_u32, c_u32, c_float, c_u32, c_u32, c_int, c_vp]),
    "cc_unquantize_from_symbols": (c_int, [c_vp, c_vp, c_int, c_u32, c_u32, c_u32, c_float, c_u32, c_u32, c_int, c_vp]),
    "cc_quantize_to_short": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_float, c_u32, c_u32, c_int, c_vp]),
    "cc_unquantize_from_short": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_float, c_u32, c_u32, c_int, c_vp]),
    "cc_get_max_abs": (c_int, [c_vp, c_vp, c_u32, c_vp]),
    "cc_compress_volume_float": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_float, c_int, c_int, PP_u32, P_u64]),
    "cc_decompress_volume_float": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_vp, c_u64, c_float, c_int, c_int]),
    "cc_compress_volume_float_quant_first": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_float, c_int, PP_u32,
                                                     P_u64]),
    "cc_decompress_volume_float_quant_first": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_vp, c_u64, c_float,
                                                       c_int]),
    "cc_decompress_volume_float_multichannel": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u32, c_int,
                                                        c_int]),
    "cc_decompress_volume_float_quant_first_multichannel": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u32,
                                                                    c_int]),
    "cc_dwt_float_forward_from_byte": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int,
                                               c_int, c_int, c_vp]),
    "cc_dwt_float_inverse_to_byte": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                             c_int, c_vp]),
    "cc_dwt_float_inverse_from_symbols_to_byte": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_float, c_int, c_int, c_int,
                                                          c_int, c_int, c_int, c_int, c_int, c_vp]),
    "cc_dwt_float_forward_lowpass_only_from_byte": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_int,
                                                            c_vp]),
    "cc_dwt_int_forward_lowpass_only": (c_int, [c_vp, c_vp, c_vp, c_int, c_int, c_int, c_int, c_vp]),
    "cc_float_to_byte_2d": (c_int, [c_vp, c_u32, c_u32, c_vp, c_u32, c_u32, c_u32, c_vp]),
    "cc_byte_to_float_2d": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_u32, c_vp]),
    "cc_quantize_difference_to_symbols_2d": (c_int, [c_vp, c_vp, c_float, c_vp, c_u32, c_u32, c_u32, c_u32, c_u32,
                                                     c_int, c_vp]),
    "cc_quantize_difference_to_symbols_roundtrip_2d": (c_int, [c_vp, c_vp, c_float, c_vp, c_u32, c_u32, c_u32, c_u32,
                                                               c_u32, c_int, c_vp]),
    "cc_unquantize_difference_from_symbols_2d": (c_int, [c_vp, c_vp, c_float, c_vp, c_u32, c_u32, c_u32, c_u32,
                                                         c_u32, c_int, c_vp]),
    "cc_convert_rgb_to_ycocg": (c_int, [c_vp, c_vp, c_int, c_int, c_vp]),
    "cc_convert_ycocg_to_rgb": (c_int, [c_vp, c_vp, c_int, c_int, c_vp]),
    "cc_get_lsb": (c_int, [c_vp, c_vp, c_u32, c_vp]),
    "cc_remove_lsb": (c_int, [c_vp, c_vp, c_u32, ctypes.c_int16, c_vp]),
    "cc_append_lsb": (c_int, [c_vp, c_vp, c_u32, ctypes.c_int16, c_vp]),
    "cc_shift_elems": (c_int, [c_vp, c_u32, ctypes.c_int16, c_vp]),
    "cc_set_timing_detail": (c_int, [c_vp, c_int]),
    "cc_get_timings": (c_int, [c_vp, P_u32]),
    "cc_get_timing_entry": (c_int, [c_vp, c_u32, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(c_float)]),
    "cc_print_timings": (c_int, [c_vp]),
    "cc_reset_timings": (c_int, [c_vp]),
    "cc_compress_volume_lossless": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, PP_u32, P_u64]),
    "cc_decompress_volume_lossless": (c_int, [c_vp, c_vp, c_u32, c_u32, c_u32, c_u32, c_vp, c_u64]),
    "cc_compress_bricks_float": (c_int, [c_vp, c_vp, c_u32, c_u32, c_int, c_vp, c_u64, P_u64, P_u32]),
    "cc_decompress_bricks_float": (c_int, [c_vp, c_vp, c_u32, c_u32, c_int, c_vp, P_u64, P_u32]),
    "cc_bricks_plan_create": (c_int, [c_vp, c_vp, c_u32, c_u32, c_int, c_vp, c_u64, ctypes.POINTER(c_vp)]),
    "cc_bricks_plan_compress": (c_int, [c_vp, P_u64, P_u32]),
    "cc_bricks_plan_decompress": (c_int, [c_vp, P_u32]),
    "cc_bricks_plan_run": (c_int, [c_vp, P_u64, P_u32, P_u32]),
    "cc_bricks_plan_launch": (c_int, [c_vp]),
    "cc_bricks_plan_wait": (c_int, [c_vp, P_u64, P_u32, P_u32]),
    "cc_bricks_plan_destroy": (None, [c_vp]),
    "cc_bricks_plan_coding_stats": (c_int, [c_vp, P_u64]),
    "cc_default_tuning": (c_int, [c_vp]),
    "cc_load_tuning": (c_int, [ctypes.c_char_p, c_vp]),
    "cc_set_instance_tuning": (c_int, [c_vp, c_vp]),
    "cc_get_instance_tuning": (c_int, [c_vp, c_vp]),
    "cc_server_create": (c_int, [c_vp, c_vp, c_u32, P_u64, P_u32, ctypes.POINTER(c_float), c_u32, c_int, c_u32, c_u64,
                                 ctypes.POINTER(c_vp)]),
    "cc_server_submit": (c_int, [c_vp, c_u64, P_u32, c_u32, ctypes.POINTER(c_vp), ctypes.c_double]),
    "cc_server_poll": (c_int, [c_vp, ctypes.c_double, P_u32]),
    "cc_server_pending": (c_u32, [c_vp]),
    "cc_server_completed": (c_u32, [c_vp, P_u64, ctypes.POINTER(ctypes.c_double), c_u32]),
    "cc_server_failed": (c_u32, [c_vp, P_u64, P_u32, c_u32]),
    "cc_server_run": (c_int, [c_vp, c_u32, ctypes.POINTER(ctypes.c_double), P_u32, P_u32, ctypes.POINTER(c_vp),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "cc_server_destroy": (None, [c_vp]),
    "cc_profile_enable": (None, [c_int]),
    "cc_profile_reset": (None, []),
    "cc_profile_collect": (c_u32, []),
    "cc_profile_entry": (c_int, [c_u32, ctypes.POINTER(ctypes.c_char_p), P_u64, ctypes.POINTER(ctypes.c_double), P_u64]),
}


class Tuning(ctypes.Structure):
    _fields_ = [("designGrid", c_u32 * 4), ("fusedXY", c_u32), ("xySlicesMax", c_u32), ("codingBlockSize", c_u32),
                ("laneDesignK", c_u32)]


class Brick(ctypes.Structure):
    """struct cc_brick of include/cudacompress_amd.h"""
    _fields_ = [("src", c_vp), ("dst", c_vp), ("sizeX", c_u32), ("sizeY", c_u32), ("sizeZ", c_u32),
                ("quantStep", c_float)]


class VolumeChannel(ctypes.Structure):
    """struct cc_volume_channel of include/cudacompress_amd.h"""
    _fields_ = [("d_image", c_vp), ("bits", c_vp), ("bitCount", c_u64), ("quantStep", c_float)]


def profile_entries():
    """Per-kernel (name, launches, total_ms, algorithmic_bytes) since the last reset."""
    lib = load()
    n = lib.cc_profile_collect()
    out = []
    for i in range(n):
        name = ctypes.c_char_p()
        launches = c_u64()
        ms = ctypes.c_double()
        nbytes = c_u64()
        if lib.cc_profile_entry(i, ctypes.byref(name), ctypes.byref(launches), ctypes.byref(ms),
                                ctypes.byref(nbytes)) == 0:
            out.append((name.value.decode(), launches.value, ms.value, nbytes.value))
    return out

_lib = None


def _share_hip_runtime_with_torch():
    """One HIP runtime per process.  PyTorch-ROCm wheels bundle their own libamdhip64.so (same
    SONAME, libamdhip64.so.7, as /opt/rocm's).  Loading that file first, by path, makes this
    library bind to it through the SONAME and lets a later `import torch` find it by inode, so
    torch tensors and this library share one runtime whichever is imported first.  Two runtimes
    in one process do not see the GPU.  No-op when torch is not installed."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if os.path.exists(path):
        ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


def load():
    """Loads the library (once).  Raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"cudacompress_amd native library not found at {LIB_PATH}; build it with "
            "`python cudacompress_amd/build.py` (hipcc, gfx950)")
    _share_hip_runtime_with_torch()
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
