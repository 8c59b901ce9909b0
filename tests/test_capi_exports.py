"""The C ABI library loads without a GPU and exports every symbol include/*.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for dirpath, _, files in os.walk(inc):
        for f in files:
            if not f.endswith(".h"):
                continue
            text = open(os.path.join(dirpath, f)).read()
            names.update(re.findall(r"CC_API\s+[\w\s\*]+?\b(cc_\w+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "cc_create_instance" in names and "cc_encode_rlhuff" in names and "cc_dwt_int_forward" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from cudacompress_amd import _capi
    lib = _capi.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes signature table covers every declaration
    assert set(declared_symbols()) == set(_capi.SIGNATURES)


def test_status_strings_and_no_gpu_error():
    from cudacompress_amd import _capi
    lib = _capi.load()
    assert lib.cc_status_string(0) == b"ok"
    assert lib.cc_status_string(6) == b"corrupt bitstream"


def test_invalid_arguments_rejected_before_device_use():
    from cudacompress_amd import _capi
    lib = _capi.load()
    h = ctypes.c_void_p()
    # codingBlockSize must be a power of two in [32, 256]; log2 <= 24
    assert lib.cc_create_instance(-1, 1, 16, 100, 0, ctypes.byref(h)) == 3
    assert lib.cc_create_instance(-1, 1, 16, 0, 25, ctypes.byref(h)) == 3
    # DWT size checks run on the host
    assert lib.cc_dwt_int_forward(None, None, None, 7, 8, 1, 0, 0, 0, 0, None) == 3
    assert b"even" in lib.cc_last_error()
    # image path, bit planes, lowpass-only transforms, multi-channel volumes, timing, plans
    assert lib.cc_dwt_int_forward_lowpass_only(None, None, None, 7, 8, 0, 0, None) == 3
    assert lib.cc_dwt_float_forward_lowpass_only_from_byte(None, None, None, 6, 8, 5, 0, 0, None) == 3
    assert lib.cc_float_to_byte_2d(None, 3, 3, None, 8, 8, 0, None) == 3
    assert b"channel" in lib.cc_last_error()
    assert lib.cc_quantize_difference_to_symbols_2d(None, None, 1.0, None, 1, 0, 8, 8, 0, 2, None) == 3
    assert lib.cc_convert_rgb_to_ycocg(None, None, 10, 5, None) == 3
    assert lib.cc_convert_ycocg_to_rgb(None, None, -1, 3, None) == 3
    assert lib.cc_get_lsb(None, None, 100, None) == 3
    assert lib.cc_get_lsb(None, None, 0, None) == 0  # nothing to do
    assert lib.cc_append_lsb(None, None, 100, 0, None) == 3
    assert lib.cc_decompress_volume_float_multichannel(None, None, 0, 8, 8, 8, 1, 0, 2) == 3
    assert lib.cc_set_timing_detail(None, 1) == 3
    assert lib.cc_bricks_plan_launch(None) == 3


def test_dropin_headers_compile():
    """The C++ drop-in headers (include/cudaCompress) compile against the C ABI declarations."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("g++ or the ROCm headers are not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(root, "include"),
                        "-I", "/opt/rocm/include", os.path.join(root, "tests", "cpp", "dropin_headers.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
