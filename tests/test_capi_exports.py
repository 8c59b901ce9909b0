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
