"""The drop-in include/cudaCompress/BitStream.h, compiled and run (CPU).

oracle/bitstream_trace.cpp runs a seeded sequence of writes and reads (every bit count 0..32 at
every offset, aligned writes of 1/2/4/8-byte elements, writeAlign past the end, caller-owned
vectors, setBitSize, reads and skips past the end, align() clamped at the end of the stream,
readAligned/skipAligned/getAlignedPointer) and prints every result.  Built against this
repository's header, its trace must equal the one of the reference's own header
(include/cudaCompress/BitStream.h:14-497): the committed fixture written by
tests/golden/make_golden.py and, where the reference tree is present, a live build of it
(oracle/_ref/bitstream_trace)."""
import gzip
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OURS = os.path.join(ROOT, "tests", "cpp", "_bin", "bitstream_trace")
REF = os.path.join(ROOT, "oracle", "_ref", "bitstream_trace")
FIXTURE = os.path.join(ROOT, "tests", "golden", "bitstream_trace_ref.txt.gz")


@pytest.fixture(scope="module")
def ours():
    src = os.path.join(ROOT, "oracle", "bitstream_trace.cpp")
    if not os.path.exists(OURS) or os.path.getmtime(OURS) < max(
            os.path.getmtime(src), os.path.getmtime(os.path.join(ROOT, "include", "cudaCompress", "BitStream.h"))):
        os.makedirs(os.path.dirname(OURS), exist_ok=True)
        subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"), src, "-o", OURS],
                       check=True)
    return subprocess.run([OURS], check=True, capture_output=True, timeout=60).stdout


def test_trace_equals_reference_fixture(ours):
    with gzip.open(FIXTURE, "rb") as f:
        want = f.read()
    a, b = ours.splitlines(), want.splitlines()
    assert len(a) == len(b)
    for i, (x, y) in enumerate(zip(a, b)):
        assert x == y, f"line {i}: {x!r} != {y!r}"


@pytest.mark.skipif(not os.path.exists(REF), reason="reference header build not present")
def test_trace_equals_live_reference_build(ours):
    assert ours == subprocess.run([REF], check=True, capture_output=True, timeout=60).stdout


def test_trace_covers_the_edge_cases(ours):
    t = ours.decode()
    assert "beyond end" in t                      # align() stepping past the end
    assert " -> 0 " in t                          # reads returning 0 bits at the end
    assert "setBitSize" in t and "writeAligned k3" in t
