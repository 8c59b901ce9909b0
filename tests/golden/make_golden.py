"""Writes tests/golden/cpu_coder.npz: input/output vectors of the reference's own CPU entropy coder.

The reference (m0bl0/cudaCompress) ships no golden vectors or known-answer tests, so these are
generated here from its src/CPU coder, compiled from the reference sources in place into
oracle/_ref/libref_cpu.so by oracle/Makefile:

    make -C oracle && python tests/golden/make_golden.py

Recorded per case (symbol streams are seeded; values < 16384 so that the CPU dialect's 14-bit
header fields hold them):
  - encodeRLHuffCPU and encodeHuffCPU bitstreams (src/CPU/EncodeCPU.cpp:98-197, :275-320)
    for coding block sizes 128 and 64;
  - runLengthZeroesEncodeCPU pairs of stream 0 (src/CPU/RunLengthCPU.cpp:8-32);
  - HuffmanTableCPU::design of stream 0 (src/CPU/HuffmanTableCPU.cpp:22-105).
Only data (inputs and outputs) is written; no reference source is stored.

It also writes tests/golden/bitstream_trace_ref.txt.gz: the output of oracle/bitstream_trace.cpp
built against the reference's own include/cudaCompress/BitStream.h (oracle/_ref/bitstream_trace),
a seeded sequence of bit stream writes and reads with every returned value, position and word.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pyoracle as O  # noqa: E402

OUT = os.path.join(HERE, "cpu_coder.npz")
TRACE_OUT = os.path.join(HERE, "bitstream_trace_ref.txt.gz")


def laplace_symbols(rng, n, scale, zero_frac):
    """Zig-zag symbols of a rounded Laplacian with extra zeros (a wavelet highpass band)."""
    v = np.rint(rng.laplace(0.0, scale, n)).astype(np.int64)
    v[rng.random(n) < zero_frac] = 0
    s = np.where(v < 0, -2 * v - 1, 2 * v)
    return np.minimum(s, 16383).astype(np.uint16)


def cases():
    rng = np.random.default_rng(20240611)
    out = []
    # (name, streams)
    out.append(("laplace_1x1000", [laplace_symbols(rng, 1000, 3.0, 0.5)]))
    out.append(("laplace_3x4096", [laplace_symbols(rng, 4096, s, z) for s, z in ((1.0, 0.7), (8.0, 0.3), (40.0, 0.1))]))
    out.append(("ragged_2x777", [laplace_symbols(rng, 777, 2.0, 0.6) for _ in range(2)]))
    # many equal counts: ties inside the heap and inside the depth sort
    t = np.repeat(np.arange(64, dtype=np.uint16), 16)
    rng.shuffle(t)
    out.append(("ties_1x1024", [t]))
    # long zero runs (fillers (0,255) for runs >= 256) and trailing zeros
    z = np.zeros(5000, dtype=np.uint16)
    z[[0, 300, 301, 2000, 4100]] = [5, 1, 2, 7, 3]
    out.append(("zero_runs_1x5000", [z]))
    out.append(("all_zero_2x300", [np.zeros(300, np.uint16), np.zeros(300, np.uint16)]))
    out.append(("single_symbol_1x1", [np.array([9], np.uint16)]))
    out.append(("wide_alphabet_1x6000", [rng.integers(0, 16384, 6000).astype(np.uint16)]))
    return out


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libref_cpu.so not built (make -C oracle with the reference tree present)")
    arrays = {}
    meta = []
    for i, (name, streams) in enumerate(cases()):
        key = f"c{i}"
        arrays[f"{key}_in"] = np.stack(streams)
        entry = {"key": key, "name": name, "nstreams": len(streams), "count": int(streams[0].size), "encodings": []}
        for rl in (True, False):
            for cbs in (128, 64):
                w, nb = O.ref_encode(streams, rl=rl, cbs=cbs)
                k = f"{key}_{'rl' if rl else 'h'}{cbs}"
                arrays[k] = w
                entry["encodings"].append({"array": k, "rl": rl, "cbs": cbs, "bits": int(nb)})
        s, zc = O.ref_rle(streams[0])
        arrays[f"{key}_rle_sym"], arrays[f"{key}_rle_zc"] = s, zc
        d = O.ref_design(streams[0])
        arrays[f"{key}_design_cpl"], arrays[f"{key}_design_syms"] = d["count_per_length"], d["symbols"]
        meta.append(entry)
    arrays["meta_json"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT}: {len(meta)} cases, {os.path.getsize(OUT)} bytes")
    write_trace()


def write_trace():
    import gzip
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "bitstream_trace")
    text = subprocess.run([exe], check=True, capture_output=True).stdout
    with gzip.GzipFile(TRACE_OUT, "wb", mtime=0) as f:
        f.write(text)
    lines = len(text.splitlines())
    print(f"wrote {TRACE_OUT}: {lines} lines")


if __name__ == "__main__":
    main()
