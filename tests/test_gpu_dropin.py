"""The C++ drop-in API run on the GPU: tests/cpp/dropin_volume.cpp (built by tests/cpp/build.sh)
calls cudaCompress::createInstance, util::dwtFloat3DForward, util::quantizeToSymbols,
encodeHuff / encodeRLHuff into a BitStream over a std::vector, and the decode path with
decodeHuff / decodeRLHuff and util::dwtFloat3DInverseFromSymbols, exactly as the reference's
compressVolumeFloat / decompressVolumeFloat (src/examples/CompressVolume.cpp:313-516) do, plus an
encodeHuff after 3 caller bits (unaligned start) and an aliased per-stream BitStream* array.
Its outputs are compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

from cudacompress_amd import synthetic as S
from oracle import pipeline as P
from oracle import pyoracle as o
from tests.gpu_util import laplace_symbols, log

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_bin", "dropin_volume")


@pytest.mark.parametrize("shape,levels", [((64, 48, 32), 2), ((32, 32, 32), 3)], ids=["64x48x32L2", "32^3L3"])
def test_cpp_dropin_volume(torch_gpu, tmp_path, shape, levels):
    assert os.path.exists(EXE), "tests/cpp/build.sh has not been run (python __graft_entry__.py build)"
    z, y, x = shape
    vol = S.gaussian_field(shape, 11)
    q0 = 0.02
    n = vol.size // 8
    rng = np.random.default_rng(31)
    streams = np.stack([laplace_symbols(rng, n, scale=s, density=d) for s, d in ((2.0, 0.4), (9.0, 0.2), (1.0, 0.8))])
    vol.tofile(tmp_path / "volume.f32")
    streams.astype(np.uint16).tofile(tmp_path / "streams.u16")
    log("C++ drop-in", shape, levels)
    r = subprocess.run([EXE, str(tmp_path), str(x), str(y), str(z), str(levels), repr(q0)], capture_output=True,
                       text=True, timeout=120)
    log(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "status.txt").read_text().strip() == "ok"
    # compressVolumeFloat through the C++ wrappers == the oracle
    words = np.fromfile(tmp_path / "volume.words", dtype=np.uint32)
    want = P.compress_volume_float(vol, levels, q0)
    assert words.size == want.size and np.array_equal(words, want)
    rec = np.fromfile(tmp_path / "volume.rec", dtype=np.float32).reshape(shape)
    ref = P.decompress_volume_float(want, shape, levels, q0)
    assert float(np.abs(rec - ref).max()) <= 1e-6 * max(1.0, float(np.abs(ref).max()))
    # unaligned start: the caller's 3 bits (101), then the Huffman segment
    pw = np.fromfile(tmp_path / "prefix.words", dtype=np.uint32)
    ow, ob = o.encode([streams[0]], rl=False, log2=14, cbs=128, start_bit=3)
    ow = ow.copy()
    ow[0] = (ow[0] & np.uint32(0x1FFFFFFF)) | np.uint32(5 << 29)
    assert np.array_equal(pw, ow[: pw.size]) and pw.size == (ob + 31) // 32
    # aliased per-stream bitstreams {b1, b2, b1}: b1 holds streams 0 and 2, b2 stream 1
    a1 = np.fromfile(tmp_path / "alias1.words", dtype=np.uint32)
    a2 = np.fromfile(tmp_path / "alias2.words", dtype=np.uint32)
    w02, _ = o.encode([streams[0], streams[2]], rl=True, log2=14, cbs=128)
    w1, _ = o.encode([streams[1]], rl=True, log2=14, cbs=128)
    assert np.array_equal(a1, w02) and np.array_equal(a2, w1)
    dec = np.fromfile(tmp_path / "alias.dec", dtype=np.uint16).reshape(3, n)
    assert np.array_equal(dec, streams)
