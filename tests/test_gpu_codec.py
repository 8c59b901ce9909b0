"""GPU parity of the entropy coder (encodeRLHuff/encodeHuff/decodeRLHuff/decodeHuff) against the
oracle (GPU dialect) and the reference's own CPU coder (CPU dialect), through the C ABI."""
import zlib

import numpy as np
import pytest

from oracle import pyoracle as o
from tests.gpu_util import laplace_symbols, log, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def inst14(cc, torch_gpu):
    inst = cc.createInstance(0, 8, 1 << 20, 128, 14)
    yield inst
    inst.close()


def gpu_encode(cc, torch, inst, streams, rl):
    dev = [to_dev(torch, s) for s in streams]
    bs = cc.BitStream()
    fn = cc.encodeRLHuff if rl else cc.encodeHuff
    ok = fn(inst, bs, dev, int(streams[0].size))
    torch.cuda.synchronize()
    return ok, bs


def gpu_decode(cc, torch, inst, words, nbits, ns, n, dtype, rl, bitpos=0):
    bs = cc.BitStream(words, nbits)
    bs.setBitPosition(bitpos)
    outs = [torch.full((max(n, 1),), -7, dtype=torch.int16 if dtype == np.uint16 else torch.int32, device="cuda")
            for _ in range(ns)]
    fn = cc.decodeRLHuff if rl else cc.decodeHuff
    assert fn(inst, bs, outs, n)
    torch.cuda.synchronize()
    return [to_host(t, dtype)[:n] for t in outs], bs.getBitPosition()


def check_roundtrip(cc, torch, inst, streams, rl, log2, cbs):
    dtype = streams[0].dtype
    ok, bs = gpu_encode(cc, torch, inst, streams, rl)
    assert ok
    ow, ob = o.encode(streams, rl=rl, log2=log2, cbs=cbs)
    gw = bs.words()
    assert bs.getBitPosition() == ob, (bs.getBitPosition(), ob)
    assert np.array_equal(gw, ow), f"first differing word {np.nonzero(gw != ow)[0][:5]}"
    dec, pos = gpu_decode(cc, torch, inst, gw, ob, len(streams), int(streams[0].size), dtype, rl)
    assert pos == ob
    for a, b in zip(dec, streams):
        assert np.array_equal(a, b)


CASES = [
    ("laplace3x10000", lambda r: [laplace_symbols(r, 10000) for _ in range(3)]),
    ("dense1x4097", lambda r: [laplace_symbols(r, 4097, scale=30, density=1.0)]),
    ("allzero2x5000", lambda r: [np.zeros(5000, np.uint16) for _ in range(2)]),
    ("allsame", lambda r: [np.full(3001, 7, np.uint16), np.full(3001, 2, np.uint16)]),
    ("len1", lambda r: [np.array([5], np.uint16), np.array([0], np.uint16)]),
    ("longruns", lambda r: [np.where(r.random(70000) < 0.0007, 3 + (r.integers(0, 5, 70000)), 0).astype(np.uint16)]),
    ("mixed_empty", lambda r: [np.zeros(9000, np.uint16), laplace_symbols(r, 9000), np.zeros(9000, np.uint16)]),
    ("odd12345", lambda r: [laplace_symbols(r, 12345, scale=100, density=0.6) for _ in range(2)]),
]


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("rl", [True, False], ids=["rlhuff", "huff"])
def test_codec_matches_oracle(cc, torch_gpu, inst14, name, gen, rl):
    log("case", name, "rl", rl)
    rng = np.random.default_rng(zlib.crc32(name.encode()) & 0xFFFF)
    check_roundtrip(cc, torch_gpu, inst14, gen(rng), rl, 14, 128)


@pytest.mark.parametrize("cbs", [32, 64, 256])
def test_coding_block_sizes(cc, torch_gpu, cbs):
    log("cbs", cbs)
    inst = cc.createInstance(0, 4, 1 << 16, cbs, 14)
    rng = np.random.default_rng(cbs)
    for rl in (True, False):
        check_roundtrip(cc, torch_gpu, inst, [laplace_symbols(rng, 20000, scale=8) for _ in range(2)], rl, 14, cbs)
    inst.close()


def test_large_streams(cc, torch_gpu):
    log("large 4 x 1M")
    inst = cc.createInstance(0, 4, 1 << 20, 128, 14)
    rng = np.random.default_rng(11)
    streams = [laplace_symbols(rng, 1 << 20, scale=s, density=d) for s, d in ((2, .1), (20, .9), (300, .5), (1, .01))]
    for rl in (True, False):
        check_roundtrip(cc, torch_gpu, inst, streams, rl, 14, 128)
    inst.close()


def test_symbol32(cc, torch_gpu):
    log("symbol32 log2=20")
    inst = cc.createInstance(0, 4, 1 << 16, 128, 20)
    assert inst.useLongSymbols
    rng = np.random.default_rng(5)
    streams = [laplace_symbols(rng, 30000, scale=3000, density=0.5, maxsym=(1 << 20) - 1, dtype=np.uint32)
               for _ in range(3)]
    for rl in (True, False):
        check_roundtrip(cc, torch_gpu, inst, streams, rl, 20, 128)
    inst.close()


def test_log2_16_pad_quirk(cc, torch_gpu):
    """Streams containing symbol 65535 with an odd count: the reference counts one 0xFFFF pad in bin
    65535 (Histogram.cu:90-96); the bitstream must match the oracle, and still decode."""
    log("log2=16 quirk")
    inst = cc.createInstance(0, 4, 1 << 16, 128, 16)
    rng = np.random.default_rng(9)
    s = laplace_symbols(rng, 20001, scale=50, density=1.0)
    s[rng.integers(0, s.size, 5)] = 65535
    s2 = laplace_symbols(rng, 20001, scale=50, density=0.4)
    s2[s2 != 0][:3] = 65535
    s2[7] = 65535
    for rl in (True, False):
        check_roundtrip(cc, torch_gpu, inst, [s, s2], rl, 16, 128)
    inst.close()


def test_symbol_overflow_returns_false(cc, torch_gpu, inst14):
    log("overflow")
    s = np.zeros(1000, np.uint16)
    s[10] = 1 << 14
    ok, _ = gpu_encode(cc, torch_gpu, inst14, [s], True)
    assert ok is False
    ok, _ = gpu_encode(cc, torch_gpu, inst14, [s], False)
    assert ok is False


def test_unaligned_start_and_per_stream_bitstreams(cc, torch_gpu, inst14):
    log("unaligned + per-stream")
    torch = torch_gpu
    rng = np.random.default_rng(3)
    streams = [laplace_symbols(rng, 7000) for _ in range(3)]
    dev = [to_dev(torch, s) for s in streams]
    # single stream with 13 bits already written
    bs = cc.BitStream()
    bs.writeBits(0x1ABC, 13)
    assert cc.encodeHuff(inst14, bs, dev, 7000)
    ow, ob = o.encode(streams, rl=False, log2=14, cbs=128, start_bit=13)
    got = bs.words()
    ow = ow.copy()
    ow[0] |= np.uint32(0x1ABC << 19)
    assert np.array_equal(got, ow) and bs.getBitPosition() == ob
    bs.setBitPosition(13)
    dec, pos = gpu_decode(cc, torch, inst14, got, ob, 3, 7000, np.uint16, False, bitpos=13)
    assert pos == ob and all(np.array_equal(a, b) for a, b in zip(dec, streams))
    # per-stream bitstreams, two of them the same object
    b1, b2 = cc.BitStream(), cc.BitStream()
    assert cc.encodeRLHuff(inst14, [b1, b2, b1], dev, 7000)
    w02, _ = o.encode([streams[0], streams[2]], rl=True, log2=14, cbs=128)
    w1, _ = o.encode([streams[1]], rl=True, log2=14, cbs=128)
    assert np.array_equal(b1.words(), w02) and np.array_equal(b2.words(), w1)
    outs = [torch.zeros(7000, dtype=torch.int16, device="cuda") for _ in range(3)]
    r1, r2 = cc.BitStream(b1.words()), cc.BitStream(b2.words())
    assert cc.decodeRLHuff(inst14, [r1, r2, r1], outs, 7000)
    torch.cuda.synchronize()
    for t, s in zip(outs, streams):
        assert np.array_equal(to_host(t, np.uint16), s)


def test_decodes_oracle_streams_with_garbage_outputs(cc, torch_gpu, inst14):
    """decodeRLHuff writes every output element: pre-filled outputs must not leak through."""
    log("decode oracle stream")
    rng = np.random.default_rng(21)
    streams = [laplace_symbols(rng, 33333, density=0.05) for _ in range(2)]
    ow, ob = o.encode(streams, rl=True, log2=14, cbs=128)
    dec, pos = gpu_decode(cc, torch_gpu, inst14, ow, ob, 2, 33333, np.uint16, True)
    assert pos == ob
    for a, b in zip(dec, streams):
        assert np.array_equal(a, b)


@pytest.mark.skipif(not o.ref_available(), reason="reference CPU coder not built")
def test_cpu_dialect_matches_reference_cpu_coder(cc, torch_gpu):
    """CPU dialect output is byte-identical to the reference's compiled src/CPU coder, and the GPU
    decoder reads the reference coder's bitstreams back to the same symbols."""
    log("cpu dialect vs reference")
    inst = cc.createInstance(0, 4, 1 << 16, 128, 14)
    inst.setDialect(cc.DIALECT_CPU)
    rng = np.random.default_rng(77)
    streams = [laplace_symbols(rng, 25000, scale=6, density=0.5), laplace_symbols(rng, 25000, scale=1.5, density=0.9),
               np.zeros(25000, np.uint16)]
    for rl in (True, False):
        ok, bs = gpu_encode(cc, torch_gpu, inst, streams, rl)
        assert ok
        rw, rb = o.ref_encode(streams, rl=rl, cbs=128)
        assert bs.getBitPosition() == rb
        assert np.array_equal(bs.words(), rw)
        dec, pos = gpu_decode(cc, torch_gpu, inst, rw, rb, 3, 25000, np.uint16, rl)
        assert pos == rb
        for a, b in zip(dec, streams):
            assert np.array_equal(a, b)
    inst.close()
