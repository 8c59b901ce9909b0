"""GPU parity of the device-resident coder (cc_encode_device: RLE, histograms, table design,
packing and layout on the GPU) against the oracle's bitstreams, through the C ABI."""
import zlib

import numpy as np
import pytest

from oracle import pyoracle as o
from tests.gpu_util import laplace_symbols, log, to_dev

pytestmark = pytest.mark.gpu


def device_words(cc, torch, inst, streams, rl, cap_words=None):
    dev = [to_dev(torch, s) for s in streams]
    n = int(streams[0].size) if streams else 0
    cap = cap_words or (sum(s.size for s in streams) * 2 + 4096)
    out = torch.full((cap,), -1, dtype=torch.int32, device="cuda")
    nw = cc.encodeDevice(inst, rl, dev, n, out)
    torch.cuda.synchronize()
    if nw is None:
        return None
    return out[:nw].cpu().numpy().view(np.uint32)


def check(cc, torch, inst, streams, rl, log2, cbs, dialect=o.DIALECT_GPU):
    got = device_words(cc, torch, inst, streams, rl)
    want, bits = o.encode(streams, rl=rl, dialect=dialect, log2=log2, cbs=cbs)
    assert got is not None
    assert got.size == want.size, (got.size, want.size)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"{bad.size} words differ, first at {bad[:5]}"


def _ties(r, n):
    t = np.repeat(np.arange(1, 300, dtype=np.uint16), n // 299 + 1)[:n]
    r.shuffle(t)
    return t


CASES = [
    ("laplace3x10000", lambda r: [laplace_symbols(r, 10000) for _ in range(3)]),
    ("dense1x4097", lambda r: [laplace_symbols(r, 4097, scale=30, density=1.0)]),
    ("allzero2x5000", lambda r: [np.zeros(5000, np.uint16) for _ in range(2)]),
    ("allsame", lambda r: [np.full(3001, 7, np.uint16), np.full(3001, 2, np.uint16)]),
    ("len1", lambda r: [np.array([5], np.uint16), np.array([0], np.uint16)]),
    ("longruns", lambda r: [np.where(r.random(70000) < 0.0007, 3 + (r.integers(0, 5, 70000)), 0).astype(np.uint16)]),
    ("run256", lambda r: [np.concatenate([np.zeros(256, np.uint16), [4], np.zeros(512, np.uint16), [9],
                                          np.zeros(300, np.uint16)]).astype(np.uint16)]),
    ("mixed_empty", lambda r: [np.zeros(9000, np.uint16), laplace_symbols(r, 9000), np.zeros(9000, np.uint16)]),
    ("odd12345", lambda r: [laplace_symbols(r, 12345, scale=100, density=0.6) for _ in range(2)]),
    ("ties", lambda r: [_ties(r, 20000), _ties(r, 20000)]),
    ("wide5000", lambda r: [laplace_symbols(r, 30000, scale=3000, density=0.9, maxsym=16383)]),
    ("bigK", lambda r: [r.integers(0, 16384, 60000).astype(np.uint16)]),
    ("big1x300000", lambda r: [laplace_symbols(r, 300000, scale=6, density=0.5)]),
]


@pytest.fixture(scope="module")
def inst14(cc, torch_gpu):
    inst = cc.createInstance(0, 8, 1 << 20, 128, 14)
    yield inst
    inst.close()


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("rl", [True, False], ids=["rlhuff", "huff"])
def test_device_coder_matches_oracle(cc, torch_gpu, inst14, name, gen, rl):
    log("device coder case", name, "rl", rl)
    rng = np.random.default_rng(zlib.crc32(name.encode()) & 0xFFFF)
    check(cc, torch_gpu, inst14, gen(rng), rl, 14, 128)


@pytest.fixture(scope="module", params=[32, 256], ids=["lanes32", "lanes256"])
def inst_lanes(cc, torch_gpu, request):
    inst = cc.createInstance(0, 8, 1 << 20, 128, 14)
    inst.setTuning(dict(cc.defaultTuning(), laneDesignK=request.param))
    yield inst
    inst.close()


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("rl", [True, False], ids=["rlhuff", "huff"])
def test_device_coder_tables_per_thread_match_oracle(cc, torch_gpu, inst_lanes, name, gen, rl):
    """Tables designed one per thread (laneDesignK): the same words as the oracle."""
    log("device coder, per-thread tables, case", name, "rl", rl)
    rng = np.random.default_rng(zlib.crc32(name.encode()) & 0xFFFF)
    check(cc, torch_gpu, inst_lanes, gen(rng), rl, 14, 128)


@pytest.mark.parametrize("cbs", [32, 256])
def test_device_coder_block_sizes(cc, torch_gpu, cbs):
    log("device coder cbs", cbs)
    inst = cc.createInstance(0, 4, 1 << 16, cbs, 14)
    rng = np.random.default_rng(cbs)
    for rl in (True, False):
        check(cc, torch_gpu, inst, [laplace_symbols(rng, 20000, scale=8) for _ in range(2)], rl, 14, cbs)
    inst.close()


def test_device_coder_log2_16_quirk(cc, torch_gpu):
    """Symbol16 at log2 16: 65535 present with an odd count (histogram pad quirk)."""
    log("device coder log2 16 quirk")
    inst = cc.createInstance(0, 4, 1 << 16, 128, 16)
    rng = np.random.default_rng(16)
    for n in (20001, 20000):
        s = laplace_symbols(rng, n, scale=2000, density=0.8)
        s[5] = 65535
        for rl in (True, False):
            check(cc, torch_gpu, inst, [s], rl, 16, 128)
    inst.close()


def test_device_coder_symbol32(cc, torch_gpu):
    log("device coder symbol32")
    inst = cc.createInstance(0, 4, 1 << 16, 128, 20)
    rng = np.random.default_rng(20)
    s = [laplace_symbols(rng, 15000, scale=50000, density=0.7, dtype=np.uint32, maxsym=(1 << 20) - 1) for _ in range(2)]
    for rl in (True, False):
        check(cc, torch_gpu, inst, s, rl, 20, 128)
    inst.close()


def test_device_coder_cpu_dialect(cc, torch_gpu):
    """CPU dialect on the device equals the oracle, which equals the reference's own CPU coder."""
    log("device coder cpu dialect")
    inst = cc.createInstance(0, 4, 1 << 16, 128, 14)
    inst.setDialect(cc.DIALECT_CPU)
    rng = np.random.default_rng(3)
    s = [laplace_symbols(rng, 9000, scale=20, density=0.5) for _ in range(3)]
    for rl in (True, False):
        check(cc, torch_gpu, inst, s, rl, 14, 128, dialect=o.DIALECT_CPU)
    inst.close()


def test_device_coder_overflow(cc, torch_gpu, inst14):
    log("device coder overflow")
    s = np.full(100, 20000, np.uint16)
    assert device_words(cc, torch_gpu, inst14, [s], True) is None


# ---------------------------------------------------------------------------- device decoder
def device_decode(cc, torch, inst, words, ns, n, dtype, rl, avail=None):
    dw = torch.from_numpy(np.ascontiguousarray(words, dtype=np.uint32).view(np.int32)).cuda()
    tdt = torch.int16 if dtype == np.uint16 else torch.int32
    outs = [torch.full((max(n, 1),), -7, dtype=tdt, device="cuda") for _ in range(ns)]
    used = cc.decodeDevice(inst, rl, dw, outs, n, wordCount=avail)
    torch.cuda.synchronize()
    return [o.cpu().numpy().view(dtype)[:n] for o in outs], used


@pytest.mark.parametrize("name,gen", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("rl", [True, False], ids=["rlhuff", "huff"])
def test_device_decoder_roundtrip(cc, torch_gpu, inst14, name, gen, rl):
    """The device decoder reads the oracle's bitstream back to the original symbols."""
    log("device decoder case", name, "rl", rl)
    rng = np.random.default_rng(zlib.crc32(name.encode()) & 0xFFFF)
    streams = gen(rng)
    words, bits = o.encode(streams, rl=rl, log2=14, cbs=128)
    dec, used = device_decode(cc, torch_gpu, inst14, words, len(streams), int(streams[0].size), streams[0].dtype, rl)
    assert used == words.size
    for a, b in zip(dec, streams):
        assert np.array_equal(a, b)


def test_device_decoder_variants(cc, torch_gpu):
    """log2 16 with the pad quirk, Symbol32, the CPU dialect (reference-coder bitstreams) and
    coding block 32."""
    log("device decoder variants")
    rng = np.random.default_rng(77)
    inst = cc.createInstance(0, 4, 1 << 16, 128, 16)
    s = laplace_symbols(rng, 20001, scale=2000, density=0.8)
    s[5] = 65535
    for rl in (True, False):
        w, _ = o.encode([s], rl=rl, log2=16, cbs=128)
        dec, used = device_decode(cc, torch_gpu, inst, w, 1, s.size, np.uint16, rl)
        assert used == w.size and np.array_equal(dec[0], s)
    inst.close()
    inst = cc.createInstance(0, 4, 1 << 16, 128, 20)
    s32 = [laplace_symbols(rng, 15000, scale=50000, density=0.7, dtype=np.uint32, maxsym=(1 << 20) - 1)]
    for rl in (True, False):
        w, _ = o.encode(s32, rl=rl, log2=20, cbs=128)
        dec, used = device_decode(cc, torch_gpu, inst, w, 1, s32[0].size, np.uint32, rl)
        assert used == w.size and np.array_equal(dec[0], s32[0])
    inst.close()
    inst = cc.createInstance(0, 4, 1 << 16, 128, 14)
    inst.setDialect(cc.DIALECT_CPU)
    sc = [laplace_symbols(rng, 9000, scale=20, density=0.5) for _ in range(3)]
    for rl in (True, False):
        w, _ = o.encode(sc, rl=rl, dialect=o.DIALECT_CPU, cbs=128)
        dec, used = device_decode(cc, torch_gpu, inst, w, 3, 9000, np.uint16, rl)
        assert used == w.size and all(np.array_equal(a, b) for a, b in zip(dec, sc))
    inst.close()
    inst = cc.createInstance(0, 4, 1 << 16, 32, 14)
    s = [laplace_symbols(rng, 20000, scale=8) for _ in range(2)]
    for rl in (True, False):
        w, _ = o.encode(s, rl=rl, log2=14, cbs=32)
        dec, used = device_decode(cc, torch_gpu, inst, w, 2, 20000, np.uint16, rl)
        assert used == w.size and all(np.array_equal(a, b) for a, b in zip(dec, s))
    inst.close()


def test_device_decoder_truncated_stream_reports_error(cc, torch_gpu, inst14):
    """A bitstream cut short is reported as corrupt (every read is bounded by the word count)."""
    log("device decoder truncated")
    rng = np.random.default_rng(5)
    s = [laplace_symbols(rng, 5000) for _ in range(2)]
    w, _ = o.encode(s, rl=True, log2=14, cbs=128)
    with pytest.raises(cc.CCError):
        device_decode(cc, torch_gpu, inst14, w, 2, 5000, np.uint16, True, avail=w.size // 2)


def test_device_encode_decode_chain(cc, torch_gpu, inst14):
    """Device encode followed by device decode, no host copy of the bitstream in between."""
    log("device encode->decode chain")
    rng = np.random.default_rng(6)
    s = [laplace_symbols(rng, 40000, scale=12, density=0.4) for _ in range(4)]
    dev = [to_dev(torch_gpu, x) for x in s]
    out = torch_gpu.zeros(200000, dtype=torch_gpu.int32, device="cuda")
    nw = cc.encodeDevice(inst14, True, dev, 40000, out)
    outs = [torch_gpu.full((40000,), -1, dtype=torch_gpu.int16, device="cuda") for _ in range(4)]
    used = cc.decodeDevice(inst14, True, out, outs, 40000, wordCount=nw)
    torch_gpu.cuda.synchronize()
    assert used == nw
    for a, b in zip(outs, s):
        assert np.array_equal(a.cpu().numpy().view(np.uint16), b)
