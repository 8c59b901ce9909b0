"""Pins the CPU oracle (CPU only; no GPU needed).

Each test checks the oracle against something other than itself:
  - golden vectors written by the reference's own src/CPU coder (tests/golden/cpu_coder.npz);
  - the compiled reference coder itself (oracle/_ref), on seeded random streams, when present;
  - closed forms and invariants that a dropped term, wrong sign or wrong index would break
    (Kraft equality, optimal code cost, filter taps derived from the CDF 9/7 lifting factors,
    exact invertibility, constant/ramp inputs, zig-zag bijection).
"""
import heapq
import itertools
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "cpu_coder.npz")


def _golden():
    g = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(g["meta_json"]).decode())
    return g, meta


# ------------------------------------------------------------------------------ golden vectors
def test_golden_cpu_dialect_bitstreams():
    g, meta = _golden()
    assert len(meta) >= 8
    for case in meta:
        streams = list(g[f"{case['key']}_in"])
        for enc in case["encodings"]:
            words, nbits = O.encode(streams, rl=enc["rl"], dialect=O.DIALECT_CPU, cbs=enc["cbs"])
            ref = g[enc["array"]]
            assert nbits == enc["bits"], (case["name"], enc)
            assert np.array_equal(words, ref), (case["name"], enc)


def test_golden_cpu_dialect_decode():
    """Decoding the reference coder's bitstreams gives the original symbols."""
    g, meta = _golden()
    for case in meta:
        streams = g[f"{case['key']}_in"]
        for enc in case["encodings"]:
            outs, end = O.decode(g[enc["array"]], enc["bits"], case["nstreams"], case["count"], rl=enc["rl"],
                                 dialect=O.DIALECT_CPU, cbs=enc["cbs"])
            assert end == enc["bits"]
            for s in range(case["nstreams"]):
                assert np.array_equal(outs[s], streams[s]), (case["name"], enc, s)


def test_golden_rle_and_design():
    g, meta = _golden()
    for case in meta:
        k = case["key"]
        s0 = g[f"{k}_in"][0]
        sym, zc = O.rle_encode(s0)
        assert np.array_equal(sym, g[f"{k}_rle_sym"]) and np.array_equal(zc, g[f"{k}_rle_zc"]), case["name"]
        hist = np.bincount(s0, minlength=int(s0.max()) + 1).astype(np.uint32)
        d = O.huff_design(hist, dialect=O.DIALECT_CPU)
        assert np.array_equal(d["count_per_length"], g[f"{k}_design_cpl"]), case["name"]
        assert np.array_equal(d["symbols"], g[f"{k}_design_syms"]), case["name"]


@pytest.mark.skipif(not O.ref_available(), reason="reference coder not built (oracle/_ref)")
def test_live_reference_coder_random():
    rng = np.random.default_rng(7)
    for trial in range(60):
        ns = int(rng.integers(1, 4))
        n = int(rng.integers(1, 3000))
        scale = float(rng.choice([0.3, 2.0, 30.0, 2000.0]))
        streams = []
        for _ in range(ns):
            v = np.rint(rng.laplace(0, scale, n)).astype(np.int64)
            v[rng.random(n) < rng.random()] = 0
            streams.append(np.minimum(np.where(v < 0, -2 * v - 1, 2 * v), 16383).astype(np.uint16))
        cbs = int(rng.choice([32, 64, 128, 256]))
        for rl in (True, False):
            rw, rb = O.ref_encode(streams, rl=rl, cbs=cbs)
            ow, ob = O.encode(streams, rl=rl, dialect=O.DIALECT_CPU, cbs=cbs)
            assert rb == ob and np.array_equal(rw, ow), (trial, rl, cbs)


# ------------------------------------------------------------------------------ Huffman design
def _optimal_cost(counts):
    """Minimum total code length, by the classic two-smallest merge (cost is tie-independent)."""
    h = [int(c) for c in counts if c > 0]
    if len(h) <= 1:
        return 0
    heapq.heapify(h)
    cost = 0
    while len(h) > 1:
        a, b = heapq.heappop(h), heapq.heappop(h)
        cost += a + b
        heapq.heappush(h, a + b)
    return cost


def _brute_min_cost(counts):
    """Minimum sum(c_i * l_i) over all length vectors satisfying Kraft's inequality (tiny n)."""
    n = len(counts)
    best = None
    for ls in itertools.product(range(1, n), repeat=n):
        if sum(2.0 ** -l for l in ls) <= 1.0 + 1e-12:
            c = sum(a * b for a, b in zip(counts, ls))
            best = c if best is None else min(best, c)
    return best


def _hist_cases(rng):
    yield np.array([5], np.uint32)
    yield np.array([0, 0, 3], np.uint32)
    yield np.array([1, 1], np.uint32)
    yield np.array([1, 1, 2, 2], np.uint32)
    yield np.array([4, 4, 4, 4, 4, 4, 4, 4], np.uint32)
    for _ in range(40):
        n = int(rng.integers(2, 400))
        h = rng.integers(0, int(rng.choice([3, 50, 10000])), n).astype(np.uint32)
        h[-1] = max(h[-1], 1)
        yield h


@pytest.mark.parametrize("dialect", [O.DIALECT_GPU, O.DIALECT_CPU])
def test_design_kraft_and_optimality(dialect):
    rng = np.random.default_rng(3)
    for h in _hist_cases(rng):
        d = O.huff_design(h, dialect=dialect)
        used = np.nonzero(h)[0]
        lens = d["lengths"][used].astype(np.int64)
        assert sorted(d["symbols"].tolist()) == used.tolist()
        if used.size == 1:
            continue
        assert np.all(lens >= 1)
        kraft = np.sum(np.ldexp(1.0, -lens))
        assert kraft == 1.0
        assert int(np.sum(lens * h[used])) == _optimal_cost(h)
        # canonical: codewords are distinct, prefix-free, and increase along the symbol table
        codes = d["codewords"][d["symbols"]].astype(np.int64)
        ls = d["lengths"][d["symbols"]].astype(np.int64)
        assert np.all(np.diff(ls) >= 0)
        left = codes << (32 - ls)
        assert np.all(np.diff(left) > 0)
        assert np.all(left[1:] >= left[:-1] + (np.int64(1) << (32 - ls[:-1])))


def test_design_brute_force_tiny():
    rng = np.random.default_rng(11)
    for _ in range(25):
        n = int(rng.integers(2, 6))
        h = rng.integers(1, 9, n).astype(np.uint32)
        for dialect in (O.DIALECT_GPU, O.DIALECT_CPU):
            d = O.huff_design(h, dialect=dialect)
            assert int(np.sum(d["lengths"].astype(np.int64) * h)) == _brute_min_cost(h.tolist())


def test_design_dialects_agree_without_ties():
    """With all leaf and merged sums distinct the heap order is forced, so both dialects (range
    constructor vs push) give the same tree; the depth sort is then the same call on the same
    sequence."""
    rng = np.random.default_rng(5)
    done = 0
    while done < 30:
        n = int(rng.integers(2, 60))
        h = rng.choice(np.arange(1, 1 << 20), n, replace=False).astype(np.uint32)
        # reject histograms with a tie among partial sums of the merge sequence
        q = sorted(h.tolist())
        seen = set(q)
        ok = True
        hh = list(q)
        heapq.heapify(hh)
        while len(hh) > 1:
            s = heapq.heappop(hh) + heapq.heappop(hh)
            if s in seen:
                ok = False
                break
            seen.add(s)
            heapq.heappush(hh, s)
        if not ok:
            continue
        a = O.huff_design(h, dialect=O.DIALECT_GPU)
        b = O.huff_design(h, dialect=O.DIALECT_CPU)
        assert np.array_equal(a["lengths"], b["lengths"])
        assert np.array_equal(a["symbols"], b["symbols"])
        done += 1


def _read_bits(words, pos, n):
    v = 0
    for i in range(n):
        p = pos + i
        v = (v << 1) | ((int(words[p // 32]) >> (31 - p % 32)) & 1)
    return v, pos + n


def test_gpu_dialect_header_fields():
    """The GPU-dialect table header read back field by field (MSB-first words): 5-bit length count,
    log2-bit counts, log2-bit symbol count, 4-bit symbol width (16 stored as 0), symbols."""
    rng = np.random.default_rng(9)
    for log2 in (14, 16):
        s = rng.integers(0, 300, 2000).astype(np.uint16)
        s[:3] = [0, 65535 if log2 == 16 else 16000, 7]
        words, _ = O.encode([s], rl=False, dialect=O.DIALECT_GPU, log2=log2, cbs=128)
        hist = np.bincount(s, minlength=int(s.max()) + 1).astype(np.uint32)
        d = O.huff_design(hist, dialect=O.DIALECT_GPU)
        pos = 0
        nlen, pos = _read_bits(words, pos, 5)
        assert nlen == d["count_per_length"].size
        for c in d["count_per_length"]:
            v, pos = _read_bits(words, pos, log2)
            assert v == c
        nsym, pos = _read_bits(words, pos, log2)
        assert nsym == d["symbols"].size
        sb, pos = _read_bits(words, pos, 4)
        want = int(s.max()).bit_length()
        assert sb == (0 if want == 16 else want)
        for sym in d["symbols"]:
            v, pos = _read_bits(words, pos, want)
            assert v == sym


# ------------------------------------------------------------------------------ RLE, histogram, pack
def test_rle_brute_force_tiny():
    """All streams of length <= 7 over {0,1,2}: pairs are (symbol, zeros before it), trailing zeros
    dropped, and decoding restores the stream."""
    for n in range(0, 8):
        for t in itertools.product((0, 1, 2), repeat=n):
            x = np.array(t, dtype=np.uint16)
            sym, zc = O.rle_encode(x)
            nz = np.nonzero(x)[0]
            assert np.array_equal(sym, x[nz])
            gaps = np.diff(np.concatenate([[-1], nz])) - 1
            assert np.array_equal(zc, gaps.astype(np.uint16))
            assert np.array_equal(O.rle_decode(sym, zc, n), x)


def test_rle_long_runs_use_fillers():
    x = np.zeros(1000, np.uint16)
    x[600] = 3
    sym, zc = O.rle_encode(x)
    # 600 zeros: two (0, 255) fillers (each stands for 255 zeros plus the zero symbol itself = 256)
    assert sym.tolist() == [0, 0, 3] and zc.tolist() == [255, 255, 600 - 512]
    assert np.array_equal(O.rle_decode(sym, zc, x.size), x)


def test_histogram_counts_and_pad_quirk():
    rng = np.random.default_rng(2)
    x = rng.integers(0, 500, 1001).astype(np.uint16)
    h = O.histogram(x, int(x.max()), 500)
    assert np.array_equal(h, np.bincount(x, minlength=500))
    # Symbol16 call whose max symbol is 65535 with an odd element count: one pad count in bin 65535
    x[5] = 65535
    h = O.histogram(x, 65535, 65536)
    want = np.bincount(x, minlength=65536)
    want[65535] += 1
    assert np.array_equal(h, want)
    h = O.histogram(x[:1000], 65535, 65536)
    assert np.array_equal(h, np.bincount(x[:1000], minlength=65536))


def test_pack_inc16_closed_form():
    v = np.cumsum(np.random.default_rng(1).integers(0, 5000, 300)).astype(np.uint32)
    inc = O.pack_inc16(v)
    assert inc[0] == (v[0] & 0xFFFF)
    assert np.array_equal(inc[1:], np.diff(v).astype(np.uint16))
    assert np.array_equal(O.unpack_inc16(inc), v)


# ------------------------------------------------------------------------------ transforms
def test_dwt53_invertible_and_closed_forms():
    rng = np.random.default_rng(4)
    for shape in [(8, 8), (16, 32), (6, 10, 4), (32, 16, 8)]:
        for x in (rng.integers(-3000, 3000, shape), rng.integers(-32768, 32768, shape)):
            x = x.astype(np.int16)
            c = O.dwt53_forward(x)
            assert np.array_equal(O.dwt53_inverse(c), x)
    # constant: highpass 0, lowpass unchanged
    c = O.dwt53_forward(np.full((16, 16), -77, np.int16))
    assert np.all(c[:8, :8] == -77) and np.all(c[8:, :] == 0) and np.all(c[:, 8:] == 0)
    # ramp along x: highpass zero except the last (mirrored) coefficient of each row
    r = np.tile(np.arange(32, dtype=np.int16) * 3 - 40, (4, 1))
    c = O.dwt53_forward(r)
    hp = c[:2, 16:]
    assert np.all(hp[:, :-1] == 0) and np.all(hp[:, -1] == 3)


def _wrap16(v):
    return ((int(v) + 32768) % 65536) - 32768


def _round_half_away(num, den):
    q, r = divmod(abs(num), den)
    q += 1 if 2 * r >= den else 0
    return q if num >= 0 else -q


def test_dwt53_rows_brute_force():
    """Rows of the 5/3 analysis against the lifting written with round-half-away-from-zero
    division.  Inputs stay within +-8000, so every X result is below 2^14 in magnitude and the Y
    pass of two equal rows is exact: it leaves row 0 as the X result and row 1 zero.  (Wrap-around
    of the int16 arithmetic is covered by the invertibility test.)"""
    rng = np.random.default_rng(13)
    for n in (2, 4, 6, 10, 16):
        for _ in range(40):
            x = rng.integers(-8000, 8001, n).astype(np.int16)
            if rng.random() < 0.3:
                x = rng.choice([-8000, -1, 0, 1, 2, 3, 8000], n).astype(np.int16)
            got = O.dwt53_forward(np.stack([x, x]))
            m = n // 2
            xi = [int(v) for v in x]
            d = [_wrap16(xi[2 * k + 1] - _round_half_away(_wrap16(xi[2 * k] + xi[_mirror(2 * k + 2, n)]), 2))
                 for k in range(m)]
            s = [_wrap16(xi[2 * k] + _round_half_away(_wrap16(d[max(k - 1, 0)] + d[k]), 4)) for k in range(m)]
            assert got[0].tolist() == s + d, (x.tolist(), got[0].tolist(), s + d)
            assert np.all(got[1] == 0)


def _cdf97_taps():
    """Analysis taps from the CDF 9/7 lifting factors (Daubechies-Sweldens), normalised to DC
    gain 1 (lowpass) and Nyquist gain 1 (highpass): lowpass taps at x[2k-4..2k+4], highpass at
    x[2k+1-3..2k+1+3]."""
    a, b, g, d = -1.586134342059924, -0.052980118572961, 0.882911075530934, 0.443506852043971
    n = 64

    def lift(x):
        x = x.copy()
        x[1::2] += a * (x[0::2] + np.roll(x[0::2], -1))
        x[0::2] += b * (x[1::2] + np.roll(x[1::2], 1))
        x[1::2] += g * (x[0::2] + np.roll(x[0::2], -1))
        x[0::2] += d * (x[1::2] + np.roll(x[1::2], 1))
        return x

    lo = np.zeros(9)
    hi = np.zeros(7)
    for j in range(-4, 5):
        e = np.zeros(n); e[32 + j] = 1.0
        lo[j + 4] = lift(e)[32]
    for j in range(-3, 4):
        e = np.zeros(n); e[33 + j] = 1.0
        hi[j + 3] = lift(e)[33]
    lo /= lo.sum()
    hi /= np.sum(hi * (-1.0) ** np.abs(np.arange(-3, 4)))
    return lo, hi


def _mirror(j, n):
    while j < 0 or j > n - 1:
        j = -j if j < 0 else 2 * (n - 1) - j
    return j


def _fwd_axis(x, axis, lo, hi):
    x = np.moveaxis(x, axis, -1)
    n = x.shape[-1]
    m = n // 2
    out = np.zeros_like(x)
    for k in range(m):
        out[..., k] = sum(lo[i] * x[..., _mirror(2 * k - 4 + i, n)] for i in range(9))
        out[..., m + k] = sum(hi[i] * x[..., _mirror(2 * k + 1 - 3 + i, n)] for i in range(7))
    return np.moveaxis(out, -1, axis)


def test_dwt97_against_lifting_derived_filters():
    lo, hi = _cdf97_taps()
    rng = np.random.default_rng(6)
    for shape in [(16, 24), (8, 12, 16)]:
        x = rng.standard_normal(shape)
        want = x
        for ax in range(x.ndim - 1, -1, -1):
            want = _fwd_axis(want, ax, lo, hi)
        got = O.dwt97_forward(x, dtype=np.float64)
        # the reference stores 7-decimal taps: about 5e-8 per tap, ~1e-6 after 2-3 passes
        assert np.max(np.abs(got - want)) < 3e-6 * np.max(np.abs(x)) * x.ndim
        got32 = O.dwt97_forward(x.astype(np.float32), dtype=np.float32)
        assert np.max(np.abs(got32 - got)) < 1e-5 * np.max(np.abs(x))


def test_dwt97_reconstruction_and_constant():
    rng = np.random.default_rng(8)
    for shape in [(32, 32), (16, 16, 16), (8, 24, 12)]:
        x = rng.standard_normal(shape)
        r64 = O.dwt97_inverse(O.dwt97_forward(x, np.float64), np.float64)
        assert np.max(np.abs(r64 - x)) < 2e-6 * x.ndim
        r32 = O.dwt97_inverse(O.dwt97_forward(x.astype(np.float32)), np.float32)
        assert np.max(np.abs(r32 - x)) < 1e-5 * x.ndim
        # linearity (exact up to rounding)
        y = rng.standard_normal(shape)
        lhs = O.dwt97_forward(2.0 * x - 3.0 * y, np.float64)
        rhs = 2.0 * O.dwt97_forward(x, np.float64) - 3.0 * O.dwt97_forward(y, np.float64)
        assert np.max(np.abs(lhs - rhs)) < 1e-12 * 10
    c = O.dwt97_forward(np.full((16, 16), 2.5), np.float64)
    # lowpass taps sum to 1.0000000 and highpass taps to -0.0000001 in the stored decimals
    assert np.allclose(c[:8, :8], 2.5, atol=1e-6)
    assert np.max(np.abs(c[8:, :])) < 1e-6 and np.max(np.abs(c[:, 8:])) < 1e-6


def test_quantizer_closed_forms():
    step = np.float32(0.01)
    inv = np.float32(1.0) / step
    rng = np.random.default_rng(12)
    v = np.concatenate([rng.standard_normal(5000).astype(np.float32),
                        np.array([0.0, -0.0, 0.0099, -0.0099, 0.01, -0.01, 0.02, 3.0e9, -3.0e9], np.float32)])
    for qtype in (0, 1):
        s = O.quantize(v, float(step), qtype=qtype, dtype=np.uint32)
        prod = (v * inv).astype(np.float32)
        if qtype == 0:
            t = prod
        else:
            t = (prod + np.sign(v).astype(np.float32) * np.float32(0.5)).astype(np.float32)
        q = np.clip(np.trunc(t.astype(np.float64)), -2 ** 31, 2 ** 31 - 1).astype(np.int64)
        want = np.where(q < 0, -2 * q - 1, 2 * q)
        assert np.array_equal(s.astype(np.int64), want)
        # reconstruction lies inside the quantization bin
        r = O.unquantize(s, float(step), qtype=qtype).astype(np.float64)
        ok = np.abs(v) < 1e6
        assert np.all(np.abs(r[ok] - v[ok]) <= float(step) * (1.0 if qtype == 0 else 0.5) + 1e-6)
    # zig-zag is a bijection on int16
    a = np.arange(-32768, 32768, dtype=np.int16)
    s = O.symbolize(a)
    assert np.array_equal(np.sort(s), np.arange(65536, dtype=np.uint16))
    assert np.array_equal(O.unsymbolize(s), a)


def test_quantize_to_short_closed_forms():
    """quantizeToShort / unquantizeFromShort: the quantizer's int as a wrapped int16; agrees with
    the symbol quantizer through the zig-zag map wherever the int fits."""
    step = np.float32(0.5)
    v = np.array([0.74, -0.76, 0.25, -0.25, 0.0, 1.5, -1.5, 0.2499, -0.2499], np.float32)
    assert O.quantize_to_short(v, float(step), qtype=1).tolist() == [1, -2, 1, -1, 0, 3, -3, 0, 0]
    assert O.quantize_to_short(v, float(step), qtype=0).tolist() == [1, -1, 0, 0, 0, 3, -3, 0, 0]
    # the int result is stored as a short: 40000 -> -25536, -40001 -> 25535
    w = np.array([40000.0, -40001.0, 32767.0, -32768.0], np.float32)
    assert O.quantize_to_short(w, 1.0, qtype=1).tolist() == [-25536, 25535, 32767, -32768]
    rng = np.random.default_rng(21)
    x = (rng.standard_normal(20000) * 300).astype(np.float32)
    for qtype in (0, 1):
        q = O.quantize_to_short(x, 0.37, qtype=qtype)
        assert np.array_equal(O.symbolize(q).astype(np.uint32), O.quantize(x, 0.37, qtype=qtype, dtype=np.uint32))
        r = O.unquantize_from_short(q, 0.37, qtype=qtype)
        assert np.array_equal(r, O.unquantize(O.symbolize(q), 0.37, qtype=qtype))
    # uniform reconstruction is the bin centre: q * step
    q = np.arange(-300, 300, dtype=np.int16)
    assert np.array_equal(O.unquantize_from_short(q, 0.25, qtype=1), (q.astype(np.float32) * np.float32(0.25)))


@pytest.mark.parametrize("levels,rle_only", [(1, False), (2, False), (2, True), (3, True)])
def test_quant_first_pipeline_reconstructs_the_quantized_volume(levels, rle_only):
    """The QuantFirst codec loses only in the quantizer (5/3 lifting and the entropy coder are
    exact), so its round trip equals unquantize(quantize(volume))."""
    from oracle import pipeline as P
    rng = np.random.default_rng(levels)
    vol = np.cumsum(rng.standard_normal((16, 24, 32)), axis=2).astype(np.float32)
    step = 0.05
    w = P.compress_volume_float_quant_first(vol, levels, step, rle_only_lvl0=rle_only)
    got = P.decompress_volume_float_quant_first(w, vol.shape, levels, step, rle_only_lvl0=rle_only)
    want = O.unquantize_from_short(O.quantize_to_short(vol, step, qtype=1), step, qtype=1)
    assert np.array_equal(got, want)
    # the RLE-only-on-level-0 flag changes the bitstream above level 0 only
    if levels > 1:
        other = P.compress_volume_float_quant_first(vol, levels, step, rle_only_lvl0=not rle_only)
        assert not np.array_equal(other, w)
