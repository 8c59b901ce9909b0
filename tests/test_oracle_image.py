"""Pinning of the image-path / bit-plane oracle (oracle/image.py) by closed forms, independent
restatements and invariants, so that a dropped term, a wrong sign or a wrong rounding fails."""
import numpy as np
import pytest

from oracle import image as im
from oracle import pyoracle as o


def test_float_to_byte_closed_forms():
    v = np.array([-1000.0, -128.5, -128.0, -0.51, -0.5, 0.0, 0.49, 0.5, 126.49, 126.5, 127.0, 1e9, np.nan],
                 dtype=np.float32)
    want = [0, 0, 0, 127, 128, 128, 128, 129, 254, 255, 255, 255, 0]
    assert im.float_to_byte(v).tolist() == want
    # on half-integers and integers: byteToFloat(floatToByte(x)) = floor(x + 0.5) clamped to [-128, 127]
    x = np.arange(-300, 300, 0.5, dtype=np.float32)
    assert np.array_equal(im.byte_to_float(im.float_to_byte(x)), np.clip(np.floor(x + 0.5), -128, 127))
    assert np.array_equal(im.byte_to_float(np.arange(256)), np.arange(256) - 128.0)


def test_float_to_byte_2d_writes_one_channel():
    rng = np.random.default_rng(1)
    plane = (rng.standard_normal((5, 7)) * 60).astype(np.float32)
    img = np.full((5, 7, 3), 9, dtype=np.uint8)
    im.float_to_byte_2d(plane, 3, 1, out=img)
    assert np.all(img[:, :, 0] == 9) and np.all(img[:, :, 2] == 9)
    assert np.array_equal(im.byte_to_float_2d(img, 1), np.clip(np.floor(plane + 0.5), -128, 127))


@pytest.mark.parametrize("qtype", [0, 1])
def test_quantize_difference_reduces_to_the_quantizer(qtype):
    rng = np.random.default_rng(2)
    data = (rng.standard_normal((6, 9)) * 40).astype(np.float32)
    step = 0.25
    # reference 128 is the zero offset: plain quantization of the data
    sym, rec = im.quantize_difference(data, step, np.full(data.shape, 128, np.uint8), qtype)
    assert np.array_equal(sym, o.quantize(data, step, qtype=qtype))
    assert np.array_equal(rec, o.unquantize(sym, step, qtype=qtype))
    # a reference c shifts the data by c - 128 before quantization
    ref = rng.integers(0, 256, data.shape).astype(np.uint8)
    sym2, rec2 = im.quantize_difference(data, step, ref, qtype)
    shifted = data - (ref.astype(np.float32) - 128.0)
    assert np.array_equal(sym2, o.quantize(shifted, step, qtype=qtype))
    assert np.array_equal(im.unquantize_difference(sym2, step, ref, qtype), rec2)
    # reconstruction error of a midtread quantizer is within one step (deadzone: within the zero bin)
    assert float(np.abs(rec2 - data).max()) <= step * (1.0 if qtype == 0 else 0.5) + 1e-4


def test_ycocg_closed_forms_and_round_trip():
    r, b = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    worst = np.zeros(3, dtype=int)
    for g in range(0, 256, 17):
        px = np.stack([r, np.full_like(r, g), b], -1).astype(np.uint8)
        y = im.rgb_to_ycocg(px).astype(np.int64)
        ri, gi, bi = r.astype(np.int64), g, b.astype(np.int64)
        # the matrix of the transform, rounded: Y = floor((R+2G+B+2)/4), Co = 127 + trunc((R-B+1)/2),
        # Cg = 127 + trunc((-R+2G-B+2)/4)
        assert np.array_equal(y[..., 0], (ri + 2 * gi + bi + 2) // 4)
        assert np.array_equal(y[..., 1], 127 + np.trunc((ri - bi + 1) / 2.0).astype(np.int64))
        assert np.array_equal(y[..., 2], 127 + np.trunc((-ri + 2 * gi - bi + 2) / 4.0).astype(np.int64))
        back = im.ycocg_to_rgb(y.astype(np.uint8)).astype(int)
        worst = np.maximum(worst, np.abs(back - px.astype(int)).max(axis=(0, 1)))
    assert worst.tolist() == [1, 1, 2]
    gray = np.stack([np.arange(256)] * 3, -1).astype(np.uint8)
    assert np.array_equal(im.rgb_to_ycocg(gray), np.stack([np.arange(256), np.full(256, 127), np.full(256, 127)], -1))
    assert np.array_equal(im.ycocg_to_rgb(im.rgb_to_ycocg(gray)), gray)
    rgba = np.array([[10, 200, 30, 77]], dtype=np.uint8)
    assert im.rgb_to_ycocg(rgba)[0, 3] == 77 and im.ycocg_to_rgb(rgba)[0, 3] == 77


def test_bit_planes():
    rng = np.random.default_rng(3)
    for n in (1, 31, 32, 33, 4096, 10007):
        d = rng.integers(-32768, 32768, n).astype(np.int16)
        bits = im.get_lsb(d)
        # independent packing: little-endian bit order within little-endian words
        packed = np.packbits((d & 1).astype(np.uint8), bitorder="little")
        packed = np.concatenate([packed, np.zeros((-packed.size) % 4, np.uint8)])
        assert np.array_equal(bits, packed.view("<u4"))
        for shift in (0, 1, -3, 32767):
            out, b2 = im.remove_lsb(d, shift)
            v = (d.astype(np.int64) + shift + 32768) % 65536 - 32768
            assert np.array_equal(out, (v // 2).astype(np.int16))  # floor division
            assert np.array_equal(out, (v >> 1).astype(np.int16))
            assert np.array_equal(b2, im.get_lsb(v.astype(np.int16)))
            assert np.array_equal(im.append_lsb(out, b2, -shift), d)
        assert np.array_equal(im.shift_elems(d, 5), ((d.astype(np.int64) + 5 + 32768) % 65536 - 32768).astype(np.int16))


def test_byte_dwt_constant_and_reconstruction():
    img = np.full((16, 24), 200, dtype=np.uint8)
    c = im.dwt97_forward_from_byte(img)
    # analysis lowpass gain 1, highpass gain 0
    assert np.allclose(c[:8, :12], 72.0, atol=1e-3) and np.allclose(c[8:, :], 0, atol=1e-3)
    assert np.allclose(c[:, 12:], 0, atol=1e-3)
    rng = np.random.default_rng(4)
    smooth = np.clip(128 + 60 * np.sin(np.arange(32)[:, None] / 5.0) * np.cos(np.arange(40)[None, :] / 7.0)
                     + rng.integers(-3, 4, (32, 40)), 0, 255).astype(np.uint8)
    assert np.array_equal(im.dwt97_inverse_to_byte(im.dwt97_forward_from_byte(smooth)), smooth)
    vol = rng.integers(0, 256, (8, 10, 12)).astype(np.uint8)
    assert np.array_equal(im.dwt97_inverse_to_byte(im.dwt97_forward_from_byte(vol)), vol)


def _mirror(j, n):
    while j < 0 or j >= n:
        j = -j if j < 0 else 2 * (n - 1) - j
    return j


def _div2(v):  # DWTCommon.cuh div2: round half away from zero
    return int(np.trunc((v + (1 if v >= 0 else -1)) / 2))


def _div4(v):
    return int(np.trunc((v + (2 if v >= 0 else -2)) / 4))


def _lift53_low(line):
    n = len(line)
    x = [int(v) for v in line]
    d = {k: x[_mirror(k, n)] - _div2(x[_mirror(k - 1, n)] + x[_mirror(k + 1, n)]) for k in range(-1, n, 2)}
    return [np.int16(x[2 * k] + _div4(d[2 * k - 1] + d[2 * k + 1])) for k in range(n // 2)]


def test_lowpass_only_53_restated():
    rng = np.random.default_rng(5)
    a = rng.integers(-900, 900, (10, 14)).astype(np.int16)
    rows = np.array([_lift53_low(r) for r in a], dtype=np.int16)
    ll = np.array([_lift53_low(c) for c in rows.T], dtype=np.int16).T
    assert np.array_equal(im.dwt53_lowpass_only(a), ll)


def test_lowpass_only_97_restated():
    taps = np.array([0.0267488, -0.0168641, -0.0782233, 0.2668641, 0.6029490, 0.2668641, -0.0782233, -0.0168641,
                     0.0267488])
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, (12, 18)).astype(np.uint8)
    x = img.astype(np.float64) - 128.0

    def low(line):
        n = len(line)
        return np.array([sum(taps[i] * line[_mirror(2 * k - 4 + i, n)] for i in range(9)) for k in range(n // 2)])

    rows = np.array([low(r) for r in x])
    ll = np.array([low(c) for c in rows.T]).T
    got = im.dwt97_lowpass_only_from_byte(img).astype(np.float64)
    assert np.abs(got - ll).max() <= 1e-4 * max(1.0, np.abs(ll).max())
