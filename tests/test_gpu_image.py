"""GPU parity of the image-path utilities, bit planes, lowpass-only transforms and the Timing
API (include/cudaCompress/util/{DWT,Quantize,YCoCg,Bits}.h, include/cudaCompress/Timing.h)
against oracle/image.py, through the C ABI."""
import numpy as np
import pytest

from cudacompress_amd import synthetic as S
from oracle import image as im
from oracle import pyoracle as o
from tests.gpu_util import log, to_dev, to_host

pytestmark = pytest.mark.gpu

F32_TOL = 1e-6


def _close(got, want):
    scale = max(1.0, float(np.abs(want).max()))
    err = float(np.abs(got.astype(np.float64) - want.astype(np.float64)).max())
    log(f"   max|err|={err:.3e} scale={scale:.3e}")
    assert err <= F32_TOL * scale


def _image(rng, shape):
    return rng.integers(0, 256, shape).astype(np.uint8)


def test_float_byte_planes(cc, torch_gpu):
    torch = torch_gpu
    log("float <-> byte planes")
    rng = np.random.default_rng(11)
    sy, sx, ch = 37, 70, 3
    plane = (rng.standard_normal((sy, sx + 5)) * 90).astype(np.float32)
    plane[0, :6] = [-128.5, -0.5, -0.51, 126.5, 1e9, -1e9]
    img = to_dev(torch, np.full((sy, sx, ch), 7, np.uint8))
    cc.util.floatToByte2D(img, ch, 1, to_dev(torch, plane), sx, sy, sx + 5)
    torch.cuda.synchronize()
    want = np.full((sy, sx, ch), 7, np.uint8)
    im.float_to_byte_2d(np.ascontiguousarray(plane[:, :sx]), ch, 1, out=want)
    got = to_host(img)
    assert np.array_equal(got, want)
    back = torch.full((sy, sx + 3), -1.0, dtype=torch.float32, device="cuda")
    cc.util.byteToFloat2D(back, img, ch, 1, sx, sy, sx + 3)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(back)[:, :sx], im.byte_to_float_2d(want, 1))
    assert np.all(to_host(back)[:, sx:] == -1.0)


@pytest.mark.parametrize("qtype", [0, 1], ids=["deadzone", "uniform"])
def test_quantize_difference(cc, torch_gpu, qtype):
    torch = torch_gpu
    log("quantize difference", qtype)
    rng = np.random.default_rng(12 + qtype)
    sy, sx, ch, c = 29, 66, 4, 2
    data = (rng.standard_normal((sy, sx + 2)) * 50).astype(np.float32)
    ref = _image(rng, (sy, sx, ch))
    step = 0.3
    sym_want, rec_want = im.quantize_difference(np.ascontiguousarray(data[:, :sx]), step, ref[:, :, c], qtype)
    dref = to_dev(torch, ref)
    sym = torch.zeros((sy, sx), dtype=torch.int16, device="cuda")
    cc.util.quantizeDifferenceToSymbols2D(sym, to_dev(torch, data), step, dref, ch, c, sx, sy, sx + 2, qtype)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(sym, np.uint16), sym_want)
    d2 = to_dev(torch, data)
    sym2 = torch.zeros_like(sym)
    cc.util.quantizeDifferenceToSymbolsRoundtrip2D(sym2, d2, step, dref, ch, c, sx, sy, sx + 2, qtype)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(sym2, np.uint16), sym_want)
    assert np.array_equal(to_host(d2)[:, :sx], rec_want)
    assert np.array_equal(to_host(d2)[:, sx:], data[:, sx:])
    out = torch.zeros((sy, sx + 1), dtype=torch.float32, device="cuda")
    cc.util.unquantizeDifferenceFromSymbols2D(out, sym, step, dref, ch, c, sx, sy, sx + 1, qtype)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(out)[:, :sx], im.unquantize_difference(sym_want, step, ref[:, :, c], qtype))


@pytest.mark.parametrize("channels", [3, 4])
def test_ycocg(cc, torch_gpu, channels):
    torch = torch_gpu
    log("ycocg", channels)
    rng = np.random.default_rng(14)
    n = 100003
    px = _image(rng, (n, channels))
    px[:4, :3] = [[0, 0, 0], [255, 255, 255], [255, 0, 255], [0, 255, 0]]
    src = to_dev(torch, px)
    y = torch.zeros_like(src)
    cc.util.convertRGBtoYCoCg(y, src, n, channels)
    back = torch.zeros_like(src)
    cc.util.convertYCoCgtoRGB(back, y, n, channels)
    torch.cuda.synchronize()
    want = im.rgb_to_ycocg(px)
    assert np.array_equal(to_host(y), want)
    assert np.array_equal(to_host(back), im.ycocg_to_rgb(want))


@pytest.mark.parametrize("n", [1, 33, 2048, 4096 * 3, 100003], ids=lambda n: f"n{n}")
def test_bit_planes(cc, torch_gpu, n):
    torch = torch_gpu
    log("bit planes", n)
    rng = np.random.default_rng(15)
    d = rng.integers(-32768, 32768, n).astype(np.int16)
    d[: min(n, 4)] = [32767, -32768, -1, 0][: min(n, 4)]
    nw = (n + 31) // 32
    sentinel = 0x5A5A5A5A
    dd = to_dev(torch, d)
    bits = torch.full((nw + 2,), sentinel, dtype=torch.int32, device="cuda")
    cc.util.getLSB(dd, bits, n)
    torch.cuda.synchronize()
    hb = to_host(bits, np.uint32)
    assert np.array_equal(hb[:nw], im.get_lsb(d)) and np.all(hb[nw:] == sentinel)
    for shift in (0, 3, -32768):
        x = to_dev(torch, d)
        b = torch.full((nw + 2,), sentinel, dtype=torch.int32, device="cuda")
        cc.util.removeLSB(x, b, n, shift)
        torch.cuda.synchronize()
        want, wb = im.remove_lsb(d, shift)
        assert np.array_equal(to_host(x), want)
        hb = to_host(b, np.uint32)
        assert np.array_equal(hb[:nw], wb) and np.all(hb[nw:] == sentinel)
        cc.util.appendLSB(x, b, n, -shift)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(x), d)
    x = to_dev(torch, d)
    cc.util.removeLSB(x, None, n, 1)
    cc.util.shiftElems(x, n, -7)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(x), im.shift_elems(im.remove_lsb(d, 1)[0], -7))


@pytest.mark.parametrize("shape,ch", [((48, 64), 1), ((40, 72), 3), ((16, 24, 32), 1), ((8, 16, 40), 4)],
                         ids=["2d", "2d-ch3", "3d", "3d-ch4"])
def test_dwt97_byte_variants(cc, torch_gpu, shape, ch):
    """ForwardFromByte equals the float transform of (byte - 128); InverseToByte equals the float
    inverse rounded and clamped; both within the float tolerance of the oracle."""
    torch = torch_gpu
    log("dwt97 byte", shape, ch)
    rng = np.random.default_rng(16)
    img = _image(rng, shape + (ch,))
    c = ch - 1  # transform the last channel
    plane = np.ascontiguousarray(img[..., c])
    is3d = len(shape) == 3
    dimg = to_dev(torch, img)
    src = dimg.view(-1)[c:]
    out = torch.zeros(shape, dtype=torch.float32, device="cuda")
    b1, b2 = torch.empty_like(out), torch.empty_like(out)
    if is3d:
        z, y, x = shape
        cc.util.dwtFloat3DForwardFromByte(out, b2, b1, src, x, y, z, ch)
    else:
        y, x = shape
        cc.util.dwtFloat2DForwardFromByte(out, b1, src, x, y, ch)
    torch.cuda.synchronize()
    coef = to_host(out)
    _close(coef, im.dwt97_forward_from_byte(plane))
    # the float path on the converted plane gives the same bits
    fout = torch.zeros_like(out)
    fsrc = to_dev(torch, im.byte_to_float(plane))
    if is3d:
        cc.util.dwtFloat3DForward(fout, b2, b1, fsrc, x, y, z)
    else:
        cc.util.dwtFloat2DForward(fout, b1, fsrc, x, y)
    torch.cuda.synchronize()
    assert np.array_equal(coef, to_host(fout))
    # inverse to byte into channel 0 of a fresh interleaved image
    dst = torch.full(shape + (ch,), 3, dtype=torch.uint8, device="cuda")
    finv = torch.zeros_like(out)
    if is3d:
        cc.util.dwtFloat3DInverseToByte(dst.view(-1), b2, b1, out, x, y, z, ch)
        cc.util.dwtFloat3DInverse(finv, b2, b1, out, x, y, z)
    else:
        cc.util.dwtFloat2DInverseToByte(dst.view(-1), b1, out, x, y, ch)
        cc.util.dwtFloat2DInverse(finv, b1, out, x, y)
    torch.cuda.synchronize()
    got = to_host(dst)
    assert np.array_equal(got[..., 0], im.float_to_byte(to_host(finv)))
    assert np.array_equal(got[..., 0], plane)  # perfect reconstruction survives the rounding
    if ch > 1:
        assert np.all(got[..., 1:] == 3)


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 48)], ids=["3d", "2d"])
def test_dwt97_inverse_from_symbols_to_byte(cc, torch_gpu, shape):
    torch = torch_gpu
    log("from symbols to byte", shape)
    coef = S.gaussian_field(shape, 7) * 40.0
    q = 0.5
    is3d = len(shape) == 3
    if is3d:
        z, y, x = shape
        bands = [(slice(bz * (z // 2), (bz + 1) * (z // 2)), slice(by * (y // 2), (by + 1) * (y // 2)),
                  slice(bx * (x // 2), (bx + 1) * (x // 2))) for b in range(1, 8)
                 for bx, by, bz in [(b % 2, b // 2 % 2, b // 4)]]
    else:
        y, x = shape
        bands = [(slice(by * (y // 2), (by + 1) * (y // 2)), slice(bx * (x // 2), (bx + 1) * (x // 2)))
                 for b in range(1, 4) for bx, by in [(b % 2, b // 2)]]
    syms = [to_dev(torch, o.quantize(np.ascontiguousarray(coef[b]).ravel(), q)) for b in bands]
    ptrs = torch.tensor([t.data_ptr() for t in syms], dtype=torch.int64, device="cuda")
    lp = to_dev(torch, coef)
    fout = torch.zeros(shape, dtype=torch.float32, device="cuda")
    bout = torch.zeros(shape, dtype=torch.uint8, device="cuda")
    b1, b2 = torch.empty_like(fout), torch.empty_like(fout)
    if is3d:
        cc.util.dwtFloat3DInverseFromSymbols(fout, b2, b1, lp, ptrs, q, x, y, z)
        cc.util.dwtFloat3DInverseFromSymbolsToByte(bout, b2, b1, lp, ptrs, q, x, y, z)
    else:
        cc.util.dwtFloat2DInverseFromSymbols(fout, b1, lp, ptrs, q, x, y)
        cc.util.dwtFloat2DInverseFromSymbolsToByte(bout, b1, lp, ptrs, q, x, y)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(bout), im.float_to_byte(to_host(fout)))


@pytest.mark.parametrize("shape,ch", [((64, 96), 1), ((30, 518), 3), ((128, 1024), 4)],
                         ids=["64x96", "30x518-ch3", "128x1024-ch4"])
def test_dwt97_lowpass_only_from_byte(cc, torch_gpu, shape, ch):
    torch = torch_gpu
    log("dwt97 lowpass only", shape, ch)
    rng = np.random.default_rng(17)
    y, x = shape
    img = _image(rng, shape + (ch,))
    plane = np.ascontiguousarray(img[..., 0])
    pitch = x // 2 + 3
    dst = torch.full((y // 2, pitch), -5.0, dtype=torch.float32, device="cuda")
    buf = torch.zeros((y, pitch), dtype=torch.float32, device="cuda")
    cc.util.dwtFloat2DForwardLowpassOnlyFromByte(dst, buf, to_dev(torch, img), x, y, ch, pitch)
    full = torch.zeros(shape, dtype=torch.float32, device="cuda")
    b1 = torch.empty_like(full)
    cc.util.dwtFloat2DForwardFromByte(full, b1, to_dev(torch, img), x, y, ch)
    torch.cuda.synchronize()
    got = to_host(dst)
    assert np.array_equal(got[:, : x // 2], to_host(full)[: y // 2, : x // 2])  # the LL band, bit for bit
    assert np.all(got[:, x // 2:] == -5.0)
    _close(got[:, : x // 2], im.dwt97_lowpass_only_from_byte(plane))


@pytest.mark.parametrize("shape", [(64, 1024), (96, 200), (6, 8)], ids=["64x1024", "96x200", "6x8"])
def test_dwt53_lowpass_only(cc, torch_gpu, shape):
    torch = torch_gpu
    log("dwt53 lowpass only", shape)
    rng = np.random.default_rng(18)
    y, x = shape
    a = rng.integers(-32768, 32768, (y, x + 4)).astype(np.int16)
    pitch = x // 2 + 1
    dst = torch.full((y // 2, pitch), 99, dtype=torch.int16, device="cuda")
    buf = torch.zeros((y, pitch), dtype=torch.int16, device="cuda")
    cc.util.dwtIntForwardLowpassOnly(dst, buf, to_dev(torch, a), x, y, pitch, x + 4)
    torch.cuda.synchronize()
    got = to_host(dst)
    assert np.array_equal(got[:, : x // 2], im.dwt53_lowpass_only(np.ascontiguousarray(a[:, :x])))
    assert np.all(got[:, x // 2:] == 99)


def test_timing_sections(cc, torch_gpu):
    torch = torch_gpu
    log("timing")
    rng = np.random.default_rng(19)
    n = 1 << 14
    streams = [to_dev(torch, (rng.integers(0, 3, n) * (rng.random(n) < 0.3)).astype(np.uint16)) for _ in range(3)]
    inst = cc.createInstance(0, 3, n, 128, 14)
    cc.setTimingDetail(inst, cc.TIMING_DETAIL_MEDIUM)
    bs = cc.BitStream()
    assert cc.encodeRLHuff(inst, bs, streams, n)
    rs = cc.BitStream(bs.words(), bs.getBitSize())
    outs = [torch.zeros(n, dtype=torch.int16, device="cuda") for _ in range(3)]
    cc.decodeRLHuff(inst, rs, outs, n)
    torch.cuda.synchronize()
    names, times = cc.getTimings(inst)
    log("   " + ", ".join(f"{a}={b:.3f}" for a, b in zip(names, times)))
    for want in ("Encode RLHuff", "Encode: Histograms", "Encode: Design Huffman Tables", "Decode RLHuff",
                 "Decode: Huffman Decode", "Decode: Run Length Decode"):
        assert want in names
    assert all(t >= 0.0 for t in times)
    call = times[names.index("Encode RLHuff")]
    stages = sum(t for a, t in zip(names, times) if a.startswith("Encode: "))
    assert stages <= call * 1.05 + 0.05
    for i in range(3):
        assert np.array_equal(to_host(outs[i]), to_host(streams[i]))
    cc.resetTimings(inst)
    assert cc.getTimings(inst) == ([], [])
    # level HIGH adds the kernels of the calls
    cc.setTimingDetail(inst, cc.TIMING_DETAIL_HIGH)
    assert cc.encodeRLHuff(inst, cc.BitStream(), streams, n)
    torch.cuda.synchronize()
    names, times = cc.getTimings(inst)
    assert "Encode RLHuff" in names and any(a.startswith("Kernel: ") for a in names)
    cc.resetTimings(inst)
    cc.setTimingDetail(inst, cc.TIMING_DETAIL_NONE)
    assert cc.encodeRLHuff(inst, cc.BitStream(), streams, n)
    assert cc.getTimings(inst) == ([], [])
    inst.close()


def test_timing_kernels_per_instance(cc, torch_gpu):
    """Two instances at TIMING_DETAIL_HIGH keep their own kernel entries: one instance's calls do
    not show up in the other's timings, lowering or destroying one leaves the other's kernel timing
    on, resetting one leaves the other's entries, and the process-wide profiler is a separate vote."""
    torch = torch_gpu
    log("timing: two instances")
    rng = np.random.default_rng(23)
    n = 1 << 13
    streams = [to_dev(torch, (rng.integers(0, 5, n) * (rng.random(n) < 0.4)).astype(np.uint16)) for _ in range(2)]
    a = cc.createInstance(0, 2, n, 128, 14)
    b = cc.createInstance(0, 2, n, 128, 14)
    cc.setTimingDetail(a, cc.TIMING_DETAIL_HIGH)
    cc.setTimingDetail(b, cc.TIMING_DETAIL_HIGH)

    def kernels(inst):
        names, _ = cc.getTimings(inst)
        return [x for x in names if x.startswith("Kernel: ")]

    assert cc.encodeRLHuff(a, cc.BitStream(), streams, n)
    torch.cuda.synchronize()
    assert kernels(a) and not kernels(b)          # b made no launches
    assert cc.encodeHuff(b, cc.BitStream(), streams, n)
    torch.cuda.synchronize()
    kb = kernels(b)
    assert kb
    cc.setTimingDetail(a, cc.TIMING_DETAIL_NONE)  # a leaves level 3: b keeps timing kernels
    cc.resetTimings(b)
    assert kernels(b) == []
    assert cc.encodeHuff(b, cc.BitStream(), streams, n)
    torch.cuda.synchronize()
    assert kernels(b)
    cc.setTimingDetail(a, cc.TIMING_DETAIL_HIGH)
    cc.resetTimings(a)                            # a's reset leaves b's entries
    assert kernels(b)
    lib = cc._capi.load()
    lib.cc_profile_reset()
    lib.cc_profile_enable(1)                      # the process-wide switch is its own vote
    a.close()                                     # destroying a leaves b's timing and the switch on
    assert cc.encodeHuff(b, cc.BitStream(), streams, n)
    torch.cuda.synchronize()
    assert kernels(b)
    assert lib.cc_profile_collect() > 0
    lib.cc_profile_enable(0)
    lib.cc_profile_reset()
    cc.setTimingDetail(b, cc.TIMING_DETAIL_NONE)
    b.close()
