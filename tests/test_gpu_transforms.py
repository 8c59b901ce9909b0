"""GPU parity of the transforms (CDF 9/7, CDF 5/3, quantizer) against the oracle, through the C ABI.

Bars: integer outputs bit-exact; fp32 DWT within 1e-6 relative to max|x| (the kernels evaluate the
same fmaf chains as the oracle, so in practice they agree exactly -- the exact-match fraction is
asserted too); quantized symbols bit-exact except coefficients within 1e-5 of a bin edge, which
are counted and must be none here since both sides quantize the same fp32 value.
"""
import numpy as np
import pytest

from cudacompress_amd import synthetic as S
from oracle import pyoracle as o
from tests.gpu_util import log, to_dev, to_host

pytestmark = pytest.mark.gpu

F32_TOL = 1e-6


def assert_close_f32(got, want):
    scale = max(1.0, float(np.abs(want).max()))
    err = float(np.abs(got.astype(np.float64) - want.astype(np.float64)).max())
    exact = float(np.mean(got == want))
    log(f"   max|err|={err:.3e} scale={scale:.3e} exact={exact:.4f}")
    assert err <= F32_TOL * scale
    assert exact > 0.99


DWT3 = [(16, 16, 16), (48, 32, 16), (64, 64, 64), (96, 16, 128), (128, 128, 32), (8, 8, 8)]
DWT2 = [(64, 48), (512, 256), (4096, 64), (6, 10)]


@pytest.mark.parametrize("shape", DWT3, ids=[f"{z}x{y}x{x}" for z, y, x in DWT3])
def test_dwt97_3d(cc, torch_gpu, shape):
    torch = torch_gpu
    log("dwt97 3d", shape)
    z, y, x = shape
    vol = S.gaussian_field(shape, 1)
    src = to_dev(torch, vol)
    dst = torch.empty_like(src)
    b1 = torch.empty_like(src)
    b2 = torch.empty_like(src)
    cc.util.dwtFloat3DForward(dst, b2, b1, src, x, y, z)
    torch.cuda.synchronize()
    want = o.dwt97_forward(vol)
    assert_close_f32(to_host(dst), want)
    # inverse of the oracle's coefficients
    csrc = to_dev(torch, want)
    rec = torch.empty_like(src)
    cc.util.dwtFloat3DInverse(rec, b2, b1, csrc, x, y, z)
    torch.cuda.synchronize()
    assert_close_f32(to_host(rec), o.dwt97_inverse(want))


@pytest.mark.parametrize("shape", DWT2, ids=[f"{y}x{x}" for y, x in DWT2])
def test_dwt97_2d(cc, torch_gpu, shape):
    torch = torch_gpu
    log("dwt97 2d", shape)
    y, x = shape
    img = S.gaussian_field(shape, 2)
    src = to_dev(torch, img)
    dst = torch.empty_like(src)
    b1 = torch.empty_like(src)
    cc.util.dwtFloat2DForward(dst, b1, src, x, y)
    torch.cuda.synchronize()
    want = o.dwt97_forward(img)
    assert_close_f32(to_host(dst), want)
    rec = torch.empty_like(src)
    cc.util.dwtFloat2DInverse(rec, b1, to_dev(torch, want), x, y)
    torch.cuda.synchronize()
    assert_close_f32(to_host(rec), o.dwt97_inverse(want))


def test_dwt97_pitched_multichannel(cc, torch_gpu):
    """srcChannelCount / dstChannelCount and explicit pitches (DWTFloat.cu:474-477)."""
    torch = torch_gpu
    log("dwt97 pitched 3 channels")
    z, y, x, ch = 16, 24, 32, 3
    data = S.gaussian_field((z, y, x * ch), 3).reshape(z, y, x, ch)
    src = to_dev(torch, data)
    pitch_x = x + 16
    dst = torch.zeros((z, y, pitch_x), dtype=torch.float32, device="cuda")
    b1 = torch.zeros_like(dst)
    b2 = torch.zeros_like(dst)
    src_ch1 = src.reshape(-1)[1:]  # channel 1 of the interleaved source
    cc.util.dwtFloat3DForward(dst, b2, b1, src_ch1, x, y, z, ch, pitch_x, pitch_x * y, x * ch, x * ch * y)
    torch.cuda.synchronize()
    want = o.dwt97_forward(np.ascontiguousarray(data[..., 1]))
    assert_close_f32(to_host(dst)[..., :x], want)
    # inverse back into channel 2 of an interleaved destination
    out = torch.zeros((z, y, x, ch), dtype=torch.float32, device="cuda")
    cc.util.dwtFloat3DInverse(out.reshape(-1)[2:], b2, b1, dst, x, y, z, ch, x * ch, x * ch * y, pitch_x, pitch_x * y)
    torch.cuda.synchronize()
    assert_close_f32(to_host(out)[..., 2], o.dwt97_inverse(want))
    assert float(np.abs(to_host(out)[..., 0]).max()) == 0.0


@pytest.mark.parametrize("shape", [(32, 32, 32), (16, 48, 64), (64, 32)], ids=["32^3", "16x48x64", "2d64x32"])
@pytest.mark.parametrize("symsize", [2, 4])
def test_dwt97_inverse_from_symbols(cc, torch_gpu, shape, symsize):
    torch = torch_gpu
    log("from symbols", shape, symsize)
    dt = np.uint16 if symsize == 2 else np.uint32
    coef = S.gaussian_field(shape, 4)
    q = np.float32(0.01)
    is3d = len(shape) == 3
    if is3d:
        z, y, x = shape
        bands = [(slice(bz * (z // 2), (bz + 1) * (z // 2)), slice(by * (y // 2), (by + 1) * (y // 2)),
                  slice(bx * (x // 2), (bx + 1) * (x // 2))) for b in range(1, 8)
                 for bx, by, bz in [(b % 2, b // 2 % 2, b // 4)]]
        lp_sl = (slice(0, z // 2), slice(0, y // 2), slice(0, x // 2))
    else:
        y, x = shape
        bands = [(slice(by * (y // 2), (by + 1) * (y // 2)), slice(bx * (x // 2), (bx + 1) * (x // 2)))
                 for b in range(1, 4) for bx, by in [(b % 2, b // 2)]]
        lp_sl = (slice(0, y // 2), slice(0, x // 2))
    syms = [o.quantize(np.ascontiguousarray(coef[b]).ravel(), q, dtype=dt) for b in bands]
    # oracle: dequantize into the octants, keep the lowpass, inverse DWT
    full = coef.copy()
    for b, s in zip(bands, syms):
        full[b] = o.unquantize(s, q).reshape(full[b].shape)
    want = o.dwt97_inverse(full)
    # device: lowpass lives inside a full-size pitched array (as in decompressVolumeFloat)
    lp = torch.zeros(shape, dtype=torch.float32, device="cuda")
    lp[lp_sl] = to_dev(torch, np.ascontiguousarray(coef[lp_sl]))
    dsyms = [to_dev(torch, s) for s in syms]
    ptrs = torch.tensor([t.data_ptr() for t in dsyms], dtype=torch.int64, device="cuda")
    out = torch.empty_like(lp)
    b1, b2 = torch.empty_like(lp), torch.empty_like(lp)
    if is3d:
        cc.util.dwtFloat3DInverseFromSymbols(out, b2, b1, lp, ptrs, float(q), x, y, z, 1, x, x * y, x, x * y,
                                             symbolSize=symsize)
    else:
        cc.util.dwtFloat2DInverseFromSymbols(out, b1, lp, ptrs, float(q), x, y, 1, x, x, symbolSize=symsize)
    torch.cuda.synchronize()
    assert_close_f32(to_host(out), want)


INT3 = [(16, 16, 16), (32, 48, 64), (8, 128, 96)]
INT2 = [(64, 48), (1024, 512), (4096, 128), (6, 8)]


def _extreme(shape, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(-32768, 32768, size=shape).astype(np.int16)
    a.flat[::7] = 32767
    a.flat[::11] = -32768
    return a


@pytest.mark.parametrize("shape", INT3 + INT2, ids=["x".join(map(str, s)) for s in INT3 + INT2])
def test_dwt53_bit_exact(cc, torch_gpu, shape):
    torch = torch_gpu
    log("dwt53", shape)
    for seed, data in ((0, (S.gaussian_field(shape, 5) * 3000).astype(np.int16)), (1, _extreme(shape, 6))):
        if len(shape) == 3:
            z, y, x = shape
        else:
            (y, x), z = shape, 1
        src = to_dev(torch, data)
        dst = torch.empty_like(src)
        buf = torch.empty_like(src)
        cc.util.dwtIntForward(dst, buf, src, x, y, z)
        torch.cuda.synchronize()
        want = o.dwt53_forward(data)
        got = to_host(dst)
        assert np.array_equal(got, want), f"seed {seed}: {np.count_nonzero(got != want)} mismatches"
        rec = torch.empty_like(src)
        cc.util.dwtIntInverse(rec, buf, dst, x, y, z)
        torch.cuda.synchronize()
        assert np.array_equal(to_host(rec), data)


def test_dwt53_pitched(cc, torch_gpu):
    torch = torch_gpu
    log("dwt53 pitched")
    z, y, x = 8, 32, 64
    data = (S.gaussian_field((z, y, x), 8) * 1000).astype(np.int16)
    big = np.zeros((z, y + 3, x + 10), np.int16)
    big[:, :y, :x] = data
    src = to_dev(torch, big)
    dst = torch.zeros((z, y + 1, x + 6), dtype=torch.int16, device="cuda")
    buf = torch.zeros_like(dst)
    cc.util.dwtIntForward(dst, buf, src, x, y, z, x + 6, (x + 6) * (y + 1), x + 10, (x + 10) * (y + 3))
    torch.cuda.synchronize()
    assert np.array_equal(to_host(dst)[:, :y, :x], o.dwt53_forward(data))


@pytest.mark.parametrize("symsize", [2, 4])
@pytest.mark.parametrize("qtype", [0, 1], ids=["deadzone", "uniform"])
def test_quantize(cc, torch_gpu, symsize, qtype):
    torch = torch_gpu
    log("quantize", symsize, qtype)
    dt = np.uint16 if symsize == 2 else np.uint32
    z, y, x = 12, 20, 70
    data = S.gaussian_field((z, y, x + 6), 9) * 3.0
    src = to_dev(torch, data)
    step = np.float32(0.013)
    syms = torch.zeros(z * y * x, dtype=torch.int16 if symsize == 2 else torch.int32, device="cuda")
    cc.util.quantizeToSymbols(syms, src, x, y, z, float(step), x + 6, (x + 6) * y, qtype)
    torch.cuda.synchronize()
    dense = np.ascontiguousarray(data[..., :x])
    want = o.quantize(dense.ravel(), step, qtype=qtype, dtype=dt)
    got = to_host(syms, dt)
    # boundary cases: coefficients within 1e-5 (relative to the step) of a bin edge
    t = dense.ravel().astype(np.float64) / float(step) + (0.5 * np.sign(dense.ravel()) if qtype else 0)
    near = np.abs(t - np.rint(t)) < 1e-5
    mism = got != want
    log(f"   mismatches={int(mism.sum())} near-boundary={int(near.sum())}")
    assert not np.any(mism & ~near)
    assert int(mism.sum()) == 0
    rec = torch.zeros((z, y, x + 6), dtype=torch.float32, device="cuda")
    cc.util.unquantizeFromSymbols(rec, syms, x, y, z, float(step), x + 6, (x + 6) * y, qtype)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(rec)[..., :x], o.unquantize(want, step, qtype=qtype).reshape(z, y, x))
    # roundtrip variant writes the dequantized value back in place
    src2 = to_dev(torch, data)
    syms2 = torch.zeros_like(syms)
    cc.util.quantizeToSymbolsRoundtrip(syms2, src2, x, y, z, float(step), x + 6, (x + 6) * y, qtype)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(syms2, dt), want)
    assert np.array_equal(to_host(src2)[..., :x], to_host(rec)[..., :x])


def test_symbolize_roundtrip_extremes(cc, torch_gpu):
    torch = torch_gpu
    log("symbolize")
    z, y, x = 3, 17, 40
    data = _extreme((z, y, x), 10)
    src = to_dev(torch, data)
    sym = torch.zeros(z * y * x, dtype=torch.int16, device="cuda")
    cc.util.symbolize(sym, src, x, y, z)
    torch.cuda.synchronize()
    want = o.symbolize(data.ravel())
    assert np.array_equal(to_host(sym, np.uint16), want)
    back = torch.zeros_like(src)
    cc.util.unsymbolize(back, sym, x, y, z)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(back), data)


def test_max_abs(cc, torch_gpu):
    torch = torch_gpu
    inst = cc.createInstance(0, 1, 1024)
    data = S.gaussian_field((64, 64, 64), 12)
    src = to_dev(torch, data)
    out = torch.zeros(1, dtype=torch.float32, device="cuda")
    cc.util.getMaxAbs(inst, src, data.size, out)
    torch.cuda.synchronize()
    assert float(out.item()) == float(np.abs(data).max())
    inst.close()


@pytest.mark.parametrize("shape", [(64, 48, 32), (16, 16, 128), (96, 64)], ids=["64x48x32", "16x16x128", "2d96x64"])
def test_dwt97_forward_f64(cc, torch_gpu, shape):
    """The FP64 twin of the 9/7 analysis (cc_dwt_double_forward) against the oracle's FP64
    transform: both sum the decimal coefficients with fused multiply-adds in filter order, so
    they agree to rounding (1e-12 relative); pitched levels use the leading corner of a volume."""
    torch = torch_gpu
    log("dwt97 f64", shape)
    vol = S.gaussian_field(shape, 7).astype(np.float64)
    dims = (shape[-1], shape[-2], shape[0] if len(shape) == 3 else 1)
    src = torch.from_numpy(vol).cuda()
    dst = torch.zeros_like(src)
    buf = torch.zeros_like(src)
    cc.util.dwtDouble3DForward(dst, buf, src, *dims)
    torch.cuda.synchronize()
    want = o.dwt97_forward(vol, dtype=np.float64)
    err = float(np.abs(to_host(dst) - want).max())
    log(f"   max|err|={err:.3e}")
    assert err <= 1e-12 * max(1.0, float(np.abs(want).max()))
    # second level on the leading corner, pitched like the float transform's levels
    half = tuple(n // 2 for n in shape)
    corner = np.ascontiguousarray(want[tuple(slice(0, h) for h in half)])
    row = shape[-1]
    slab = shape[-1] * shape[-2] if len(shape) == 3 else 0
    d2 = torch.zeros_like(src)
    cc.util.dwtDouble3DForward(d2, buf, torch.from_numpy(want).cuda(), dims[0] // 2, dims[1] // 2,
                               max(1, dims[2] // 2) if len(shape) == 3 else 1, row, slab)
    torch.cuda.synchronize()
    got = to_host(d2)[tuple(slice(0, h) for h in half)]
    want2 = o.dwt97_forward(corner, dtype=np.float64)
    assert float(np.abs(got - want2).max()) <= 1e-12 * max(1.0, float(np.abs(want2).max()))
