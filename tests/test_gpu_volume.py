"""GPU parity of the volume codecs (compressVolumeFloat, compressVolumeFloatQuantFirst, their
multi-channel decoders and
compressVolumeLossless callers of
src/examples/CompressVolume.cpp) against the oracle composition in oracle/pipeline.py."""
import numpy as np
import pytest

from cudacompress_amd import synthetic as S
from oracle import pipeline as P
from tests.gpu_util import log, to_dev, to_host

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 2), ((64, 48, 32), 2), ((64, 64, 64), 3)],
                         ids=["32^3L2", "64x48x32L2", "64^3L3"])
@pytest.mark.parametrize("rle_lvl0", [False, True], ids=["rle_all", "rle_lvl0"])
def test_volume_float(cc, torch_gpu, shape, levels, rle_lvl0):
    torch = torch_gpu
    log("volume float", shape, levels, rle_lvl0)
    vol = S.gaussian_field(shape, 1)
    q0 = 0.01
    inst = cc.createInstance(0, 7, int(np.prod(shape)) // 8, 128, 14)
    src = to_dev(torch, vol)
    words = cc.compressVolumeFloat(inst, src, shape[2], shape[1], shape[0], levels, q0, rle_lvl0)
    want = P.compress_volume_float(vol, levels, q0, rle_only_lvl0=rle_lvl0)
    assert words is not None
    assert words.size == want.size and np.array_equal(words, want)
    out = torch.full(shape, 123.0, dtype=torch.float32, device="cuda")
    cc.decompressVolumeFloat(inst, out, shape[2], shape[1], shape[0], levels, words, q0, rle_lvl0)
    torch.cuda.synchronize()
    rec = to_host(out)
    ref = P.decompress_volume_float(want, shape, levels, q0, rle_only_lvl0=rle_lvl0)
    err = float(np.abs(rec - ref).max())
    log(f"   words={words.size} psnr={S.psnr(vol, rec):.2f} max|gpu-oracle|={err:.3e}")
    assert err <= 1e-6 * max(1.0, float(np.abs(ref).max()))
    inst.close()


@pytest.mark.parametrize("shape,levels", [((256, 256), 3), ((128, 64), 2), ((32, 32, 32), 2)],
                         ids=["2d256L3", "2d128x64L2", "3d32L2"])
def test_volume_lossless(cc, torch_gpu, shape, levels):
    torch = torch_gpu
    log("lossless", shape, levels)
    if len(shape) == 2:
        img = S.fbm_heightfield(shape[0], 0.8, 2)[:, : shape[1]].copy()
        sx, sy, sz = shape[1], shape[0], 1
    else:
        img = (S.gaussian_field(shape, 3) * 2000).astype(np.int16)
        sx, sy, sz = shape[2], shape[1], shape[0]
    inst = cc.createInstance(0, 7, int(np.prod(shape)) // 4, 128, 16)
    src = to_dev(torch, img)
    words = cc.compressVolumeLossless(inst, src, sx, sy, sz, levels)
    want = P.compress_lossless(img, levels)
    assert np.array_equal(words, want)
    out = torch.full(img.shape, -1, dtype=torch.int16, device="cuda")
    cc.decompressVolumeLossless(inst, out, sx, sy, sz, levels, words)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(out), img)
    inst.close()


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 2), ((64, 48, 32), 2), ((32, 64, 64), 3), ((16, 16, 32), 1)],
                         ids=["32^3L2", "64x48x32L2", "32x64x64L3", "16x16x32L1"])
@pytest.mark.parametrize("rle_lvl0", [False, True], ids=["rle_all", "rle_lvl0"])
def test_volume_float_quant_first(cc, torch_gpu, shape, levels, rle_lvl0):
    """compressVolumeFloatQuantFirst / decompressVolumeFloatQuantFirst against the oracle: words
    equal, reconstruction equal (exact: the codec loses only in the quantizer)."""
    torch = torch_gpu
    log("volume quant-first", shape, levels, rle_lvl0)
    vol = S.gaussian_field(shape, 5) * 40.0
    step = 0.05
    inst = cc.createInstance(0, 7, int(np.prod(shape)) // 8, 128, 14)
    src = to_dev(torch, vol)
    words = cc.compressVolumeFloatQuantFirst(inst, src, shape[2], shape[1], shape[0], levels, step, rle_lvl0)
    want = P.compress_volume_float_quant_first(vol, levels, step, rle_only_lvl0=rle_lvl0)
    assert words is not None
    assert words.size == want.size and np.array_equal(words, want)
    out = torch.full(shape, 123.0, dtype=torch.float32, device="cuda")
    cc.decompressVolumeFloatQuantFirst(inst, out, shape[2], shape[1], shape[0], levels, words, step, rle_lvl0)
    torch.cuda.synchronize()
    rec = to_host(out)
    ref = P.decompress_volume_float_quant_first(want, shape, levels, step, rle_only_lvl0=rle_lvl0)
    log(f"   words={words.size} psnr={S.psnr(vol, rec):.2f}")
    assert np.array_equal(rec, ref)
    inst.close()


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 2), ((64, 48, 32), 3)], ids=["32^3L2", "64x48x32L3"])
@pytest.mark.parametrize("rle_lvl0", [False, True], ids=["rle_all", "rle_lvl0"])
def test_volume_float_multichannel(cc, torch_gpu, shape, levels, rle_lvl0):
    """decompressVolumeFloatMultiChannel over three channels with their own steps: each channel
    equals its single-channel decompression exactly and the oracle within 1e-6 relative."""
    torch = torch_gpu
    log("volume float multichannel", shape, levels, rle_lvl0)
    steps = [0.01, 0.03, 0.02]
    vols = [S.gaussian_field(shape, 11 + c) * (1.0 + c) for c in range(3)]
    bits = [P.compress_volume_float(v, levels, q, rle_only_lvl0=rle_lvl0) for v, q in zip(vols, steps)]
    inst = cc.createInstance(0, 7 * len(vols), int(np.prod(shape)) // 8, 128, 14)
    outs = [torch.full(shape, 123.0, dtype=torch.float32, device="cuda") for _ in vols]
    cc.decompressVolumeFloatMultiChannel(inst, list(zip(outs, bits, steps)), shape[2], shape[1], shape[0], levels,
                                         rle_lvl0)
    one = torch.full(shape, -7.0, dtype=torch.float32, device="cuda")
    for c, (b, q) in enumerate(zip(bits, steps)):
        cc.decompressVolumeFloat(inst, one, shape[2], shape[1], shape[0], levels, b, q, rle_lvl0)
        torch.cuda.synchronize()
        rec = to_host(outs[c])
        assert np.array_equal(rec, to_host(one)), f"channel {c}"
        ref = P.decompress_volume_float(b, shape, levels, q, rle_only_lvl0=rle_lvl0)
        assert float(np.abs(rec - ref).max()) <= 1e-6 * max(1.0, float(np.abs(ref).max())), f"channel {c}"
    inst.close()


@pytest.mark.parametrize("shape,levels", [((32, 32, 32), 2), ((32, 64, 64), 3)], ids=["32^3L2", "32x64x64L3"])
@pytest.mark.parametrize("rle_lvl0", [False, True], ids=["rle_all", "rle_lvl0"])
def test_volume_float_quant_first_multichannel(cc, torch_gpu, shape, levels, rle_lvl0):
    """decompressVolumeFloatQuantFirstMultiChannel over three channels: each channel equals the
    oracle's decompression of its bitstream exactly."""
    torch = torch_gpu
    log("volume quant-first multichannel", shape, levels, rle_lvl0)
    steps = [0.05, 0.2, 0.1]
    vols = [S.gaussian_field(shape, 21 + c) * 40.0 for c in range(3)]
    bits = [P.compress_volume_float_quant_first(v, levels, q, rle_only_lvl0=rle_lvl0) for v, q in zip(vols, steps)]
    inst = cc.createInstance(0, 7 * len(vols), int(np.prod(shape)) // 8, 128, 14)
    outs = [torch.full(shape, 123.0, dtype=torch.float32, device="cuda") for _ in vols]
    cc.decompressVolumeFloatQuantFirstMultiChannel(inst, list(zip(outs, bits, steps)), shape[2], shape[1], shape[0],
                                                   levels, rle_lvl0)
    torch.cuda.synchronize()
    for c, (b, q) in enumerate(zip(bits, steps)):
        ref = P.decompress_volume_float_quant_first(b, shape, levels, q, rle_only_lvl0=rle_lvl0)
        assert np.array_equal(to_host(outs[c]), ref), f"channel {c}"
    inst.close()


@pytest.mark.parametrize("qtype", [0, 1], ids=["deadzone", "uniform"])
def test_quantize_to_short(cc, torch_gpu, qtype):
    """util::quantizeToShort / unquantizeFromShort with source / destination pitches, including
    values whose int wraps in int16."""
    from oracle import pyoracle as o
    torch = torch_gpu
    log("quantize to short", qtype)
    z, y, x = 10, 18, 50
    data = (S.gaussian_field((z, y, x + 6), 4) * 900.0).astype(np.float32)
    data[0, 0, :4] = [40000.0 * 0.05, -40001.0 * 0.05, 1640.0, -1640.0]
    step = 0.05
    src = to_dev(torch, data)
    q = torch.zeros((z, y, x), dtype=torch.int16, device="cuda")
    cc.util.quantizeToShort(q, src, x, y, z, step, x + 6, (x + 6) * y, qtype)
    torch.cuda.synchronize()
    want = o.quantize_to_short(np.ascontiguousarray(data[..., :x]), step, qtype=qtype)
    assert np.array_equal(to_host(q), want)
    rec = torch.zeros((z, y, x + 6), dtype=torch.float32, device="cuda")
    cc.util.unquantizeFromShort(rec, q, x, y, z, step, x + 6, (x + 6) * y, qtype)
    torch.cuda.synchronize()
    assert np.array_equal(to_host(rec)[..., :x], o.unquantize_from_short(want, step, qtype=qtype))
