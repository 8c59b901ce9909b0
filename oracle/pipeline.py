"""CPU composition of the volume codecs out of oracle primitives (test infrastructure only).

Restates the caller layouts of src/examples/CompressVolume.cpp (compressVolumeFloat :313-409,
decompressVolumeFloat :412-516, compressVolumeLossless :234-309, compressVolumeFloatQuantFirst
:651-757, decompressVolumeFloatQuantFirst :759-856) and, for 2D lossless fields, the band layout of
src/examples/CompressHeightfield.cpp:116-126, on top of oracle/liboracle.so.
Used by tests/ as the checker and by bench.py as the timed CPU baseline.
"""
import numpy as np

from . import pyoracle as o


def _bands3d(level_shape):
    z, y, x = level_shape
    bz, by, bx = z // 2, y // 2, x // 2
    out = []
    for b in range(1, 8):  # CompressVolume.cpp:375-379
        ox, oy, oz = b % 2, b // 2 % 2, b // 4
        out.append((slice(oz * bz, (oz + 1) * bz), slice(oy * by, (oy + 1) * by), slice(ox * bx, (ox + 1) * bx)))
    return out


def _bands2d(level_shape):
    y, x = level_shape
    by, bx = y // 2, x // 2
    out = []
    for b in range(1, 4):  # CompressHeightfield.cpp:119-121
        ox, oy = b % 2, b // 2
        out.append((slice(oy * by, (oy + 1) * by), slice(ox * bx, (ox + 1) * bx)))
    return out


def _step(q0, divisor):
    return np.float32(np.float32(q0) / np.float32(divisor))


def compress_volume_float(vol, levels, q0, log2=14, cbs=128, dtype=np.uint16, rle_only_lvl0=False, coef_out=None):
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    Z, Y, X = vol.shape
    coef = vol.copy()
    for lv in range(levels):
        s = (slice(0, Z >> lv), slice(0, Y >> lv), slice(0, X >> lv))
        coef[s] = o.dwt97_forward(coef[s])
    if coef_out is not None:
        coef_out.append(coef.copy())
    words = []
    f = 1 << levels
    lp = coef[: Z // f, : Y // f, : X // f]
    q = _step(q0, f // 2)
    w, _ = o.encode([o.quantize(lp.ravel(), q, dtype=dtype)], rl=False, log2=log2, cbs=cbs)
    words.append(w)
    for level in range(levels - 1, -1, -1):
        q = _step(q0, f // 2)
        region = coef[: Z // (f // 2), : Y // (f // 2), : X // (f // 2)]
        streams = [o.quantize(np.ascontiguousarray(region[b]).ravel(), q, dtype=dtype) for b in _bands3d(region.shape)]
        rl = not (level > 0 and rle_only_lvl0)
        w, _ = o.encode(streams, rl=rl, log2=log2, cbs=cbs)
        words.append(w)
        f //= 2
    return np.concatenate(words)


def decompress_volume_float(words, shape, levels, q0, log2=14, cbs=128, dtype=np.uint16, rle_only_lvl0=False):
    Z, Y, X = shape
    nbits = words.size * 32
    vol = np.zeros(shape, dtype=np.float32)
    f = 1 << levels
    (lp,), pos = o.decode(words, nbits, 1, (Z // f) * (Y // f) * (X // f), dtype=dtype, rl=False, log2=log2, cbs=cbs)
    vol[: Z // f, : Y // f, : X // f] = o.unquantize(lp, _step(q0, f // 2)).reshape(Z // f, Y // f, X // f)
    for level in range(levels - 1, -1, -1):
        count = (Z // f) * (Y // f) * (X // f)
        rl = not (level > 0 and rle_only_lvl0)
        streams, pos = o.decode(words, nbits, 7, count, dtype=dtype, rl=rl, log2=log2, cbs=cbs, bitpos=pos)
        f //= 2
        q = _step(q0, f)
        region = vol[: Z // f, : Y // f, : X // f]
        for b, s in zip(_bands3d(region.shape), streams):
            region[b] = o.unquantize(s, q).reshape(region[b].shape)
        vol[: Z // f, : Y // f, : X // f] = o.dwt97_inverse(region)
    return vol


def compress_lossless(img, levels, log2=16, cbs=128):
    img = np.ascontiguousarray(img, dtype=np.int16)
    is3d = img.ndim == 3
    bands = _bands3d if is3d else _bands2d
    cur = img
    level_words = []
    for level in range(levels):
        coef = o.dwt53_forward(cur)
        streams = [o.symbolize(np.ascontiguousarray(coef[b]).ravel()) for b in bands(coef.shape)]
        w, _ = o.encode(streams, rl=True, log2=log2, cbs=cbs)
        level_words.append(w)
        cur = np.ascontiguousarray(coef[tuple(slice(0, n // 2) for n in coef.shape)])
    w, _ = o.encode([o.symbolize(cur.ravel())], rl=False, log2=log2, cbs=cbs)
    return np.concatenate([w] + level_words[::-1])


def decompress_lossless(words, shape, levels, log2=16, cbs=128):
    is3d = len(shape) == 3
    bands = _bands3d if is3d else _bands2d
    nbits = words.size * 32
    lshape = tuple(n >> levels for n in shape)
    (ll,), pos = o.decode(words, nbits, 1, int(np.prod(lshape)), rl=False, log2=log2, cbs=cbs)
    cur = o.unsymbolize(ll).reshape(lshape)
    nb = 7 if is3d else 3
    for level in range(levels - 1, -1, -1):
        cshape = tuple(n >> level for n in shape)
        coef = np.zeros(cshape, dtype=np.int16)
        coef[tuple(slice(0, n // 2) for n in cshape)] = cur
        count = int(np.prod([n // 2 for n in cshape]))
        streams, pos = o.decode(words, nbits, nb, count, rl=True, log2=log2, cbs=cbs, bitpos=pos)
        for b, s in zip(bands(cshape), streams):
            coef[b] = o.unsymbolize(s).reshape(coef[b].shape)
        cur = o.dwt53_inverse(coef)
    return cur


def compress_volume_float_quant_first(vol, levels, step, log2=14, cbs=128, rle_only_lvl0=False):
    """compressVolumeFloatQuantFirst (CompressVolume.cpp:651-757): uniform quantization to int16,
    `levels` 5/3 levels, the lowpass band symbolized and Huffman coded, then per level (coarsest
    first) the 7 symbolized highpass bands RL+Huffman coded (Huffman only above level 0 with
    rle_only_lvl0)."""
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    Z, Y, X = vol.shape
    coef = o.quantize_to_short(vol, step, qtype=1)
    for lv in range(levels):
        s = (slice(0, Z >> lv), slice(0, Y >> lv), slice(0, X >> lv))
        coef[s] = o.dwt53_forward(np.ascontiguousarray(coef[s]))
    f = 1 << levels
    lp = coef[: Z // f, : Y // f, : X // f]
    w, _ = o.encode([o.symbolize(np.ascontiguousarray(lp).ravel())], rl=False, log2=log2, cbs=cbs)
    words = [w]
    for level in range(levels - 1, -1, -1):
        region = coef[: Z // (f // 2), : Y // (f // 2), : X // (f // 2)]
        streams = [o.symbolize(np.ascontiguousarray(region[b]).ravel()) for b in _bands3d(region.shape)]
        w, _ = o.encode(streams, rl=not (level > 0 and rle_only_lvl0), log2=log2, cbs=cbs)
        words.append(w)
        f //= 2
    return np.concatenate(words)


def decompress_volume_float_quant_first(words, shape, levels, step, log2=14, cbs=128, rle_only_lvl0=False):
    """decompressVolumeFloatQuantFirst (CompressVolume.cpp:759-856): the inverse of the above, then
    unquantizeFromShort (uniform)."""
    Z, Y, X = shape
    nbits = words.size * 32
    coef = np.zeros(shape, dtype=np.int16)
    f = 1 << levels
    (lp,), pos = o.decode(words, nbits, 1, (Z // f) * (Y // f) * (X // f), rl=False, log2=log2, cbs=cbs)
    coef[: Z // f, : Y // f, : X // f] = o.unsymbolize(lp).reshape(Z // f, Y // f, X // f)
    for level in range(levels - 1, -1, -1):
        count = (Z // f) * (Y // f) * (X // f)
        rl = not (level > 0 and rle_only_lvl0)
        streams, pos = o.decode(words, nbits, 7, count, rl=rl, log2=log2, cbs=cbs, bitpos=pos)
        f //= 2
        region = coef[: Z // f, : Y // f, : X // f]
        for b, st in zip(_bands3d(region.shape), streams):
            region[b] = o.unsymbolize(st).reshape(region[b].shape)
        coef[: Z // f, : Y // f, : X // f] = o.dwt53_inverse(np.ascontiguousarray(region))
    return o.unquantize_from_short(coef, step, qtype=1)


# ---------------------------------------------------------------------------- precision reports
def dwt97_levels(vol, levels, dtype=np.float32):
    """`levels` levels of the 9/7 analysis in `dtype` with the in-place layout of
    compressVolumeFloat (CompressVolume.cpp:342-352): level l transforms the leading
    (Z, Y, X) >> l corner."""
    coef = np.array(vol, dtype=dtype, order="C", copy=True)
    for lv in range(levels):
        s = tuple(slice(0, n >> lv) for n in coef.shape)
        coef[s] = o.dwt97_forward(coef[s], dtype=dtype)
    return coef


def band_steps(shape, levels, q0):
    """(slices, step) of every quantized band in bitstream order: the lowpass band at
    q0 / 2^(levels-1), then per level (coarsest first) the 7 highpass bands at q0 / 2^level
    (CompressVolume.cpp:355-386)."""
    Z, Y, X = shape
    f = 1 << levels
    out = [((slice(0, Z // f), slice(0, Y // f), slice(0, X // f)), float(_step(q0, f // 2)))]
    for level in range(levels - 1, -1, -1):
        region = (Z // (f // 2), Y // (f // 2), X // (f // 2))
        q = float(_step(q0, f // 2))
        out += [(b, q) for b in _bands3d(region)]
        f //= 2
    return out


def precision_report(vol, levels, q0, boundary_tol=1e-5):
    """FP32-vs-FP64 error of the multi-level 9/7 analysis and the quantizer's sensitivity to it.

    dwt_max_abs / dwt_rel_l2: max |c32 - c64| and ||c32 - c64|| / ||c64|| over all coefficients;
    near_boundary: coefficients whose FP64 value lies within boundary_tol (in units of the band's
    step) of a deadzone bin edge |c| / step = k >= 1 (Quantize.cuh:76-120 truncates);
    symbols_differ_fp64: quantized symbols of the FP32 coefficients that differ from those of the
    FP64 coefficients quantized the same way."""
    c32 = dwt97_levels(vol, levels, np.float32)
    c64 = dwt97_levels(vol, levels, np.float64)
    d = c32.astype(np.float64) - c64
    near = 0
    differ = 0
    for sl, q in band_steps(vol.shape, levels, q0):
        a64 = np.abs(c64[sl]).ravel() / np.float64(np.float32(q))
        k = np.rint(a64)
        near += int(np.count_nonzero((k >= 1) & (np.abs(a64 - k) < boundary_tol)))
        s32 = o.quantize(np.ascontiguousarray(c32[sl]).ravel(), q, dtype=np.uint32)
        t64 = np.trunc(c64[sl].ravel() / np.float64(np.float32(q))).astype(np.int64)
        s64 = np.where(t64 >= 0, 2 * t64, -2 * t64 - 1).astype(np.uint32)
        differ += int(np.count_nonzero(s32 != s64))
    return {"dwt_max_abs": float(np.abs(d).max()),
            "dwt_rel_l2": float(np.sqrt((d * d).sum() / max((c64 * c64).sum(), 1e-300))),
            "near_boundary": near, "symbols_differ_fp64": differ, "coefficients": int(c64.size)}


def volume_float_calls(vol, levels, q0, dtype=np.uint16):
    """The coder calls of compressVolumeFloat (CompressVolume.cpp:355-400) as (rl, streams): the
    lowpass band Huffman coded, then per level (coarsest first) the 7 highpass bands run-length +
    Huffman coded.  The entropy half of the pipeline, for timing coders on the same symbols."""
    coef = dwt97_levels(vol, levels, np.float32)
    calls = []
    bs = band_steps(vol.shape, levels, q0)
    calls.append((False, [o.quantize(np.ascontiguousarray(coef[bs[0][0]]).ravel(), bs[0][1], dtype=dtype)]))
    for lv in range(levels):
        part = bs[1 + 7 * lv: 8 + 7 * lv]
        calls.append((True, [o.quantize(np.ascontiguousarray(coef[sl]).ravel(), q, dtype=dtype) for sl, q in part]))
    return calls
