"""CPU restatement of the image-path and bit-plane utilities (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/.  The product package never imports this module.

Follows:
  * floatToByte / byteToFloat            src/cudaCompress/util/Quantize.cuh:15-23
  * floatToByte2D / byteToFloat2D         src/cudaCompress/util/Quantize.cu:19-43, QuantizeKernels.cui:11-35
  * quantizeDifference*2D                 src/cudaCompress/util/QuantizeKernels.cui:228-283
  * RGB <-> YCoCg                         src/cudaCompress/util/YCoCg.cu:23-47
  * getLSB / removeLSB / appendLSB / shiftElems   src/cudaCompress/util/Bits.cu:18-172
  * byte DWT variants                     src/cudaCompress/util/DWTFloat.cu:648-711 with the
                                          conversions of DWTCommon.cuh:16-34
  * lowpass-only transforms               src/cudaCompress/util/DWTFloat2DLowpassKernels.cui,
                                          DWTInt2DLowpassKernels.cui (the LL band of one 2D level)
"""
import numpy as np

from . import pyoracle as o


def byte_to_float(b):
    return np.asarray(b, dtype=np.uint8).astype(np.float32) - np.float32(128.0)


def float_to_byte(v):
    """uchar(min(max(v + 128.5f, 0), 255)): fp32 add, clamp, truncation toward zero."""
    t = np.asarray(v, dtype=np.float32) + np.float32(128.5)
    t = np.where(np.isnan(t), np.float32(0.0), t)  # fmaxf(NaN, 0) = 0
    return np.clip(t, 0.0, 255.0).astype(np.uint8)


def float_to_byte_2d(data, channel_count, channel, out=None):
    """Channel `channel` of an interleaved (y, x, channel_count) byte image from a (y, x) plane."""
    sy, sx = data.shape
    if out is None:
        out = np.zeros((sy, sx, channel_count), dtype=np.uint8)
    out[:, :, channel] = float_to_byte(data)
    return out


def byte_to_float_2d(img, channel):
    return byte_to_float(img[:, :, channel])


def quantize_difference(data, step, ref_plane, qtype=0):
    """Symbols (uint16) of data - byteToFloat(ref); and the roundtrip reconstruction."""
    diff = np.asarray(data, dtype=np.float32) - byte_to_float(ref_plane)
    sym = o.quantize(diff, step, qtype=qtype, dtype=np.uint16)
    rec = o.unquantize(sym, step, qtype=qtype) + byte_to_float(ref_plane)
    return sym, rec


def unquantize_difference(sym, step, ref_plane, qtype=0):
    return o.unquantize(np.asarray(sym, dtype=np.uint16), step, qtype=qtype) + byte_to_float(ref_plane)


def _cdiv(a, b):
    """C integer division (truncation toward zero) of int arrays."""
    q = np.abs(a) // b
    return np.where(a < 0, -q, q)


def rgb_to_ycocg(px):
    """px: (..., 3 or 4) uint8; the fourth channel is copied."""
    p = np.asarray(px, dtype=np.int32)
    r, g, b = p[..., 0], p[..., 1], p[..., 2]
    out = np.array(px, dtype=np.uint8, copy=True)
    out[..., 0] = _cdiv(r + 2 * g + b + 2, 4)
    out[..., 1] = _cdiv(r - b + 1, 2) + 127
    out[..., 2] = _cdiv(-r + 2 * g - b + 2, 4) + 127
    return out


def ycocg_to_rgb(px):
    p = np.asarray(px, dtype=np.int32)
    y, co, cg = p[..., 0], p[..., 1], p[..., 2]
    out = np.array(px, dtype=np.uint8, copy=True)
    out[..., 0] = np.clip(y + co - cg, 0, 255)
    out[..., 1] = np.clip(y + cg - 127, 0, 255)
    out[..., 2] = np.clip(y - co - cg + 254, 0, 255)
    return out


def get_lsb(data):
    """Word w bit i = LSB of element 32*w + i; ceil(n/32) words, missing elements read as 0."""
    d = np.asarray(data, dtype=np.int16).ravel()
    n = d.size
    bits = np.zeros(((n + 31) // 32) * 32, dtype=np.uint64)
    bits[:n] = (d.astype(np.int32) & 1).astype(np.uint64)
    bits = bits.reshape(-1, 32) << np.arange(32, dtype=np.uint64)
    return bits.sum(axis=1).astype(np.uint32)


def remove_lsb(data, shift=0):
    """v = int16(v + shift); bit = v & 1; v = (v + (v < 0 ? -1 : 0)) / 2 with C division."""
    v = (np.asarray(data, dtype=np.int16).astype(np.int32) + int(shift)).astype(np.int16).astype(np.int32)
    bits = get_lsb(v.astype(np.int16))
    out = _cdiv(v + np.where(v < 0, -1, 0), 2).astype(np.int16)
    return out, bits


def append_lsb(data, bits, shift=0):
    """v = int16(int16((2*v) | bit) + shift)."""
    d = np.asarray(data, dtype=np.int16).astype(np.int32).ravel()
    n = d.size
    e = np.arange(n)
    b = ((np.asarray(bits, dtype=np.uint32)[e >> 5] >> (e & 31).astype(np.uint32)) & 1).astype(np.int32)
    v = ((d * 2) | b).astype(np.int16).astype(np.int32)
    return (v + int(shift)).astype(np.int16).reshape(np.shape(data))


def shift_elems(data, shift):
    return (np.asarray(data, dtype=np.int16).astype(np.int32) + int(shift)).astype(np.int16)


def dwt97_forward_from_byte(img):
    """One 9/7 level of (byte - 128) on a (y, x) or (z, y, x) byte array."""
    return o.dwt97_forward(byte_to_float(img))


def dwt97_inverse_to_byte(coeffs):
    return float_to_byte(o.dwt97_inverse(coeffs))


def dwt97_lowpass_only_from_byte(img):
    """The LL band of one 2D 9/7 level of (byte - 128), shape (y/2, x/2)."""
    c = dwt97_forward_from_byte(img)
    return c[: img.shape[0] // 2, : img.shape[1] // 2]


def dwt53_lowpass_only(a):
    """The LL band of one 2D 5/3 level, shape (y/2, x/2)."""
    c = o.dwt53_forward(a)
    return c[: a.shape[0] // 2, : a.shape[1] // 2]
