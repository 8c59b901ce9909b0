"""ctypes bindings for the CPU oracle (liboracle.so) and the reference CPU coder (_ref/libref_cpu.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by the CPU-baseline
leg of bench.py.  The product package never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liboracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref_cpu.so")

DIALECT_GPU = 0
DIALECT_CPU = 1

_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class Params(ctypes.Structure):
    _fields_ = [("dialect", ctypes.c_int), ("log2", ctypes.c_int), ("cbs", ctypes.c_int), ("symbol_size", ctypes.c_int)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run `make -C oracle`)")
        _lib = ctypes.CDLL(ORACLE_SO)
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference CPU coder not built: {REF_SO}")
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _ptr_array(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _sym_size(dtype):
    return np.dtype(dtype).itemsize


# ---------------------------------------------------------------------------- entropy coder
def encode(streams, rl=True, dialect=DIALECT_GPU, log2=14, cbs=128, start_bit=0):
    """Encode a list of equal-length 1D symbol arrays (uint16 or uint32) into one word array.

    Returns (words uint32 array, total bit count).  Raises on coder error.
    """
    streams = [np.ascontiguousarray(s) for s in streams]
    ssz = _sym_size(streams[0].dtype)
    count = int(streams[0].size)
    p = Params(dialect, log2, cbs, ssz)
    out = _u32p()
    nw = ctypes.c_uint64()
    nb = ctypes.c_uint64()
    rc = lib().orc_encode(int(rl), ctypes.byref(p), _ptr_array(streams), len(streams), count, start_bit,
                          ctypes.byref(out), ctypes.byref(nw), ctypes.byref(nb))
    if rc != 0:
        raise CoderError(rc)
    words = np.ctypeslib.as_array(out, shape=(max(nw.value, 1),))[: nw.value].copy()
    lib().orc_free(out)
    return words, nb.value


def decode(words, nbits, nstreams, count, dtype=np.uint16, rl=True, dialect=DIALECT_GPU, log2=14, cbs=128, bitpos=0):
    """Decode nstreams streams of `count` symbols; returns (list of arrays, end bit position)."""
    words = np.ascontiguousarray(words, dtype=np.uint32)
    outs = [np.zeros(count, dtype=dtype) for _ in range(nstreams)]
    p = Params(dialect, log2, cbs, _sym_size(dtype))
    pos = ctypes.c_uint64(bitpos)
    rc = lib().orc_decode(int(rl), ctypes.byref(p), _ptr(words), ctypes.c_uint64(nbits), ctypes.byref(pos),
                          _ptr_array(outs), nstreams, count)
    if rc != 0:
        raise CoderError(rc)
    return outs, pos.value


class CoderError(RuntimeError):
    def __init__(self, code):
        super().__init__(f"oracle coder error {code}")
        self.code = code


def rle_encode(syms):
    syms = np.ascontiguousarray(syms)
    n = syms.size
    out_s = np.zeros(max(n, 1), dtype=syms.dtype)
    out_z = np.zeros(max(n, 1), dtype=syms.dtype)
    m = ctypes.c_uint32()
    lib().orc_rle_encode(_sym_size(syms.dtype), _ptr(syms), n, _ptr(out_s), _ptr(out_z), ctypes.byref(m))
    return out_s[: m.value].copy(), out_z[: m.value].copy()


def rle_decode(syms, zc, n):
    syms = np.ascontiguousarray(syms)
    zc = np.ascontiguousarray(zc, dtype=syms.dtype)
    out = np.zeros(n, dtype=syms.dtype)
    rc = lib().orc_rle_decode(_sym_size(syms.dtype), _ptr(syms), _ptr(zc), syms.size, _ptr(out), n)
    if rc != 0:
        raise CoderError(rc)
    return out


def histogram(syms, call_symbol_max, nbins):
    syms = np.ascontiguousarray(syms)
    h = np.zeros(nbins, dtype=np.uint32)
    lib().orc_histogram(_sym_size(syms.dtype), _ptr(syms), syms.size, call_symbol_max, _ptr(h), nbins)
    return h


def huff_design(hist, dialect=DIALECT_GPU):
    """Design a table from a histogram whose last bin is the stream's max symbol.

    Returns dict(lengths, codewords (by symbol), count_per_length, symbols (by codeword index)).
    """
    hist = np.ascontiguousarray(hist, dtype=np.uint32)
    smax = hist.size - 1
    lengths = np.zeros(hist.size, dtype=np.uint32)
    codes = np.zeros(hist.size, dtype=np.uint32)
    cpl = np.zeros(32, dtype=np.uint32)
    syms = np.zeros(hist.size, dtype=np.uint32)
    nl = ctypes.c_uint32()
    ns = ctypes.c_uint32()
    rc = lib().orc_huff_design(dialect, _ptr(hist), smax, _ptr(lengths), _ptr(codes), ctypes.byref(nl), _ptr(cpl),
                               _ptr(syms), ctypes.byref(ns))
    if rc != 0:
        raise CoderError(rc)
    return dict(lengths=lengths, codewords=codes, count_per_length=cpl[: nl.value].copy(), symbols=syms[: ns.value].copy())


def pack_inc16(values):
    values = np.ascontiguousarray(values, dtype=np.uint32)
    out = np.zeros(values.size, dtype=np.uint16)
    lib().orc_pack_inc16(_ptr(values), _ptr(out), values.size)
    return out


def unpack_inc16(incs):
    incs = np.ascontiguousarray(incs, dtype=np.uint16)
    out = np.zeros(incs.size, dtype=np.uint32)
    lib().orc_unpack_inc16(_ptr(incs), _ptr(out), incs.size)
    return out


# ---------------------------------------------------------------------------- transforms
def _dims(a):
    if a.ndim == 2:
        return a.shape[1], a.shape[0], 1
    return a.shape[2], a.shape[1], a.shape[0]


def dwt97_forward(a, dtype=np.float32):
    """One CDF 9/7 level on a C-ordered (z, y, x) or (y, x) array; returns a new array."""
    a = np.array(a, dtype=dtype, order="C", copy=True)
    sx, sy, sz = _dims(a)
    fn = lib().orc_dwt97_forward_f32 if dtype == np.float32 else lib().orc_dwt97_forward_f64
    fn(_ptr(a), sx, sy, sz)
    return a


def dwt97_inverse(a, dtype=np.float32):
    a = np.array(a, dtype=dtype, order="C", copy=True)
    sx, sy, sz = _dims(a)
    fn = lib().orc_dwt97_inverse_f32 if dtype == np.float32 else lib().orc_dwt97_inverse_f64
    fn(_ptr(a), sx, sy, sz)
    return a


def dwt53_forward(a):
    a = np.array(a, dtype=np.int16, order="C", copy=True)
    sx, sy, sz = _dims(a)
    lib().orc_dwt53_forward(_ptr(a), sx, sy, sz)
    return a


def dwt53_inverse(a):
    a = np.array(a, dtype=np.int16, order="C", copy=True)
    sx, sy, sz = _dims(a)
    lib().orc_dwt53_inverse(_ptr(a), sx, sy, sz)
    return a


def quantize(data, step, qtype=0, dtype=np.uint16):
    data = np.ascontiguousarray(data, dtype=np.float32)
    out = np.zeros(data.shape, dtype=dtype)
    lib().orc_quantize(_ptr(data), data.size, ctypes.c_float(step), qtype, _sym_size(dtype), _ptr(out))
    return out


def unquantize(symbols, step, qtype=0):
    symbols = np.ascontiguousarray(symbols)
    out = np.zeros(symbols.shape, dtype=np.float32)
    lib().orc_unquantize(_ptr(symbols), symbols.size, ctypes.c_float(step), qtype, _sym_size(symbols.dtype), _ptr(out))
    return out


def quantize_to_short(data, step, qtype=0):
    """quantizeToShort (src/cudaCompress/util/QuantizeKernels.cui:136-151): int16 (wrapped)."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    out = np.zeros(data.shape, dtype=np.int16)
    lib().orc_quantize_to_short(_ptr(data), data.size, ctypes.c_float(step), qtype, _ptr(out))
    return out


def unquantize_from_short(quant, step, qtype=0):
    """unquantizeFromShort (src/cudaCompress/util/QuantizeKernels.cui:154-172)."""
    quant = np.ascontiguousarray(quant, dtype=np.int16)
    out = np.zeros(quant.shape, dtype=np.float32)
    lib().orc_unquantize_from_short(_ptr(quant), quant.size, ctypes.c_float(step), qtype, _ptr(out))
    return out


def symbolize(data):
    data = np.ascontiguousarray(data, dtype=np.int16)
    out = np.zeros(data.shape, dtype=np.uint16)
    lib().orc_symbolize(_ptr(data), data.size, _ptr(out))
    return out


def unsymbolize(symbols):
    symbols = np.ascontiguousarray(symbols, dtype=np.uint16)
    out = np.zeros(symbols.shape, dtype=np.int16)
    lib().orc_unsymbolize(_ptr(symbols), symbols.size, _ptr(out))
    return out


# ---------------------------------------------------------------------------- reference coder
def ref_encode(streams, rl=True, cbs=128):
    streams = [np.ascontiguousarray(s, dtype=np.uint16) for s in streams]
    out = _u32p()
    nw = ctypes.c_uint64()
    nb = ctypes.c_uint64()
    rc = ref().ref_encode(int(rl), cbs, _ptr_array(streams), len(streams), int(streams[0].size), ctypes.byref(out),
                          ctypes.byref(nw), ctypes.byref(nb))
    if rc != 0:
        raise RuntimeError(f"reference encode failed: {rc}")
    words = np.ctypeslib.as_array(out, shape=(max(nw.value, 1),))[: nw.value].copy()
    ref().ref_free(out)
    return words, nb.value


def ref_decode(words, nbits, nstreams, count, rl=True, cbs=128):
    words = np.ascontiguousarray(words, dtype=np.uint32)
    outs = [np.zeros(count, dtype=np.uint16) for _ in range(nstreams)]
    pos = ctypes.c_uint64()
    rc = ref().ref_decode(int(rl), cbs, _ptr(words), ctypes.c_uint64(nbits), _ptr_array(outs), nstreams, count,
                          ctypes.byref(pos))
    if rc != 0:
        raise RuntimeError(f"reference decode failed: {rc}")
    return outs, pos.value


def ref_design(syms):
    syms = np.ascontiguousarray(syms, dtype=np.uint16)
    cpl = np.zeros(32, dtype=np.uint32)
    out = np.zeros(65536, dtype=np.uint32)
    nl = ctypes.c_uint32()
    ns = ctypes.c_uint32()
    ref().ref_design(_ptr(syms), syms.size, ctypes.byref(nl), _ptr(cpl), _ptr(out), ctypes.byref(ns))
    return dict(count_per_length=cpl[: nl.value].copy(), symbols=out[: ns.value].copy())


def ref_rle(syms):
    syms = np.ascontiguousarray(syms, dtype=np.uint16)
    s = np.zeros(max(syms.size, 1), dtype=np.uint16)
    z = np.zeros(max(syms.size, 1), dtype=np.uint16)
    m = ctypes.c_uint32()
    ref().ref_rle(_ptr(syms), syms.size, _ptr(s), _ptr(z), ctypes.byref(m))
    return s[: m.value].copy(), z[: m.value].copy()
