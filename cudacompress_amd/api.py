"""Python mirror of the cudaCompress C++ API over the C ABI (include/cudacompress_amd.h).

Names, argument order and meaning follow the reference headers (include/cudaCompress/Instance.h,
Encode.h, BitStream.h, util/DWT.h, util/Quantize.h).  Device arrays are torch tensors (or raw
device addresses as ints); streams are torch.cuda.Stream objects, raw hipStream_t ints, or None
(the null stream).  Encode/decode return False on symbol overflow, as the reference does, and
raise CCError for every other failure.
"""
import ctypes

import numpy as np

from . import _capi

CC_OK = 0
CC_ERR_SYMBOL_OVERFLOW = 1
CC_ERR_OUT_OF_MEMORY = 2
CC_ERR_INVALID_ARG = 3
CC_ERR_DEVICE = 4
CC_ERR_CODEWORD_TOO_LONG = 5
CC_ERR_CORRUPT_STREAM = 6

DIALECT_GPU = 0
DIALECT_CPU = 1

QUANTIZE_DEADZONE = 0
QUANTIZE_UNIFORM = 1


class CCError(RuntimeError):
    def __init__(self, status, where=""):
        lib = _capi.load()
        msg = lib.cc_status_string(status).decode()
        detail = lib.cc_last_error().decode()
        super().__init__(f"{where}: {msg}" + (f" ({detail})" if detail else ""))
        self.status = status


def _check(status, where):
    if status != CC_OK:
        raise CCError(status, where)


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _stream(s):
    if s is None:
        return None
    if isinstance(s, int):
        return s
    return s.cuda_stream


def _elem_size(t):
    return t.element_size() if hasattr(t, "element_size") else None


# ------------------------------------------------------------------------------------ Instance
def tuning_to_dict(t):
    return {"designGrid": [int(g) for g in t.designGrid], "fusedXY": int(t.fusedXY), "xySlicesMax": int(t.xySlicesMax),
            "codingBlockSize": int(t.codingBlockSize), "laneDesignK": int(t.laneDesignK)}


def tuning_from_dict(d):
    t = _capi.Tuning()
    for i, g in enumerate(d["designGrid"]):
        t.designGrid[i] = int(g)
    t.fusedXY, t.xySlicesMax, t.codingBlockSize = int(d["fusedXY"]), int(d["xySlicesMax"]), int(d["codingBlockSize"])
    t.laneDesignK = int(d.get("laneDesignK", 0))
    return t


def loadTuning(path=None):
    """A tuning table parsed by the library (cc_load_tuning); None: the library's own table."""
    t = _capi.Tuning()
    _check(_capi.load().cc_load_tuning(None if path is None else str(path).encode(), ctypes.byref(t)), "cc_load_tuning")
    return tuning_to_dict(t)


def defaultTuning():
    t = _capi.Tuning()
    _check(_capi.load().cc_default_tuning(ctypes.byref(t)), "cc_default_tuning")
    return tuning_to_dict(t)


class Instance:
    """cudaCompress::Instance (include/cudaCompress/Instance.h:14-45)."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle)
        self._stream = None

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _capi.load().cc_destroy_instance(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # getInstance* (Instance.h:36-41)
    @property
    def cudaDevice(self):
        return _capi.load().cc_get_instance_device(self._h)

    @property
    def streamCountMax(self):
        return _capi.load().cc_get_instance_stream_count_max(self._h)

    @property
    def elemCountPerStreamMax(self):
        return _capi.load().cc_get_instance_elem_count_per_stream_max(self._h)

    @property
    def codingBlockSize(self):
        return _capi.load().cc_get_instance_coding_block_size(self._h)

    @property
    def log2HuffmanDistinctSymbolCountMax(self):
        return _capi.load().cc_get_instance_log2_huffman_distinct_symbol_count_max(self._h)

    @property
    def useLongSymbols(self):
        return bool(_capi.load().cc_get_instance_use_long_symbols(self._h))

    @property
    def symbolSize(self):
        return 4 if self.useLongSymbols else 2

    def setStream(self, stream):
        """setInstanceCudaStream (Instance.h:44)."""
        self._stream = stream
        _capi.load().cc_set_instance_stream(self._h, _stream(stream))

    def getStream(self):
        return _capi.load().cc_get_instance_stream(self._h)

    def setDialect(self, dialect):
        _check(_capi.load().cc_set_instance_dialect(self._h, dialect), "setDialect")

    def getTuning(self):
        """The kernel shapes brick batches and plans of this instance use (cc_get_instance_tuning)."""
        t = _capi.Tuning()
        _check(_capi.load().cc_get_instance_tuning(self.handle, ctypes.byref(t)), "cc_get_instance_tuning")
        return tuning_to_dict(t)

    def setTuning(self, tuning=None):
        """Sets the kernel shapes (a dict as getTuning returns; None: the library's tuning table)."""
        t = None if tuning is None else ctypes.byref(tuning_from_dict(tuning))
        _check(_capi.load().cc_set_instance_tuning(self.handle, t), "cc_set_instance_tuning")

    def setBrickGroups(self, groups):
        """Concurrent groups of a brick batch (0: chosen from the batch size)."""
        _check(_capi.load().cc_set_instance_brick_groups(self._h, groups), "setBrickGroups")


def createInstance(cudaDevice=-1, streamCountMax=1, elemCountPerStreamMax=1, codingBlockSize=0,
                   log2HuffmanDistinctSymbolCountMax=0):
    """createInstance (Instance.h:26-31).  Raises CCError instead of returning nullptr."""
    lib = _capi.load()
    h = ctypes.c_void_p()
    _check(lib.cc_create_instance(cudaDevice, streamCountMax, elemCountPerStreamMax, codingBlockSize,
                                  log2HuffmanDistinctSymbolCountMax, ctypes.byref(h)), "createInstance")
    return Instance(h.value)


def destroyInstance(inst):
    inst.close()


# ------------------------------------------------------------------------------------ BitStream
class BitStream:
    """Host bit stream with the semantics of include/cudaCompress/BitStream.h.

    Words are uint32, bits are written MSB-first; aligned reads/writes round the position up to
    a 32-bit word first and afterwards.
    """

    def __init__(self, words=None, bit_size=None):
        if words is None:
            self._w = np.zeros(0, dtype=np.uint32)
            self._size = 0
        else:
            self._w = np.ascontiguousarray(words, dtype=np.uint32).copy()
            self._size = int(self._w.size * 32 if bit_size is None else bit_size)
        self._pos = 0

    # -- queries
    def getBitPosition(self):
        return self._pos

    def setBitPosition(self, pos):
        if pos > self._size:
            self._pos = self._size
            return False
        self._pos = int(pos)
        return True

    def getBitSize(self):
        return self._size

    def getRaw(self):
        n = (self._size + 31) // 32
        return self._w[:n]

    def getRawSizeUInts(self):
        return (self._size + 31) // 32

    def words(self):
        return self.getRaw().copy()

    # -- writing
    def _grow(self, bits):
        need = (bits + 31) // 32
        if need > self._w.size:
            nw = np.zeros(max(need, 2 * self._w.size), dtype=np.uint32)
            nw[: self._w.size] = self._w
            self._w = nw
        if bits > self._size:
            self._size = bits

    def writeBits(self, data, count):
        data = int(data) & ((1 << count) - 1) if count < 32 else int(data) & 0xFFFFFFFF
        self._grow(self._pos + count)
        for i in range(count):
            bit = (data >> (count - 1 - i)) & 1
            p = self._pos + i
            sh = 31 - (p % 32)
            w = int(self._w[p // 32])
            self._w[p // 32] = (w & ~(1 << sh) & 0xFFFFFFFF) | (bit << sh)
        self._pos += count

    def writeAlign(self, nbytes=4):
        bits = nbytes * 8
        off = self._pos % bits
        if off:
            self._pos += bits - off
        self._grow(self._pos)

    def writeAligned(self, arr):
        arr = np.ascontiguousarray(arr)
        self.writeAlign(4)
        self.writeAlign(arr.itemsize)
        if arr.size == 0:
            return
        nbytes = arr.nbytes
        start = self._pos // 8
        self._grow(self._pos + nbytes * 8)
        self._w.view(np.uint8)[start:start + nbytes] = arr.view(np.uint8)
        self._pos += nbytes * 8
        self.writeAlign(4)

    def _append_segment(self, words, nbits, start_bit):
        """Places an encoded segment (nbits counted from bit 0 of words[0], the first start_bit of
        which are placeholders) at the current position, which must be start_bit mod 32."""
        assert self._pos % 32 == start_bit
        nw = (nbits + 31) // 32
        w0 = self._pos // 32
        self._grow(self._pos - start_bit + nbits)
        if nw:
            seg = np.asarray(words[:nw], dtype=np.uint32)
            if start_bit:
                keep = np.uint32((0xFFFFFFFF << (32 - start_bit)) & 0xFFFFFFFF)
                self._w[w0] = (self._w[w0] & keep) | seg[0]
                self._w[w0 + 1:w0 + nw] = seg[1:]
            else:
                self._w[w0:w0 + nw] = seg
        self._pos += nbits - start_bit

    # -- reading
    def readBits(self, count):
        v = 0
        for _ in range(count):
            if self._pos >= self._size:
                break
            v = (v << 1) | ((int(self._w[self._pos // 32]) >> (31 - self._pos % 32)) & 1)
            self._pos += 1
        return v

    def align(self, nbytes=4):
        bits = nbytes * 8
        off = self._pos % bits
        if off:
            self._pos = min(self._pos + bits - off, self._size)
        return True

    def readAligned(self, dtype, count):
        dtype = np.dtype(dtype)
        self.align(4)
        self.align(dtype.itemsize)
        start = self._pos // 8
        avail = (self._size - self._pos) // 8 // dtype.itemsize
        count = min(count, avail)
        out = self._w.view(np.uint8)[start:start + count * dtype.itemsize].view(dtype).copy()
        self._pos += count * dtype.itemsize * 8
        self.align(4)
        return out


# ------------------------------------------------------------------------------------ coding
def _encode(fn_name, inst, bitStreams, streams, count):
    lib = _capi.load()
    ns = len(streams)
    single = isinstance(bitStreams, BitStream)
    targets = [bitStreams] * ns if single else list(bitStreams)
    if len(targets) != ns:
        raise ValueError("one bit stream per symbol stream is required")
    size = inst.symbolSize
    for s in streams:
        es = _elem_size(s)
        if es is not None and es != size:
            raise ValueError(f"symbol streams must have {size}-byte elements for this instance")
    ptrs = (ctypes.c_void_p * max(ns, 1))(*[_ptr(s) for s in streams])
    seen = set()
    start = (ctypes.c_uint32 * max(ns, 1))()
    for i, b in enumerate(targets):
        if id(b) not in seen:
            start[i] = b.getBitPosition() % 32
            seen.add(id(b))
        else:
            start[i] = 0
    words = _capi.P_u32()
    beg = _capi.P_u64()
    bits = _capi.P_u64()
    st = getattr(lib, fn_name)(inst.handle, size, ptrs, ns, count, start, ctypes.byref(words), ctypes.byref(beg),
                               ctypes.byref(bits))
    if st == CC_ERR_SYMBOL_OVERFLOW:
        return False
    _check(st, fn_name)
    if ns == 0:
        return True
    total = int(beg[ns - 1]) + (int(bits[ns - 1]) + 31) // 32
    all_words = np.ctypeslib.as_array(words, shape=(max(total, 1),))[:total]
    for i, b in enumerate(targets):
        s0 = int(beg[i])
        nb = int(bits[i])
        b._append_segment(all_words[s0:s0 + (nb + 31) // 32], nb, int(start[i]))
    return True


def _decode(fn_name, inst, bitStreams, streams, count):
    lib = _capi.load()
    ns = len(streams)
    single = isinstance(bitStreams, BitStream)
    sources = [bitStreams] if single else []
    s2b = None
    if not single:
        index = {}
        mapping = []
        for b in bitStreams:
            if id(b) not in index:
                index[id(b)] = len(sources)
                sources.append(b)
            mapping.append(index[id(b)])
        s2b = (ctypes.c_uint32 * max(ns, 1))(*mapping)
    nbs = len(sources)
    raws = [np.ascontiguousarray(b._w[: max(b.getRawSizeUInts(), 1)]) for b in sources]
    wptrs = (ctypes.c_void_p * nbs)(*[r.ctypes.data for r in raws])
    bits = (ctypes.c_uint64 * nbs)(*[b.getBitSize() for b in sources])
    pos = (ctypes.c_uint64 * nbs)(*[b.getBitPosition() for b in sources])
    outs = (ctypes.c_void_p * max(ns, 1))(*[_ptr(s) for s in streams])
    st = getattr(lib, fn_name)(inst.handle, inst.symbolSize, wptrs, bits, pos, nbs, s2b, outs, ns, count)
    _check(st, fn_name)
    for b, p in zip(sources, pos):
        b._pos = int(p)
    return True


def encodeRLHuff(inst, bitStream, symbolStreams, symbolCountPerStream):
    """encodeRLHuff (Encode.h:28, :42; per-stream bit streams :48, :51)."""
    return _encode("cc_encode_rlhuff", inst, bitStream, symbolStreams, symbolCountPerStream)


def encodeDevice(inst, rl, symbolStreams, symbolCountPerStream, out):
    """Device-resident encodeRLHuff (rl=True) / encodeHuff: the bitstream of one call is written to
    the device tensor `out` (int32/uint32 words) by the GPU pipeline alone.  Returns the word count,
    or None when a symbol exceeds 2^log2HuffmanDistinctSymbolCountMax."""
    lib = _capi.load()
    ns = len(symbolStreams)
    ptrs = (ctypes.c_void_p * max(ns, 1))(*[_ptr(s) for s in symbolStreams])
    nw = ctypes.c_uint64()
    cap = out.numel() * out.element_size() // 4
    st = lib.cc_encode_device(inst.handle, 1 if rl else 0, inst.symbolSize, ptrs, ns, symbolCountPerStream, _ptr(out),
                              cap, ctypes.byref(nw))
    if st == CC_ERR_SYMBOL_OVERFLOW:
        return None
    _check(st, "cc_encode_device")
    return int(nw.value)


def decodeDevice(inst, rl, words, symbolStreams, symbolCountPerStream, wordCount=None):
    """Device-resident decodeRLHuff (rl=True) / decodeHuff of one call whose bitstream is the
    device tensor `words` (32-bit words).  Returns the number of words the call occupies."""
    lib = _capi.load()
    ns = len(symbolStreams)
    ptrs = (ctypes.c_void_p * max(ns, 1))(*[_ptr(s) for s in symbolStreams])
    nw = ctypes.c_uint64()
    avail = wordCount if wordCount is not None else words.numel() * words.element_size() // 4
    st = lib.cc_decode_device(inst.handle, 1 if rl else 0, inst.symbolSize, _ptr(words), avail, ptrs, ns,
                              symbolCountPerStream, ctypes.byref(nw))
    _check(st, "cc_decode_device")
    return int(nw.value)


def decodeDeviceNoAlloc(inst, rl, words, symbolStreams, symbolCountPerStream, wordCount=None):
    """decodeDevice with no allocation: the decode kernels' tables and arena live in the library's
    static per-device scratch array (32 MiB; calls on one device must not overlap).  Raises
    CCError (status CC_ERR_INVALID_ARG) for a call too large for that array."""
    lib = _capi.load()
    ns = len(symbolStreams)
    ptrs = (ctypes.c_void_p * max(ns, 1))(*[_ptr(s) for s in symbolStreams])
    nw = ctypes.c_uint64()
    avail = wordCount if wordCount is not None else words.numel() * words.element_size() // 4
    st = lib.cc_decode_device_noalloc(inst.handle, 1 if rl else 0, inst.symbolSize, _ptr(words), avail, ptrs, ns,
                                      symbolCountPerStream, ctypes.byref(nw))
    _check(st, "cc_decode_device_noalloc")
    return int(nw.value)


def _bricks(srcs, dsts, shapes, steps):
    n = len(shapes)
    arr = (_capi.Brick * max(n, 1))()
    for i, (z, y, x) in enumerate(shapes):
        arr[i].src = _ptr(srcs[i]) if srcs is not None else None
        arr[i].dst = _ptr(dsts[i]) if dsts is not None else None
        arr[i].sizeX, arr[i].sizeY, arr[i].sizeZ = int(x), int(y), int(z)
        arr[i].quantStep = float(steps[i])
    return arr


def compressBricksFloat(inst, sources, shapes, quantSteps, numLevels, out, symbolSize=2):
    """Batched compressVolumeFloat of many bricks (device tensors, shapes (z, y, x)) into the
    device word tensor `out`.  Returns (offsets, status): brick i is out[offsets[i]:offsets[i+1]]."""
    lib = _capi.load()
    n = len(shapes)
    arr = _bricks(sources, None, shapes, quantSteps)
    offs = np.zeros(n + 1, dtype=np.uint64)
    stat = np.zeros(max(n, 1), dtype=np.uint32)
    cap = out.numel() * out.element_size() // 4
    st = lib.cc_compress_bricks_float(inst.handle, arr, n, numLevels, symbolSize, _ptr(out), cap,
                                      offs.ctypes.data_as(_capi.P_u64), stat.ctypes.data_as(_capi.P_u32))
    if st not in (CC_OK, CC_ERR_SYMBOL_OVERFLOW, CC_ERR_CODEWORD_TOO_LONG):
        _check(st, "cc_compress_bricks_float")
    return offs, stat[:n]


def decompressBricksFloat(inst, dests, shapes, quantSteps, numLevels, words, offsets, symbolSize=2):
    """Batched decompressVolumeFloat of the bricks written by compressBricksFloat."""
    lib = _capi.load()
    n = len(shapes)
    arr = _bricks(None, dests, shapes, quantSteps)
    offs = np.ascontiguousarray(offsets, dtype=np.uint64)
    stat = np.zeros(max(n, 1), dtype=np.uint32)
    st = lib.cc_decompress_bricks_float(inst.handle, arr, n, numLevels, symbolSize, _ptr(words),
                                        offs.ctypes.data_as(_capi.P_u64), stat.ctypes.data_as(_capi.P_u32))
    _check(st, "cc_decompress_bricks_float")
    return stat[:n]


class BricksPlan:
    """A brick batch captured into hipGraphs (cc_bricks_plan_*): compress() and decompress()
    replay the captured pipelines on the same bricks and output buffer; run() replays both with
    one host synchronisation (cc_bricks_plan_run)."""

    def __init__(self, inst, sources, dests, shapes, quantSteps, numLevels, out, symbolSize=2):
        self._lib = _capi.load()
        self._n = len(shapes)
        self._arr = _bricks(sources, dests, shapes, quantSteps)
        self._keep = (sources, dests, out)
        h = ctypes.c_void_p()
        cap = out.numel() * out.element_size() // 4
        _check(self._lib.cc_bricks_plan_create(inst.handle, self._arr, self._n, numLevels, symbolSize, _ptr(out), cap,
                                               ctypes.byref(h)), "cc_bricks_plan_create")
        self.handle = h

    def compress(self):
        offs = np.zeros(self._n + 1, dtype=np.uint64)
        stat = np.zeros(max(self._n, 1), dtype=np.uint32)
        st = self._lib.cc_bricks_plan_compress(self.handle, offs.ctypes.data_as(_capi.P_u64),
                                               stat.ctypes.data_as(_capi.P_u32))
        if st not in (CC_OK, CC_ERR_SYMBOL_OVERFLOW, CC_ERR_CODEWORD_TOO_LONG):
            _check(st, "cc_bricks_plan_compress")
        return offs, stat[:self._n]

    def decompress(self):
        stat = np.zeros(max(self._n, 1), dtype=np.uint32)
        _check(self._lib.cc_bricks_plan_decompress(self.handle, stat.ctypes.data_as(_capi.P_u32)),
               "cc_bricks_plan_decompress")
        return stat[:self._n]

    def run(self):
        """Compression then decompression; returns (offsets, compress status, decompress status)."""
        offs = np.zeros(self._n + 1, dtype=np.uint64)
        cst = np.zeros(max(self._n, 1), dtype=np.uint32)
        dst = np.zeros(max(self._n, 1), dtype=np.uint32)
        st = self._lib.cc_bricks_plan_run(self.handle, offs.ctypes.data_as(_capi.P_u64), cst.ctypes.data_as(_capi.P_u32),
                                          dst.ctypes.data_as(_capi.P_u32))
        if st not in (CC_OK, CC_ERR_SYMBOL_OVERFLOW, CC_ERR_CODEWORD_TOO_LONG):
            _check(st, "cc_bricks_plan_run")
        return offs, cst[:self._n], dst[:self._n]

    def launch(self):
        """Enqueues compression and decompression on the plan's stream and returns at once."""
        _check(self._lib.cc_bricks_plan_launch(self.handle), "cc_bricks_plan_launch")

    def wait(self):
        """Waits for the last launch; returns (offsets, compress status, decompress status)."""
        offs = np.zeros(self._n + 1, dtype=np.uint64)
        cst = np.zeros(max(self._n, 1), dtype=np.uint32)
        dst = np.zeros(max(self._n, 1), dtype=np.uint32)
        st = self._lib.cc_bricks_plan_wait(self.handle, offs.ctypes.data_as(_capi.P_u64),
                                           cst.ctypes.data_as(_capi.P_u32), dst.ctypes.data_as(_capi.P_u32))
        if st not in (CC_OK, CC_ERR_SYMBOL_OVERFLOW, CC_ERR_CODEWORD_TOO_LONG):
            _check(st, "cc_bricks_plan_wait")
        return offs, cst[:self._n], dst[:self._n]

    def codingStats(self):
        """Counts of the last compression (cc_bricks_plan_coding_stats): dict with N (symbols),
        M (zero-run pairs), N_H (Huffman-only symbols) and C (compressed bytes)."""
        s = np.zeros(4, dtype=np.uint64)
        _check(self._lib.cc_bricks_plan_coding_stats(self.handle, s.ctypes.data_as(_capi.P_u64)),
               "cc_bricks_plan_coding_stats")
        return {"N": int(s[0]), "M": int(s[1]), "N_H": int(s[2]), "C": int(s[3]) * 4}

    def close(self):
        if self.handle:
            self._lib.cc_bricks_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Decode serving loop over resident compressed bricks (cc_server_*).

    words: device word tensor holding the bricks' bitstreams; ranges: (n, 2) word ranges;
    shapes: (z, y, x) per brick; steps: quantization step per brick."""

    def __init__(self, inst, words, ranges, shapes, steps, numLevels, slots=2, workspaceBytes=32 << 30,
                 symbolSize=2):
        self._lib = _capi.load()
        self._words = words
        n = len(shapes)
        rng = np.ascontiguousarray(np.asarray(ranges, dtype=np.uint64).reshape(n, 2))
        dims = np.ascontiguousarray(np.array([(x, y, z) for (z, y, x) in shapes], dtype=np.uint32).reshape(n, 3))
        st = np.ascontiguousarray(np.asarray(steps, dtype=np.float32))
        h = ctypes.c_void_p()
        _check(self._lib.cc_server_create(inst.handle, _ptr(words), n, rng.ctypes.data_as(_capi.P_u64),
                                          dims.ctypes.data_as(_capi.P_u32),
                                          st.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), numLevels, symbolSize,
                                          slots, workspaceBytes, ctypes.byref(h)), "cc_server_create")
        self.handle = h

    def submit(self, request_id, brick_ids, outputs, arrival):
        ids = np.ascontiguousarray(brick_ids, dtype=np.uint32)
        outs = (ctypes.c_void_p * len(outputs))(*[_ptr(o) for o in outputs])
        _check(self._lib.cc_server_submit(self.handle, request_id, ids.ctypes.data_as(_capi.P_u32), ids.size, outs,
                                          float(arrival)), "cc_server_submit")

    def poll(self, now):
        n = ctypes.c_uint32()
        _check(self._lib.cc_server_poll(self.handle, float(now), ctypes.byref(n)), "cc_server_poll")
        return n.value

    def pending(self):
        return self._lib.cc_server_pending(self.handle)

    def completed(self, maxn=65536):
        ids = np.zeros(maxn, dtype=np.uint64)
        t = np.zeros(maxn, dtype=np.float64)
        k = self._lib.cc_server_completed(self.handle, ids.ctypes.data_as(_capi.P_u64),
                                          t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), maxn)
        return ids[:k], t[:k]

    def failed(self, maxn=65536):
        ids = np.zeros(maxn, dtype=np.uint64)
        stt = np.zeros(maxn, dtype=np.uint32)
        k = self._lib.cc_server_failed(self.handle, ids.ctypes.data_as(_capi.P_u64), stt.ctypes.data_as(_capi.P_u32),
                                       maxn)
        return ids[:k], stt[:k]

    def run(self, arrivals, requests, outputs):
        """Replays a request trace in real time (cc_server_run).  requests: lists of brick ids;
        outputs: matching lists of device output pointers or tensors.  Returns (latencies in
        seconds, -1 for failed requests; elapsed seconds to the last completion)."""
        n = len(requests)
        arr = np.ascontiguousarray(arrivals, dtype=np.float64)
        cnt = np.array([len(r) for r in requests], dtype=np.uint32)
        ids = np.ascontiguousarray(np.concatenate([np.asarray(r, dtype=np.uint32) for r in requests])
                                   if n else np.zeros(0, np.uint32))
        flat = [o for req in outputs for o in req]
        outs = (ctypes.c_void_p * max(len(flat), 1))(*[int(o) if isinstance(o, (int, np.integer)) else _ptr(o)
                                                        for o in flat])
        lat = np.zeros(n, dtype=np.float64)
        el = ctypes.c_double()
        _check(self._lib.cc_server_run(self.handle, n, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       cnt.ctypes.data_as(_capi.P_u32), ids.ctypes.data_as(_capi.P_u32), outs,
                                       lat.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(el)),
               "cc_server_run")
        return lat, el.value

    def close(self):
        if self.handle:
            self._lib.cc_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def decodeRLHuff(inst, bitStream, symbolStreams, symbolCountPerStream):
    """decodeRLHuff (Encode.h:29, :43).  Outputs need not be zeroed."""
    return _decode("cc_decode_rlhuff", inst, bitStream, symbolStreams, symbolCountPerStream)


def encodeHuff(inst, bitStream, symbolStreams, symbolCountPerStream):
    """encodeHuff (Encode.h:31, :45).  Does not write padding into the inputs."""
    return _encode("cc_encode_huff", inst, bitStream, symbolStreams, symbolCountPerStream)


def decodeHuff(inst, bitStream, symbolStreams, symbolCountPerStream):
    """decodeHuff (Encode.h:32, :46).  Outputs need no padding."""
    return _decode("cc_decode_huff", inst, bitStream, symbolStreams, symbolCountPerStream)


# ------------------------------------------------------------------------------------ util
class util:
    """cudaCompress::util (include/cudaCompress/util/DWT.h, Quantize.h, YCoCg.h, Bits.h)."""

    @staticmethod
    def dwtFloat3DForward(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, srcChannelCount=1,
                          dstRowPitch=0, dstSlicePitch=0, srcRowPitch=0, srcSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_forward(_ptr(dpDest), _ptr(dpBuffer2), _ptr(dpBuffer1), _ptr(dpSource),
                                                 sizeX, sizeY, sizeZ, srcChannelCount, dstRowPitch, dstSlicePitch,
                                                 srcRowPitch, srcSlicePitch, _stream(stream)), "dwtFloat3DForward")

    @staticmethod
    def dwtDouble3DForward(dpDest, dpBuffer, dpSource, sizeX, sizeY, sizeZ, rowPitch=0, slicePitch=0, stream=None):
        """One 9/7 analysis level in float64 (cc_dwt_double_forward; no reference counterpart): the
        FP64 twin of dwtFloat3DForward / dwtFloat2DForward (sizeZ 1) that states the FP32 error."""
        _check(_capi.load().cc_dwt_double_forward(_ptr(dpDest), _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, sizeZ,
                                                  rowPitch, slicePitch, _stream(stream)), "dwtDouble3DForward")

    @staticmethod
    def dwtFloat2DForward(dpDest, dpBuffer, dpSource, sizeX, sizeY, srcChannelCount=1, dstRowPitch=0, srcRowPitch=0,
                          stream=None):
        _check(_capi.load().cc_dwt_float_forward(_ptr(dpDest), None, _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, 1,
                                                 srcChannelCount, dstRowPitch, 0, srcRowPitch, 0, _stream(stream)),
               "dwtFloat2DForward")

    @staticmethod
    def dwtFloat3DInverse(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, dstChannelCount=1,
                          dstRowPitch=0, dstSlicePitch=0, srcRowPitch=0, srcSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_inverse(_ptr(dpDest), _ptr(dpBuffer2), _ptr(dpBuffer1), _ptr(dpSource),
                                                 sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch,
                                                 srcRowPitch, srcSlicePitch, _stream(stream)), "dwtFloat3DInverse")

    @staticmethod
    def dwtFloat2DInverse(dpDest, dpBuffer, dpSource, sizeX, sizeY, dstChannelCount=1, dstRowPitch=0, srcRowPitch=0,
                          stream=None):
        _check(_capi.load().cc_dwt_float_inverse(_ptr(dpDest), None, _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, 1,
                                                 dstChannelCount, dstRowPitch, 0, srcRowPitch, 0, _stream(stream)),
               "dwtFloat2DInverse")

    @staticmethod
    def dwtFloat3DInverseFromSymbols(dpDest, dpBuffer2, dpBuffer1, dpLowpass, dppHighpass, quantStep, sizeX, sizeY,
                                     sizeZ, dstChannelCount=1, dstRowPitch=0, dstSlicePitch=0, lowpassRowPitch=0,
                                     lowpassSlicePitch=0, stream=None, symbolSize=2):
        """dppHighpass: device array (tensor of int64 addresses) of 7 band pointers."""
        _check(_capi.load().cc_dwt_float_inverse_from_symbols(
            _ptr(dpDest), _ptr(dpBuffer2), _ptr(dpBuffer1), _ptr(dpLowpass), _ptr(dppHighpass), symbolSize,
            quantStep, sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch, lowpassRowPitch,
            lowpassSlicePitch, _stream(stream)), "dwtFloat3DInverseFromSymbols")

    @staticmethod
    def dwtFloat2DInverseFromSymbols(dpDest, dpBuffer, dpLowpass, dppHighpass, quantStep, sizeX, sizeY,
                                     dstChannelCount=1, dstRowPitch=0, lowpassRowPitch=0, stream=None, symbolSize=2):
        _check(_capi.load().cc_dwt_float_inverse_from_symbols(
            _ptr(dpDest), None, _ptr(dpBuffer), _ptr(dpLowpass), _ptr(dppHighpass), symbolSize, quantStep, sizeX,
            sizeY, 1, dstChannelCount, dstRowPitch, 0, lowpassRowPitch, 0, _stream(stream)),
            "dwtFloat2DInverseFromSymbols")

    @staticmethod
    def dwtIntForward(dpDest, dpBuffer, dpSource, sizeX, sizeY, sizeZ, dstRowPitch=0, dstSlicePitch=0, srcRowPitch=0,
                      srcSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_int_forward(_ptr(dpDest), _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, sizeZ,
                                               dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch,
                                               _stream(stream)), "dwtIntForward")

    @staticmethod
    def dwtIntInverse(dpDest, dpBuffer, dpSource, sizeX, sizeY, sizeZ, dstRowPitch=0, dstSlicePitch=0, srcRowPitch=0,
                      srcSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_int_inverse(_ptr(dpDest), _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, sizeZ,
                                               dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch,
                                               _stream(stream)), "dwtIntInverse")

    @staticmethod
    def symbolize(dpSymbols, dpData, sizeX, sizeY, sizeZ, rowPitchSrc=0, slicePitchSrc=0, stream=None):
        _check(_capi.load().cc_symbolize(_ptr(dpSymbols), _ptr(dpData), sizeX, sizeY, sizeZ, rowPitchSrc,
                                         slicePitchSrc, _stream(stream)), "symbolize")

    @staticmethod
    def unsymbolize(dpData, dpSymbols, sizeX, sizeY, sizeZ, rowPitchDst=0, slicePitchDst=0, stream=None):
        _check(_capi.load().cc_unsymbolize(_ptr(dpData), _ptr(dpSymbols), sizeX, sizeY, sizeZ, rowPitchDst,
                                           slicePitchDst, _stream(stream)), "unsymbolize")

    @staticmethod
    def quantizeToSymbols(dpSymbols, dpData, sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc=0, slicePitchSrc=0,
                          quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_quantize_to_symbols(_ptr(dpSymbols), dpSymbols.element_size(), _ptr(dpData), sizeX,
                                                   sizeY, sizeZ, quantizationStep, rowPitchSrc, slicePitchSrc,
                                                   quantType, _stream(stream)), "quantizeToSymbols")

    @staticmethod
    def quantizeToSymbolsRoundtrip(dpSymbols, dpData, sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc=0,
                                   slicePitchSrc=0, quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_quantize_to_symbols_roundtrip(_ptr(dpSymbols), dpSymbols.element_size(), _ptr(dpData),
                                                             sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc,
                                                             slicePitchSrc, quantType, _stream(stream)),
               "quantizeToSymbolsRoundtrip")

    @staticmethod
    def unquantizeFromSymbols(dpData, dpSymbols, sizeX, sizeY, sizeZ, quantizationStep, rowPitchDst=0,
                              slicePitchDst=0, quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_unquantize_from_symbols(_ptr(dpData), _ptr(dpSymbols), dpSymbols.element_size(),
                                                       sizeX, sizeY, sizeZ, quantizationStep, rowPitchDst,
                                                       slicePitchDst, quantType, _stream(stream)),
               "unquantizeFromSymbols")

    @staticmethod
    def quantizeToShort(dpQuant, dpData, sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc=0, slicePitchSrc=0,
                        quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_quantize_to_short(_ptr(dpQuant), _ptr(dpData), sizeX, sizeY, sizeZ, quantizationStep,
                                                 rowPitchSrc, slicePitchSrc, quantType, _stream(stream)),
               "quantizeToShort")

    @staticmethod
    def unquantizeFromShort(dpData, dpQuant, sizeX, sizeY, sizeZ, quantizationStep, rowPitchDst=0, slicePitchDst=0,
                            quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_unquantize_from_short(_ptr(dpData), _ptr(dpQuant), sizeX, sizeY, sizeZ,
                                                     quantizationStep, rowPitchDst, slicePitchDst, quantType,
                                                     _stream(stream)), "unquantizeFromShort")

    @staticmethod
    def getMaxAbs(inst, dpImage, elemCount, dpValMax):
        _check(_capi.load().cc_get_max_abs(inst.handle, _ptr(dpImage), elemCount, _ptr(dpValMax)), "getMaxAbs")

    # ---- image path (DWT.h:81-131, Quantize.h:25-61, YCoCg.h) and bit planes (Bits.h) ----
    @staticmethod
    def dwtFloat3DForwardFromByte(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, srcChannelCount=1,
                                  dstRowPitch=0, dstSlicePitch=0, srcRowPitch=0, srcSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_forward_from_byte(
            _ptr(dpDest), _ptr(dpBuffer2), _ptr(dpBuffer1), _ptr(dpSource), sizeX, sizeY, sizeZ, srcChannelCount,
            dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch, _stream(stream)), "dwtFloat3DForwardFromByte")

    @staticmethod
    def dwtFloat2DForwardFromByte(dpDest, dpBuffer, dpSource, sizeX, sizeY, srcChannelCount=1, dstRowPitch=0,
                                  srcRowPitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_forward_from_byte(
            _ptr(dpDest), None, _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, 1, srcChannelCount, dstRowPitch, 0,
            srcRowPitch, 0, _stream(stream)), "dwtFloat2DForwardFromByte")

    @staticmethod
    def dwtFloat3DInverseToByte(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, dstChannelCount=1,
                                dstRowPitch=0, dstSlicePitch=0, srcRowPitch=0, srcSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_inverse_to_byte(
            _ptr(dpDest), _ptr(dpBuffer2), _ptr(dpBuffer1), _ptr(dpSource), sizeX, sizeY, sizeZ, dstChannelCount,
            dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch, _stream(stream)), "dwtFloat3DInverseToByte")

    @staticmethod
    def dwtFloat2DInverseToByte(dpDest, dpBuffer, dpSource, sizeX, sizeY, dstChannelCount=1, dstRowPitch=0,
                                srcRowPitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_inverse_to_byte(
            _ptr(dpDest), None, _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, 1, dstChannelCount, dstRowPitch, 0,
            srcRowPitch, 0, _stream(stream)), "dwtFloat2DInverseToByte")

    @staticmethod
    def dwtFloat3DInverseFromSymbolsToByte(dpDest, dpBuffer2, dpBuffer1, dpLowpass, dppHighpass, quantStep, sizeX,
                                           sizeY, sizeZ, dstChannelCount=1, dstRowPitch=0, dstSlicePitch=0,
                                           lowpassRowPitch=0, lowpassSlicePitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_inverse_from_symbols_to_byte(
            _ptr(dpDest), _ptr(dpBuffer2), _ptr(dpBuffer1), _ptr(dpLowpass), _ptr(dppHighpass), quantStep, sizeX,
            sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch, lowpassRowPitch, lowpassSlicePitch,
            _stream(stream)), "dwtFloat3DInverseFromSymbolsToByte")

    @staticmethod
    def dwtFloat2DInverseFromSymbolsToByte(dpDest, dpBuffer, dpLowpass, dppHighpass, quantStep, sizeX, sizeY,
                                           dstChannelCount=1, dstRowPitch=0, lowpassRowPitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_inverse_from_symbols_to_byte(
            _ptr(dpDest), None, _ptr(dpBuffer), _ptr(dpLowpass), _ptr(dppHighpass), quantStep, sizeX, sizeY, 1,
            dstChannelCount, dstRowPitch, 0, lowpassRowPitch, 0, _stream(stream)),
            "dwtFloat2DInverseFromSymbolsToByte")

    @staticmethod
    def dwtFloat2DForwardLowpassOnlyFromByte(dpDest, dpBuffer, dpSource, sizeX, sizeY, srcChannelCount=1,
                                             dstRowPitch=0, srcRowPitch=0, stream=None):
        _check(_capi.load().cc_dwt_float_forward_lowpass_only_from_byte(
            _ptr(dpDest), _ptr(dpBuffer), _ptr(dpSource), sizeX, sizeY, srcChannelCount, dstRowPitch, srcRowPitch,
            _stream(stream)), "dwtFloat2DForwardLowpassOnlyFromByte")

    @staticmethod
    def dwtIntForwardLowpassOnly(dpDest, dpBuffer, dpSource, sizeX, sizeY, dstRowPitch=0, srcRowPitch=0,
                                 stream=None):
        _check(_capi.load().cc_dwt_int_forward_lowpass_only(_ptr(dpDest), _ptr(dpBuffer), _ptr(dpSource), sizeX,
                                                            sizeY, dstRowPitch, srcRowPitch, _stream(stream)),
               "dwtIntForwardLowpassOnly")

    @staticmethod
    def floatToByte2D(dpQuant, channelCount, channel, dpData, sizeX, sizeY, rowPitchSrc=0, stream=None):
        _check(_capi.load().cc_float_to_byte_2d(_ptr(dpQuant), channelCount, channel, _ptr(dpData), sizeX, sizeY,
                                                rowPitchSrc, _stream(stream)), "floatToByte2D")

    @staticmethod
    def byteToFloat2D(dpData, dpQuant, channelCount, channel, sizeX, sizeY, rowPitchDst=0, stream=None):
        _check(_capi.load().cc_byte_to_float_2d(_ptr(dpData), _ptr(dpQuant), channelCount, channel, sizeX, sizeY,
                                                rowPitchDst, _stream(stream)), "byteToFloat2D")

    @staticmethod
    def quantizeDifferenceToSymbols2D(dpSymbols, dpData, quantizationStep, dpReference, channelCount, channel, sizeX,
                                      sizeY, rowPitchSrc=0, quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_quantize_difference_to_symbols_2d(
            _ptr(dpSymbols), _ptr(dpData), quantizationStep, _ptr(dpReference), channelCount, channel, sizeX, sizeY,
            rowPitchSrc, quantType, _stream(stream)), "quantizeDifferenceToSymbols2D")

    @staticmethod
    def quantizeDifferenceToSymbolsRoundtrip2D(dpSymbols, dpData, quantizationStep, dpReference, channelCount,
                                               channel, sizeX, sizeY, rowPitchSrc=0, quantType=QUANTIZE_DEADZONE,
                                               stream=None):
        _check(_capi.load().cc_quantize_difference_to_symbols_roundtrip_2d(
            _ptr(dpSymbols), _ptr(dpData), quantizationStep, _ptr(dpReference), channelCount, channel, sizeX, sizeY,
            rowPitchSrc, quantType, _stream(stream)), "quantizeDifferenceToSymbolsRoundtrip2D")

    @staticmethod
    def unquantizeDifferenceFromSymbols2D(dpData, dpSymbols, quantizationStep, dpReference, channelCount, channel,
                                          sizeX, sizeY, rowPitchDst=0, quantType=QUANTIZE_DEADZONE, stream=None):
        _check(_capi.load().cc_unquantize_difference_from_symbols_2d(
            _ptr(dpData), _ptr(dpSymbols), quantizationStep, _ptr(dpReference), channelCount, channel, sizeX, sizeY,
            rowPitchDst, quantType, _stream(stream)), "unquantizeDifferenceFromSymbols2D")

    @staticmethod
    def convertRGBtoYCoCg(dpTarget, dpData, pixelCount, channels=3, stream=None):
        _check(_capi.load().cc_convert_rgb_to_ycocg(_ptr(dpTarget), _ptr(dpData), pixelCount, channels,
                                                    _stream(stream)), "convertRGBtoYCoCg")

    @staticmethod
    def convertYCoCgtoRGB(dpTarget, dpData, pixelCount, channels=3, stream=None):
        _check(_capi.load().cc_convert_ycocg_to_rgb(_ptr(dpTarget), _ptr(dpData), pixelCount, channels,
                                                    _stream(stream)), "convertYCoCgtoRGB")

    @staticmethod
    def getLSB(dpData, dpBits, elemCount, stream=None):
        _check(_capi.load().cc_get_lsb(_ptr(dpData), _ptr(dpBits), elemCount, _stream(stream)), "getLSB")

    @staticmethod
    def removeLSB(dpData, dpBits, elemCount, shift=0, stream=None):
        _check(_capi.load().cc_remove_lsb(_ptr(dpData), _ptr(dpBits), elemCount, shift, _stream(stream)), "removeLSB")

    @staticmethod
    def appendLSB(dpData, dpBits, elemCount, shift=0, stream=None):
        _check(_capi.load().cc_append_lsb(_ptr(dpData), _ptr(dpBits), elemCount, shift, _stream(stream)), "appendLSB")

    @staticmethod
    def shiftElems(dpData, elemCount, shift, stream=None):
        _check(_capi.load().cc_shift_elems(_ptr(dpData), elemCount, shift, _stream(stream)), "shiftElems")


# ------------------------------------------------------------------------------------ timing
TIMING_DETAIL_NONE, TIMING_DETAIL_LOW, TIMING_DETAIL_MEDIUM, TIMING_DETAIL_HIGH = 0, 1, 2, 3


def setTimingDetail(inst, detail):
    """setTimingDetail (include/cudaCompress/Timing.h:23)."""
    _check(_capi.load().cc_set_timing_detail(inst.handle, int(detail)), "setTimingDetail")


def getTimings(inst):
    """getTimings (include/cudaCompress/Timing.h:25): (names, times in ms)."""
    lib = _capi.load()
    n = ctypes.c_uint32()
    _check(lib.cc_get_timings(inst.handle, ctypes.byref(n)), "getTimings")
    names, times = [], []
    for i in range(n.value):
        name = ctypes.c_char_p()
        ms = ctypes.c_float()
        _check(lib.cc_get_timing_entry(inst.handle, i, ctypes.byref(name), ctypes.byref(ms)), "getTimings")
        names.append(name.value.decode())
        times.append(ms.value)
    return names, times


def printTimings(inst):
    _check(_capi.load().cc_print_timings(inst.handle), "printTimings")


def resetTimings(inst):
    _check(_capi.load().cc_reset_timings(inst.handle), "resetTimings")


# ------------------------------------------------------------------------------------ volume codecs
def compressVolumeFloat(inst, dpImage, sizeX, sizeY, sizeZ, numLevels, quantizationStepLevel0,
                        doRLEOnlyOnLvl0=False):
    """compressVolumeFloat (src/examples/CompressVolume.cpp:313); returns the bitstream words,
    or None on symbol overflow (the reference returns false)."""
    lib = _capi.load()
    words = _capi.P_u32()
    nbits = ctypes.c_uint64()
    st = lib.cc_compress_volume_float(inst.handle, _ptr(dpImage), sizeX, sizeY, sizeZ, numLevels,
                                      quantizationStepLevel0, int(doRLEOnlyOnLvl0), inst.symbolSize,
                                      ctypes.byref(words), ctypes.byref(nbits))
    if st == CC_ERR_SYMBOL_OVERFLOW:
        return None
    _check(st, "compressVolumeFloat")
    n = nbits.value // 32
    return np.ctypeslib.as_array(words, shape=(max(n, 1),))[:n].copy()


def decompressVolumeFloat(inst, dpImage, sizeX, sizeY, sizeZ, numLevels, bits, quantizationStepLevel0,
                          doRLEOnlyOnLvl0=False):
    """decompressVolumeFloat (src/examples/CompressVolume.cpp:412)."""
    bits = np.ascontiguousarray(bits, dtype=np.uint32)
    _check(_capi.load().cc_decompress_volume_float(inst.handle, _ptr(dpImage), sizeX, sizeY, sizeZ, numLevels,
                                                   bits.ctypes.data, bits.size * 32, quantizationStepLevel0,
                                                   int(doRLEOnlyOnLvl0), inst.symbolSize), "decompressVolumeFloat")


def compressVolumeFloatQuantFirst(inst, dpImage, sizeX, sizeY, sizeZ, numLevels, quantizationStep,
                                  doRLEOnlyOnLvl0=False):
    """compressVolumeFloatQuantFirst (src/examples/CompressVolume.cpp:651); returns the bitstream
    words, or None on symbol overflow (the reference returns false)."""
    lib = _capi.load()
    words = _capi.P_u32()
    nbits = ctypes.c_uint64()
    st = lib.cc_compress_volume_float_quant_first(inst.handle, _ptr(dpImage), sizeX, sizeY, sizeZ, numLevels,
                                                  quantizationStep, int(doRLEOnlyOnLvl0), ctypes.byref(words),
                                                  ctypes.byref(nbits))
    if st == CC_ERR_SYMBOL_OVERFLOW:
        return None
    _check(st, "compressVolumeFloatQuantFirst")
    n = nbits.value // 32
    return np.ctypeslib.as_array(words, shape=(max(n, 1),))[:n].copy()


def decompressVolumeFloatQuantFirst(inst, dpImage, sizeX, sizeY, sizeZ, numLevels, bits, quantizationStep,
                                    doRLEOnlyOnLvl0=False):
    """decompressVolumeFloatQuantFirst (src/examples/CompressVolume.cpp:759)."""
    bits = np.ascontiguousarray(bits, dtype=np.uint32)
    _check(_capi.load().cc_decompress_volume_float_quant_first(inst.handle, _ptr(dpImage), sizeX, sizeY, sizeZ,
                                                               numLevels, bits.ctypes.data, bits.size * 32,
                                                               quantizationStep, int(doRLEOnlyOnLvl0)),
           "decompressVolumeFloatQuantFirst")


def _channels(channels):
    """(dpImage, bits, quantStep) triples -> a cc_volume_channel array and the word arrays it points
    into (kept alive by the caller for the duration of the call)."""
    keep = [np.ascontiguousarray(b, dtype=np.uint32) for _, b, _ in channels]
    arr = (_capi.VolumeChannel * len(channels))()
    for i, ((img, _, q), w) in enumerate(zip(channels, keep)):
        arr[i] = _capi.VolumeChannel(_ptr(img), w.ctypes.data, w.size * 32, q)
    return arr, keep


def decompressVolumeFloatMultiChannel(inst, channels, sizeX, sizeY, sizeZ, numLevels, doRLEOnlyOnLvl0=False,
                                      symbolSize=2):
    """decompressVolumeFloatMultiChannel (src/examples/CompressVolume.cpp:518); `channels` holds
    (dpImage, bits, quantizationStepLevel0) per channel, as VolumeChannel does
    (src/examples/CompressVolume.h:36)."""
    if not channels:
        raise ValueError("decompressVolumeFloatMultiChannel: no channels")
    arr, keep = _channels(channels)
    _check(_capi.load().cc_decompress_volume_float_multichannel(inst.handle, ctypes.addressof(arr), len(channels),
                                                                sizeX, sizeY, sizeZ, numLevels, int(doRLEOnlyOnLvl0),
                                                                symbolSize),
           "decompressVolumeFloatMultiChannel")
    del keep


def decompressVolumeFloatQuantFirstMultiChannel(inst, channels, sizeX, sizeY, sizeZ, numLevels,
                                                doRLEOnlyOnLvl0=False):
    """decompressVolumeFloatQuantFirstMultiChannel (src/examples/CompressVolume.cpp:863); `channels`
    holds (dpImage, bits, quantizationStep) per channel."""
    if not channels:
        raise ValueError("decompressVolumeFloatQuantFirstMultiChannel: no channels")
    arr, keep = _channels(channels)
    _check(_capi.load().cc_decompress_volume_float_quant_first_multichannel(
        inst.handle, ctypes.addressof(arr), len(channels), sizeX, sizeY, sizeZ, numLevels, int(doRLEOnlyOnLvl0)),
        "decompressVolumeFloatQuantFirstMultiChannel")
    del keep


def compressVolumeLossless(inst, dpImage, sizeX, sizeY, sizeZ, numLevels):
    """compressVolumeLossless (src/examples/CompressVolume.cpp:234); sizeZ == 1 for 2D fields."""
    lib = _capi.load()
    words = _capi.P_u32()
    nbits = ctypes.c_uint64()
    st = lib.cc_compress_volume_lossless(inst.handle, _ptr(dpImage), sizeX, sizeY, sizeZ, numLevels,
                                         ctypes.byref(words), ctypes.byref(nbits))
    if st == CC_ERR_SYMBOL_OVERFLOW:
        return None
    _check(st, "compressVolumeLossless")
    n = nbits.value // 32
    return np.ctypeslib.as_array(words, shape=(max(n, 1),))[:n].copy()


def decompressVolumeLossless(inst, dpImage, sizeX, sizeY, sizeZ, numLevels, bits):
    """decompressVolumeLossless (src/examples/CompressVolume.cpp:279)."""
    bits = np.ascontiguousarray(bits, dtype=np.uint32)
    _check(_capi.load().cc_decompress_volume_lossless(inst.handle, _ptr(dpImage), sizeX, sizeY, sizeZ, numLevels,
                                                      bits.ctypes.data, bits.size * 32), "decompressVolumeLossless")
