"""cudacompress_amd: MI355X (gfx950) implementation of the cudaCompress wavelet compression path.

The compute lives in libcudacompress_amd.so (hand-written HIP kernels behind a C ABI,
include/cudacompress_amd.h); this package is the Python host-side mirror of the reference's
C++ API.  Importing the package loads the native library and fails if it has not been built.
"""
from . import _capi

_capi.load()

from .api import (  # noqa: E402
    BitStream,
    BricksPlan,
    CCError,
    CC_ERR_CODEWORD_TOO_LONG,
    CC_ERR_CORRUPT_STREAM,
    CC_ERR_DEVICE,
    CC_ERR_INVALID_ARG,
    CC_ERR_OUT_OF_MEMORY,
    CC_ERR_SYMBOL_OVERFLOW,
    CC_OK,
    DIALECT_CPU,
    DIALECT_GPU,
    Instance,
    QUANTIZE_DEADZONE,
    QUANTIZE_UNIFORM,
    Server,
    compressBricksFloat,
    compressVolumeFloat,
    compressVolumeFloatQuantFirst,
    compressVolumeLossless,
    createInstance,
    defaultTuning,
    loadTuning,
    decodeDevice,
    decodeDeviceNoAlloc,
    decodeHuff,
    decodeRLHuff,
    decompressBricksFloat,
    decompressVolumeFloat,
    decompressVolumeFloatQuantFirst,
    decompressVolumeFloatMultiChannel,
    setTimingDetail,
    getTimings,
    printTimings,
    resetTimings,
    TIMING_DETAIL_NONE,
    TIMING_DETAIL_LOW,
    TIMING_DETAIL_MEDIUM,
    TIMING_DETAIL_HIGH,
    decompressVolumeFloatQuantFirstMultiChannel,
    decompressVolumeLossless,
    destroyInstance,
    encodeDevice,
    encodeHuff,
    encodeRLHuff,
    util,
)

LIB_PATH = _capi.LIB_PATH
