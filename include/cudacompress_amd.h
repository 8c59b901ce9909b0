/*
 * cudacompress_amd -- C ABI of the MI355X (gfx950) implementation of the cudaCompress
 * wavelet compression path.
 *
 * Every entry point below replaces one reference interface (m0bl0/cudaCompress, paths relative
 * to that repository); the cited lines are the declaration being replaced.  Arguments keep the
 * reference's meaning and order; device pointers are plain pointers, HIP streams are passed as
 * `void*` (hipStream_t), and every function returns a status code instead of printing and
 * continuing.  The C++ drop-in API in include/cudaCompress/ is a set of inline wrappers over
 * these functions.
 *
 * Threading: an instance is not thread-safe (as in the reference); use one per host thread.
 */
#ifndef CUDACOMPRESS_AMD_H
#define CUDACOMPRESS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CC_API __attribute__((visibility("default")))

typedef enum cc_status {
    CC_OK = 0,
    CC_ERR_SYMBOL_OVERFLOW = 1,    /* a symbol >= 2^log2HuffmanDistinctSymbolCountMax (ref: encode* returns false) */
    CC_ERR_OUT_OF_MEMORY = 2,
    CC_ERR_INVALID_ARG = 3,
    CC_ERR_DEVICE = 4,             /* a HIP runtime call failed */
    CC_ERR_CODEWORD_TOO_LONG = 5,  /* a Huffman codeword > 31 bits: not representable in the 5-bit header field */
    CC_ERR_CORRUPT_STREAM = 6      /* bitstream inconsistent with the requested decode */
} cc_status;

/* Bitstream dialects.  GPU: the reference library's layout (src/cudaCompress/HuffmanTable.cu:470-485).
 * CPU: the reference CPU coder's layout (src/CPU/HuffmanTableCPU.cpp:108-135; Symbol16 only). */
enum { CC_DIALECT_GPU = 0, CC_DIALECT_CPU = 1 };

typedef struct cc_instance cc_instance;

CC_API const char* cc_status_string(cc_status s);
/* Text of the last error raised on the calling thread (empty when none). */
CC_API const char* cc_last_error(void);

/* ---- Instance: include/cudaCompress/Instance.h:26-45 -------------------------------------- */

/* createInstance, include/cudaCompress/Instance.h:26-31.  codingBlockSize 0 -> 128,
 * log2HuffmanDistinctSymbolCountMax 0 -> 14 (src/cudaCompress/InstanceImpl.cpp:65-75). */
CC_API cc_status cc_create_instance(int device, uint32_t streamCountMax, uint32_t elemCountPerStreamMax,
                                    uint32_t codingBlockSize, uint32_t log2HuffmanDistinctSymbolCountMax,
                                    cc_instance** out);
/* destroyInstance, include/cudaCompress/Instance.h:33 */
CC_API void cc_destroy_instance(cc_instance* inst);
/* getInstance*, include/cudaCompress/Instance.h:36-41 */
CC_API int cc_get_instance_device(const cc_instance* inst);
CC_API uint32_t cc_get_instance_stream_count_max(const cc_instance* inst);
CC_API uint32_t cc_get_instance_elem_count_per_stream_max(const cc_instance* inst);
CC_API uint32_t cc_get_instance_coding_block_size(const cc_instance* inst);
CC_API uint32_t cc_get_instance_log2_huffman_distinct_symbol_count_max(const cc_instance* inst);
CC_API int cc_get_instance_use_long_symbols(const cc_instance* inst);
/* setInstanceCudaStream/getInstanceCudaStream, include/cudaCompress/Instance.h:44-45
 * (implemented as setInstanceStream/getInstanceStream, src/cudaCompress/Instance.cpp:59-67) */
CC_API void cc_set_instance_stream(cc_instance* inst, void* stream);
CC_API void* cc_get_instance_stream(const cc_instance* inst);
/* Bitstream dialect of this instance (CC_DIALECT_GPU by default).  No reference counterpart as a
 * switch: the CPU dialect is what src/CPU/EncodeCPU.cpp:98-354 reads and writes. */
CC_API cc_status cc_set_instance_dialect(cc_instance* inst, int dialect);
/* Groups of an eager brick batch (cc_compress/decompress_bricks_float): contiguous brick ranges
 * run as concurrent pipelines on their own streams, their bitstreams still back to back in brick
 * order.  0 or 1 (default): one pipeline.  Plans always capture one.  No reference counterpart. */
CC_API cc_status cc_set_instance_brick_groups(cc_instance* inst, uint32_t groups);

/* ---- Entropy coding: include/cudaCompress/Encode.h:28-57 ---------------------------------- */

/* encodeRLHuff / encodeHuff (all four overloads each), include/cudaCompress/Encode.h:28,31,34,37,42,45,48,51.
 * symbolSize: 2 (Symbol16) or 4 (Symbol32); must match the instance (log2 > 16 <=> 4).
 * d_symbolStreams: host array of streamCount device pointers, symbolCountPerStream symbols each.
 * startBitOffsets: per stream, the bit offset (0..31) inside the first word at which that stream's
 *   data will be placed by the caller (NULL = all 0); bits before it in the first word are zero.
 * On success *outWords points at instance-owned host memory valid until the next encode call on
 * this instance; stream i occupies words [outSegmentWordBegin[i], ...) and outSegmentBits[i] bits
 * counted from bit 0 of its first word (so they include startBitOffsets[i] leading zero bits).
 * Every segment ends 32-bit aligned.  Unlike the reference, encodeHuff does not write padding
 * into the caller's symbol arrays. */
CC_API cc_status cc_encode_rlhuff(cc_instance* inst, int symbolSize, const void* const* d_symbolStreams,
                                  uint32_t streamCount, uint32_t symbolCountPerStream,
                                  const uint32_t* startBitOffsets, const uint32_t** outWords,
                                  const uint64_t** outSegmentWordBegin, const uint64_t** outSegmentBits);
CC_API cc_status cc_encode_huff(cc_instance* inst, int symbolSize, const void* const* d_symbolStreams,
                                uint32_t streamCount, uint32_t symbolCountPerStream,
                                const uint32_t* startBitOffsets, const uint32_t** outWords,
                                const uint64_t** outSegmentWordBegin, const uint64_t** outSegmentBits);

/* decodeRLHuff / decodeHuff, include/cudaCompress/Encode.h:29,32,35,38,43,46,49,52.
 * bitstreamCount host bitstreams (words, size in bits, current position in bits -- advanced on
 * return).  streamToBitstream[i] names the bitstream stream i is read from (NULL: all streams
 * from bitstream 0, back to back, as the single-BitStream overloads do).
 * decodeRLHuff outputs need not be zeroed by the caller (the reference requires it; this
 * implementation writes every element).  decodeHuff outputs need no padding.
 * Returns with the work enqueued on the instance stream (as the reference). */
CC_API cc_status cc_decode_rlhuff(cc_instance* inst, int symbolSize, const uint32_t* const* bitstreamWords,
                                  const uint64_t* bitstreamBits, uint64_t* bitstreamPos, uint32_t bitstreamCount,
                                  const uint32_t* streamToBitstream, void* const* d_symbolStreams,
                                  uint32_t streamCount, uint32_t symbolCountPerStream);
CC_API cc_status cc_decode_huff(cc_instance* inst, int symbolSize, const uint32_t* const* bitstreamWords,
                                const uint64_t* bitstreamBits, uint64_t* bitstreamPos, uint32_t bitstreamCount,
                                const uint32_t* streamToBitstream, void* const* d_symbolStreams,
                                uint32_t streamCount, uint32_t symbolCountPerStream);

/* Device-resident encodeRLHuff (rl != 0) / encodeHuff (rl == 0): the bitstream of one call --
 * the same words cc_encode_rlhuff / cc_encode_huff return with all start bit offsets 0 -- is
 * written to device memory d_out (capacity outCapacityWords) by the GPU pipeline alone (RLE,
 * histograms, Huffman table design, packing and layout all on the device).  *outWords receives
 * the word count.  Synchronises the instance stream once, at the end.  No reference counterpart
 * as an entry point: it is encodeRLHuff/encodeHuff (include/cudaCompress/Encode.h:28-52) with the
 * BitStream kept in HBM. */
CC_API cc_status cc_encode_device(cc_instance* inst, int rl, int symbolSize, const void* const* d_symbolStreams,
                                  uint32_t streamCount, uint32_t symbolCountPerStream, uint32_t* d_out,
                                  uint64_t outCapacityWords, uint64_t* outWords);

/* Device-resident decodeRLHuff (rl != 0) / decodeHuff (rl == 0) of one call whose bitstream
 * starts at d_words (word aligned, wordCount words readable): table headers, block offsets,
 * codewords and zero runs are all read and decoded on the GPU into the streamCount device output
 * streams.  *outWordsConsumed receives the call's length in words.  A malformed stream returns
 * CC_ERR_CORRUPT_STREAM.  Synchronises the instance stream once, at the end.  decodeRLHuff /
 * decodeHuff (include/cudaCompress/Encode.h:29-52) with the BitStream in HBM. */
CC_API cc_status cc_decode_device(cc_instance* inst, int rl, int symbolSize, const uint32_t* d_words,
                                  uint64_t wordCount, void* const* d_symbolStreams, uint32_t streamCount,
                                  uint32_t symbolCountPerStream, uint64_t* outWordsConsumed);

/* cc_decode_device with no allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (serving loops and other code that must not call hipMalloc).
 * Same arguments, results and errors as cc_decode_device, but every buffer the decode kernels
 * need beyond the bitstream and the outputs (call, item and wave tables, sequence records, job
 * state, decode arena) lives in one static device array of 32 MiB per device (held by each device
 * the library runs on until the process ends); nothing is allocated, on the device or pinned on
 * the host.  At most 128 streams per call; a call whose tables and arena exceed the array returns
 * CC_ERR_INVALID_ARG before any launch.  Every call on a device shares that device's array: calls
 * that overlap on one device race.  With kernel timing on (timing detail 3 or cc_profile_enable)
 * the timer creates HIP events. */
CC_API cc_status cc_decode_device_noalloc(cc_instance* inst, int rl, int symbolSize, const uint32_t* d_words,
                                          uint64_t wordCount, void* const* d_symbolStreams, uint32_t streamCount,
                                          uint32_t symbolCountPerStream, uint64_t* outWordsConsumed);

/* ---- Timing: include/cudaCompress/Timing.h:15-27 -------------------------------------------- */
/* setTimingDetail: 0 none, 1 whole coder calls ("Encode RLHuff", ...), 2 also their stages
 * ("Encode: Histograms", ...), 3 also every kernel launch ("Kernel: <name>"; kernel timing is
 * process-wide, cc_profile_enable: switched on by the instance that enters level 3 and off when it
 * leaves it).  Sections are hipEvent pairs on the instance stream. */
CC_API cc_status cc_set_timing_detail(cc_instance* inst, int detail);
/* getTimings: waits for the recorded sections and snapshots (name, accumulated ms) pairs, read
 * with cc_get_timing_entry(i) until the next cc_get_timings call. */
CC_API cc_status cc_get_timings(cc_instance* inst, uint32_t* count);
CC_API cc_status cc_get_timing_entry(cc_instance* inst, uint32_t i, const char** name, float* ms);
/* printTimings / resetTimings */
CC_API cc_status cc_print_timings(cc_instance* inst);
CC_API cc_status cc_reset_timings(cc_instance* inst);

/* ---- DWT: include/cudaCompress/util/DWT.h:21-155 ------------------------------------------ */

/* dwtFloat3DForward / dwtFloat2DForward (sizeZ == 1, dpBuffer2 unused), util/DWT.h:21-31.
 * Pitches in elements; 0 selects the reference defaults (src/cudaCompress/util/DWTFloat.cu:474-477). */
CC_API cc_status cc_dwt_float_forward(float* d_dst, float* d_buffer2, float* d_buffer1, const float* d_src,
                                      int sizeX, int sizeY, int sizeZ, int srcChannelCount,
                                      int dstRowPitch, int dstSlicePitch, int srcRowPitch, int srcSlicePitch,
                                      void* stream);
/* dwtFloat3DInverse / dwtFloat2DInverse, util/DWT.h:35-45 */
/* No reference counterpart: one level of the CDF 9/7 analysis of dwtFloat3DForward /
 * dwtFloat2DForward (src/cudaCompress/util/DWTFloat.cu:85-202) in double precision, with the
 * filter coefficients as decimal doubles, to state the error of the fp32 transform against an
 * fp64 one.  d_dst and d_src share rowPitch and slicePitch (elements; 0: dense) and must not
 * overlap; d_buffer holds sizeX*sizeY*sizeZ at the same pitches.  sizeZ = 1: 2D. */
CC_API cc_status cc_dwt_double_forward(double* d_dst, double* d_buffer, const double* d_src, int sizeX, int sizeY,
                                       int sizeZ, int rowPitch, int slicePitch, void* stream);
CC_API cc_status cc_dwt_float_inverse(float* d_dst, float* d_buffer2, float* d_buffer1, const float* d_src,
                                      int sizeX, int sizeY, int sizeZ, int dstChannelCount,
                                      int dstRowPitch, int dstSlicePitch, int srcRowPitch, int srcSlicePitch,
                                      void* stream);
/* dwtFloat{2D,3D}InverseFromSymbols (ushort and uint highpass), util/DWT.h:48-79.
 * dd_highpass: DEVICE array of 3 (2D) or 7 (3D) device pointers to dense highpass symbol bands. */
CC_API cc_status cc_dwt_float_inverse_from_symbols(float* d_dst, float* d_buffer2, float* d_buffer1,
                                                   const float* d_lowpass, const void* const* dd_highpass,
                                                   int symbolSize, float quantStep, int sizeX, int sizeY, int sizeZ,
                                                   int dstChannelCount, int dstRowPitch, int dstSlicePitch,
                                                   int lowpassRowPitch, int lowpassSlicePitch, void* stream);
/* dwtIntForward / dwtIntInverse, util/DWT.h:139-151 (sizeZ == 1: 2D) */
CC_API cc_status cc_dwt_int_forward(int16_t* d_dst, int16_t* d_buffer, const int16_t* d_src,
                                    int sizeX, int sizeY, int sizeZ, int dstRowPitch, int dstSlicePitch,
                                    int srcRowPitch, int srcSlicePitch, void* stream);
CC_API cc_status cc_dwt_int_inverse(int16_t* d_dst, int16_t* d_buffer, const int16_t* d_src,
                                    int sizeX, int sizeY, int sizeZ, int dstRowPitch, int dstSlicePitch,
                                    int srcRowPitch, int srcSlicePitch, void* stream);
/* dwtFloat{2D,3D}ForwardFromByte, util/DWT.h:81-94: byte samples enter as value - 128. */
CC_API cc_status cc_dwt_float_forward_from_byte(float* d_dst, float* d_buffer2, float* d_buffer1,
                                                const uint8_t* d_src, int sizeX, int sizeY, int sizeZ,
                                                int srcChannelCount, int dstRowPitch, int dstSlicePitch,
                                                int srcRowPitch, int srcSlicePitch, void* stream);
/* dwtFloat{2D,3D}InverseToByte, util/DWT.h:96-107: outputs rounded half up, clamped to [0,255]
 * after adding 128. */
CC_API cc_status cc_dwt_float_inverse_to_byte(uint8_t* d_dst, float* d_buffer2, float* d_buffer1,
                                              const float* d_src, int sizeX, int sizeY, int sizeZ,
                                              int dstChannelCount, int dstRowPitch, int dstSlicePitch,
                                              int srcRowPitch, int srcSlicePitch, void* stream);
/* dwtFloat{2D,3D}InverseFromSymbolsToByte (ushort highpass), util/DWT.h:109-122 */
CC_API cc_status cc_dwt_float_inverse_from_symbols_to_byte(uint8_t* d_dst, float* d_buffer2, float* d_buffer1,
                                                           const float* d_lowpass,
                                                           const uint16_t* const* dd_highpass, float quantStep,
                                                           int sizeX, int sizeY, int sizeZ, int dstChannelCount,
                                                           int dstRowPitch, int dstSlicePitch, int lowpassRowPitch,
                                                           int lowpassSlicePitch, void* stream);
/* dwtFloat2DForwardLowpassOnlyFromByte, util/DWT.h:125-131: the LL band only, (sizeX/2) x
 * (sizeY/2) at row pitch dstRowPitch (default sizeX/2); d_buffer holds sizeY rows of the same
 * pitch. */
CC_API cc_status cc_dwt_float_forward_lowpass_only_from_byte(float* d_dst, float* d_buffer, const uint8_t* d_src,
                                                             int sizeX, int sizeY, int srcChannelCount,
                                                             int dstRowPitch, int srcRowPitch, void* stream);
/* dwtIntForwardLowpassOnly, util/DWT.h:152-156: the LL band of the 2D 5/3 transform (every even
 * size; the reference computes whole 512 x 64 tiles only). */
CC_API cc_status cc_dwt_int_forward_lowpass_only(int16_t* d_dst, int16_t* d_buffer, const int16_t* d_src, int sizeX,
                                                 int sizeY, int dstRowPitch, int srcRowPitch, void* stream);

/* ---- Quantization: include/cudaCompress/util/Quantize.h:29-65 ----------------------------- */

/* floatToByte2D / byteToFloat2D, util/Quantize.h:25-26: channel `channel` of an image of
 * channelCount interleaved byte channels <-> a float plane of row pitch rowPitch (0: sizeX). */
CC_API cc_status cc_float_to_byte_2d(uint8_t* d_quant, uint32_t channelCount, uint32_t channel, const float* d_data,
                                     uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchSrc, void* stream);
CC_API cc_status cc_byte_to_float_2d(float* d_data, const uint8_t* d_quant, uint32_t channelCount, uint32_t channel,
                                     uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchDst, void* stream);
/* quantizeDifferenceToSymbols2D / ...Roundtrip2D / unquantizeDifferenceFromSymbols2D,
 * util/Quantize.h:58-61: symbols of (data - (reference byte - 128)). */
CC_API cc_status cc_quantize_difference_to_symbols_2d(uint16_t* d_symbols, const float* d_data, float quantizationStep,
                                                      const uint8_t* d_reference, uint32_t channelCount,
                                                      uint32_t channel, uint32_t sizeX, uint32_t sizeY,
                                                      uint32_t rowPitchSrc, int quantType, void* stream);
CC_API cc_status cc_quantize_difference_to_symbols_roundtrip_2d(uint16_t* d_symbols, float* d_data,
                                                                float quantizationStep, const uint8_t* d_reference,
                                                                uint32_t channelCount, uint32_t channel,
                                                                uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchSrc,
                                                                int quantType, void* stream);
CC_API cc_status cc_unquantize_difference_from_symbols_2d(float* d_data, const uint16_t* d_symbols,
                                                          float quantizationStep, const uint8_t* d_reference,
                                                          uint32_t channelCount, uint32_t channel, uint32_t sizeX,
                                                          uint32_t sizeY, uint32_t rowPitchDst, int quantType,
                                                          void* stream);

/* ---- Colour: include/cudaCompress/util/YCoCg.h:16-20 ---------------------------------------- */
/* convertRGBtoYCoCg / convertYCoCgtoRGB on uchar3 (channels 3) or uchar4 (channels 4, the fourth
 * channel copied) pixels. */
CC_API cc_status cc_convert_rgb_to_ycocg(uint8_t* d_target, const uint8_t* d_data, int pixelCount, int channels,
                                         void* stream);
CC_API cc_status cc_convert_ycocg_to_rgb(uint8_t* d_target, const uint8_t* d_data, int pixelCount, int channels,
                                         void* stream);

/* ---- Bit planes: include/cudaCompress/util/Bits.h:16-27 ------------------------------------- */
/* Word w of d_bits holds bit i = least significant bit of element 32*w + i; ceil(elemCount/32)
 * words (any elemCount; the reference needs multiples of 4096).  cc_remove_lsb: v += shift, the
 * bit is taken, v = floor(v / 2); d_bits may be NULL.  cc_append_lsb: v = 2*v + bit + shift. */
CC_API cc_status cc_get_lsb(const int16_t* d_data, uint32_t* d_bits, uint32_t elemCount, void* stream);
CC_API cc_status cc_remove_lsb(int16_t* d_data, uint32_t* d_bits, uint32_t elemCount, int16_t shift, void* stream);
CC_API cc_status cc_append_lsb(int16_t* d_data, const uint32_t* d_bits, uint32_t elemCount, int16_t shift,
                               void* stream);
CC_API cc_status cc_shift_elems(int16_t* d_data, uint32_t elemCount, int16_t shift, void* stream);

/* symbolize / unsymbolize, util/Quantize.h:29-30 */
CC_API cc_status cc_symbolize(uint16_t* d_symbols, const int16_t* d_data, uint32_t sizeX, uint32_t sizeY,
                              uint32_t sizeZ, uint32_t rowPitchSrc, uint32_t slicePitchSrc, void* stream);
CC_API cc_status cc_unsymbolize(int16_t* d_data, const uint16_t* d_symbols, uint32_t sizeX, uint32_t sizeY,
                                uint32_t sizeZ, uint32_t rowPitchDst, uint32_t slicePitchDst, void* stream);
/* quantizeToSymbols / quantizeToSymbolsRoundtrip / unquantizeFromSymbols (ushort and uint,
 * 3D and 2D), util/Quantize.h:34-50.  quantType: 0 deadzone, 1 uniform. */
CC_API cc_status cc_quantize_to_symbols(void* d_symbols, int symbolSize, const float* d_data, uint32_t sizeX,
                                        uint32_t sizeY, uint32_t sizeZ, float quantizationStep, uint32_t rowPitchSrc,
                                        uint32_t slicePitchSrc, int quantType, void* stream);
CC_API cc_status cc_quantize_to_symbols_roundtrip(void* d_symbols, int symbolSize, float* d_data, uint32_t sizeX,
                                                  uint32_t sizeY, uint32_t sizeZ, float quantizationStep,
                                                  uint32_t rowPitchSrc, uint32_t slicePitchSrc, int quantType,
                                                  void* stream);
CC_API cc_status cc_unquantize_from_symbols(float* d_data, const void* d_symbols, int symbolSize, uint32_t sizeX,
                                            uint32_t sizeY, uint32_t sizeZ, float quantizationStep,
                                            uint32_t rowPitchDst, uint32_t slicePitchDst, int quantType, void* stream);
/* quantizeToShort / unquantizeFromShort, util/Quantize.h:54-55 */
CC_API cc_status cc_quantize_to_short(int16_t* d_quant, const float* d_data, uint32_t sizeX, uint32_t sizeY,
                                      uint32_t sizeZ, float quantizationStep, uint32_t rowPitchSrc,
                                      uint32_t slicePitchSrc, int quantType, void* stream);
CC_API cc_status cc_unquantize_from_short(float* d_data, const int16_t* d_quant, uint32_t sizeX, uint32_t sizeY,
                                          uint32_t sizeZ, float quantizationStep, uint32_t rowPitchDst,
                                          uint32_t slicePitchDst, int quantType, void* stream);
/* getMaxAbs, util/Quantize.h:64 (result written to device memory, instance stream) */
CC_API cc_status cc_get_max_abs(cc_instance* inst, const float* d_image, uint32_t elemCount, float* d_valMax);

/* ---- Volume codec callers: src/examples/CompressVolume.cpp ------------------------------- */

/* compressVolumeFloat, src/examples/CompressVolume.cpp:313-409 (Symbol16 when symbolSize == 2,
 * Symbol32 when 4).  Output held by the instance like cc_encode_*. */
CC_API cc_status cc_compress_volume_float(cc_instance* inst, const float* d_image, uint32_t sizeX, uint32_t sizeY,
                                          uint32_t sizeZ, uint32_t numLevels, float quantStepLevel0,
                                          int doRLEOnlyOnLvl0, int symbolSize, const uint32_t** outWords,
                                          uint64_t* outBits);
/* decompressVolumeFloat, src/examples/CompressVolume.cpp:412-516 */
CC_API cc_status cc_decompress_volume_float(cc_instance* inst, float* d_image, uint32_t sizeX, uint32_t sizeY,
                                            uint32_t sizeZ, uint32_t numLevels, const uint32_t* bits,
                                            uint64_t bitCount, float quantStepLevel0, int doRLEOnlyOnLvl0,
                                            int symbolSize);
/* compressVolumeLossless, src/examples/CompressVolume.cpp:234-277.  sizeZ == 1 uses the
 * heightfield's three-band layout per level (src/examples/CompressHeightfield.cpp:116-126). */
CC_API cc_status cc_compress_volume_lossless(cc_instance* inst, const int16_t* d_image, uint32_t sizeX,
                                             uint32_t sizeY, uint32_t sizeZ, uint32_t numLevels,
                                             const uint32_t** outWords, uint64_t* outBits);
/* decompressVolumeLossless, src/examples/CompressVolume.cpp:279-309 */
CC_API cc_status cc_decompress_volume_lossless(cc_instance* inst, int16_t* d_image, uint32_t sizeX, uint32_t sizeY,
                                               uint32_t sizeZ, uint32_t numLevels, const uint32_t* bits,
                                               uint64_t bitCount);

/* compressVolumeFloatQuantFirst, src/examples/CompressVolume.cpp:651-757: uniform quantization to
 * int16, numLevels 5/3 levels, Symbol16 streams (log2 <= 16).  Output held by the instance. */
CC_API cc_status cc_compress_volume_float_quant_first(cc_instance* inst, const float* d_image, uint32_t sizeX,
                                                      uint32_t sizeY, uint32_t sizeZ, uint32_t numLevels,
                                                      float quantizationStep, int doRLEOnlyOnLvl0,
                                                      const uint32_t** outWords, uint64_t* outBits);
/* decompressVolumeFloatQuantFirst, src/examples/CompressVolume.cpp:759-856 (the inverse transforms
 * start from the buffer holding the coarsest level for any level count; the reference's choice,
 * buffer numLevels % 2, is that buffer only for an even count). */
CC_API cc_status cc_decompress_volume_float_quant_first(cc_instance* inst, float* d_image, uint32_t sizeX,
                                                        uint32_t sizeY, uint32_t sizeZ, uint32_t numLevels,
                                                        const uint32_t* bits, uint64_t bitCount,
                                                        float quantizationStep, int doRLEOnlyOnLvl0);
/* One channel of a multi-channel decode: VolumeChannel, src/examples/CompressVolume.h:36-42
 * (bitCount widened to 64 bits).  quantStep is quantizationStepLevel0 for the float codec and the
 * uniform step for QuantFirst. */
typedef struct cc_volume_channel {
    float* d_image;        /* device, dense sizeX*sizeY*sizeZ output */
    const uint32_t* bits;  /* host bitstream of this channel */
    uint64_t bitCount;
    float quantStep;
} cc_volume_channel;
/* decompressVolumeFloatMultiChannel, src/examples/CompressVolume.cpp:518-648: one decode call per
 * level over every channel's streams; channel c's image equals cc_decompress_volume_float of its
 * bitstream alone. */
CC_API cc_status cc_decompress_volume_float_multichannel(cc_instance* inst, const cc_volume_channel* channels,
                                                         uint32_t channelCount, uint32_t sizeX, uint32_t sizeY,
                                                         uint32_t sizeZ, uint32_t numLevels, int doRLEOnlyOnLvl0,
                                                         int symbolSize);
/* decompressVolumeFloatQuantFirstMultiChannel, src/examples/CompressVolume.cpp:863-985; channel c's
 * image equals cc_decompress_volume_float_quant_first of its bitstream alone (each channel's
 * inverse transforms start from the buffer holding its coarsest level). */
CC_API cc_status cc_decompress_volume_float_quant_first_multichannel(cc_instance* inst,
                                                                     const cc_volume_channel* channels,
                                                                     uint32_t channelCount, uint32_t sizeX,
                                                                     uint32_t sizeY, uint32_t sizeZ,
                                                                     uint32_t numLevels, int doRLEOnlyOnLvl0);

/* ---- Brick batches ------------------------------------------------------------------------ */
/* No single reference counterpart: a batch is compressVolumeFloat / decompressVolumeFloat
 * (src/examples/CompressVolume.cpp:313-516, doRLEOnlyOnLvl0 = false) applied to many bricks of
 * different sizes in one pipeline of launches (DWT passes over all bricks at once, quantization
 * fused into the last pass, the device-resident coder for all bricks' streams).  Brick i's
 * bitstream is word for word what cc_compress_volume_float returns for that brick alone. */
typedef struct cc_brick {
    const float* src;   /* device, dense sizeX*sizeY*sizeZ (compress input) */
    float* dst;         /* device, dense sizeX*sizeY*sizeZ (decompress output) */
    uint32_t sizeX, sizeY, sizeZ;  /* each divisible by 2^numLevels; sizeZ > 1 */
    float quantStep;    /* level-0 quantization step */
} cc_brick;

/* Compresses `count` bricks into d_out (device, outCapacityWords words): brick i occupies words
 * [brickWordOffsets[i], brickWordOffsets[i+1]) (host array of count+1 entries).  brickStatus
 * (host, count entries, may be NULL) receives each brick's status; the return value is CC_OK only
 * when every brick succeeded.  Synchronises the instance stream once, at the end. */
CC_API cc_status cc_compress_bricks_float(cc_instance* inst, const cc_brick* bricks, uint32_t count,
                                          uint32_t numLevels, int symbolSize, uint32_t* d_out,
                                          uint64_t outCapacityWords, uint64_t* brickWordOffsets,
                                          uint32_t* brickStatus);
/* Decompresses the bricks written by cc_compress_bricks_float (same bricks, levels and steps)
 * from d_in into each brick's dst. */
CC_API cc_status cc_decompress_bricks_float(cc_instance* inst, const cc_brick* bricks, uint32_t count,
                                            uint32_t numLevels, int symbolSize, const uint32_t* d_in,
                                            const uint64_t* brickWordOffsets, uint32_t* brickStatus);

/* Plans: the same batch captured once into two hipGraphs (compression: every DWT pass and coder
 * kernel of all bricks; decompression: the coder's decode kernels, the inverse passes, and the
 * per-brick bitstream positions taken on the device from the last compression) and replayed.
 * The bricks' src/dst pointers, sizes, steps, numLevels and d_out are fixed at creation; the data
 * behind them may change between replays.  Creation runs one compression eagerly to size the
 * coder's arena (with headroom); a replay whose data needs more arena fails with
 * CC_ERR_OUT_OF_MEMORY and the plan must be recreated.  No reference counterpart: the reference
 * has no batched or graph-captured path. */
typedef struct cc_bricks_plan cc_bricks_plan;
CC_API cc_status cc_bricks_plan_create(cc_instance* inst, const cc_brick* bricks, uint32_t count, uint32_t numLevels,
                                       int symbolSize, uint32_t* d_out, uint64_t outCapacityWords,
                                       cc_bricks_plan** plan);
/* Replays the compression graph; fills the host arrays like cc_compress_bricks_float. */
CC_API cc_status cc_bricks_plan_compress(cc_bricks_plan* plan, uint64_t* brickWordOffsets, uint32_t* brickStatus);
/* Replays the decompression graph (bitstreams of the plan's last compression) into the dst bricks. */
CC_API cc_status cc_bricks_plan_decompress(cc_bricks_plan* plan, uint32_t* brickStatus);
/* Replays compression and then decompression with one host synchronisation at the end (the
 * decompression takes the stream positions from device memory, so nothing waits for the host in
 * between); fills brickWordOffsets / compressStatus like cc_bricks_plan_compress and
 * decompressStatus like cc_bricks_plan_decompress.  A compression error is reported first. */
CC_API cc_status cc_bricks_plan_run(cc_bricks_plan* plan, uint64_t* brickWordOffsets, uint32_t* compressStatus,
                                    uint32_t* decompressStatus);
/* cc_bricks_plan_run split in two: _launch enqueues compression and decompression on the plan's
 * stream and returns at once; _wait synchronises that stream and fills the results of the last
 * launch.  Plans have their own streams and buffers, so launches of different plans run
 * concurrently (one plan's serial table designs overlap another's streaming passes). */
CC_API cc_status cc_bricks_plan_launch(cc_bricks_plan* plan);
CC_API cc_status cc_bricks_plan_wait(cc_bricks_plan* plan, uint64_t* brickWordOffsets, uint32_t* compressStatus,
                                     uint32_t* decompressStatus);
CC_API void cc_bricks_plan_destroy(cc_bricks_plan* plan);
/* Counts of the plan's last compression (no reference counterpart), from which the algorithmic
 * bytes of every stage follow: stats[0] symbols coded (N), stats[1] zero-run pairs of the
 * run-length streams (M; each pair is one Huffman symbol and one zero count), stats[2] symbols of
 * the Huffman-only streams (N_H), stats[3] compressed words (C / 4).  Synchronises the plan. */
CC_API cc_status cc_bricks_plan_coding_stats(cc_bricks_plan* plan, uint64_t* stats);

/* ---- Decode serving ----------------------------------------------------------------------- */
/* A request loop over compressed bricks resident in device memory (no reference counterpart).
 * Brick i is the bitstream d_words[brickWordRanges[2i], brickWordRanges[2i+1]) of a brick of
 * brickSizes[3i..3i+2] (x, y, z) compressed with quantSteps[i] and numLevels (as written by
 * cc_compress_bricks_float).  Requests name bricks and caller-owned device output bricks;
 * cc_server_poll admits pending bricks in arrival order onto free slots -- one batch per slot, up
 * to workspaceBytes / slots of decode workspace (at least one brick) -- asynchronously on that
 * slot's stream, and records a request as completed at time `now` (the caller's clock) once the
 * batches holding all of its bricks have finished.  Batches of different slots run concurrently;
 * a request's bricks may be spread over several.  Requests decode to the same values whatever
 * they are batched with.  A server is driven from one host thread. */
typedef struct cc_server cc_server;
CC_API cc_status cc_server_create(cc_instance* inst, const uint32_t* d_words, uint32_t brickCount,
                                  const uint64_t* brickWordRanges, const uint32_t* brickSizes,
                                  const float* quantSteps, uint32_t numLevels, int symbolSize, uint32_t slots,
                                  uint64_t workspaceBytes, cc_server** out);
CC_API cc_status cc_server_submit(cc_server* server, uint64_t requestId, const uint32_t* brickIds, uint32_t count,
                                  float* const* d_outputs, double arrivalTime);
CC_API cc_status cc_server_poll(cc_server* server, double now, uint32_t* batchesLaunched);
/* batches in flight plus requests not yet finished */
CC_API uint32_t cc_server_pending(const cc_server* server);
/* moves up to `max` completed requests (id, completion time) out; returns how many */
CC_API uint32_t cc_server_completed(cc_server* server, uint64_t* requestIds, double* times, uint32_t max);
/* moves up to `max` failed requests (id, decoder status code) out */
CC_API uint32_t cc_server_failed(cc_server* server, uint64_t* requestIds, uint32_t* status, uint32_t max);
/* Drives an idle server through a request trace in real time: request r (id r) arrives
 * arrivals[r] seconds after the call starts (non-decreasing) and asks for counts[r] bricks,
 * brickIds / d_outputs holding the requests' lists one after the other.  The call submits each
 * request at its arrival time, polls continuously and returns when every request has finished:
 * latencies[r] = completion - arrival in seconds (-1 for a failed request), *elapsed = time of
 * the last completion. */
CC_API cc_status cc_server_run(cc_server* server, uint32_t requestCount, const double* arrivals, const uint32_t* counts,
                               const uint32_t* brickIds, float* const* d_outputs, double* latencies, double* elapsed);
CC_API void cc_server_destroy(cc_server* server);

/* ---- Tuning (no reference counterpart) ---------------------------------------------------- */
/* Kernel-shape parameters of the brick batches and the device coder, chosen by the project's
 * autotuner (cudacompress_amd/autotune.py) and kept in a tuning table.  They replace fixed
 * choices of the reference (the 32x4-thread DWT blocks of src/cudaCompress/util/DWTFloat.cu:92-100,
 * the default coding block size of src/cudaCompress/InstanceImpl.cpp:65-69).  None of them changes
 * a bitstream except codingBlockSize, which the table only recommends: an instance's coding block
 * size is fixed at creation. */
typedef struct cc_tuning {
    uint32_t designGrid[4];   /* resident workgroups of the LDS table-design classes (tables of at most
                                 256, 1024, 2048, 4096 distinct symbols) */
    uint32_t fusedXY;         /* 1: brick levels whose rows have 8..512 samples use the fused x+y
                                 passes; 0: separate x and y passes everywhere */
    uint32_t xySlicesMax;     /* most z-slices one fused x+y workgroup serves (0: as many as its
                                 threads cover, one thread per x pair) */
    uint32_t codingBlockSize; /* recommended coding block size (32..256) */
    uint32_t laneDesignK;     /* tables of at most this many distinct symbols (0, 32, 64, 128 or 256;
                                 0: none) are designed one per thread, the others one per wave */
} cc_tuning;
/* The built-in defaults. */
CC_API cc_status cc_default_tuning(cc_tuning* out);
/* Parses a tuning table ("key value..." lines: design_grid g0 g1 g2 g3, fused_xy b, xy_slices_max k,
 * coding_block_size c, lane_design_k k; '#' comments).  path NULL: the library's table, $CUDACOMPRESS_AMD_TUNING or
 * else tuning/gfx950.txt beside the library's lib/ directory; keys it lacks keep their defaults, and
 * a missing file gives the defaults with CC_OK. */
CC_API cc_status cc_load_tuning(const char* path, cc_tuning* out);
/* An instance starts with the library's table; brick batches and plans created afterwards use the
 * instance's tuning.  NULL restores the table. */
CC_API cc_status cc_set_instance_tuning(cc_instance* inst, const cc_tuning* tuning);
CC_API cc_status cc_get_instance_tuning(const cc_instance* inst, cc_tuning* out);

/* ---- Kernel timing ------------------------------------------------------------------------ */
/* No reference counterpart (the reference times host-side sections with CUDA events,
 * src/cudaCompress/util/CudaTimer.cpp:65-169).  When enabled, every kernel launch of this library
 * is bracketed by hipEvents on its stream; collect() waits for the recorded events and aggregates
 * per kernel name: launch count, total milliseconds and algorithmic bytes (one read of each input
 * plus one write of each output, summed over launches). */
CC_API void cc_profile_enable(int enable);
CC_API void cc_profile_reset(void);
CC_API uint32_t cc_profile_collect(void);
CC_API cc_status cc_profile_entry(uint32_t index, const char** name, uint64_t* launches, double* totalMs,
                                  uint64_t* algorithmicBytes);

#ifdef __cplusplus
}
#endif

#endif
