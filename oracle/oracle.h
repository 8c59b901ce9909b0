/*
 * CPU oracle for the cudaCompress wavelet-compression path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py may load it; the product
 * library (cudacompress_amd) never links, loads or calls it.
 *
 * Every function restates the reference algorithm cited next to it
 * (paths relative to the reference repository m0bl0/cudaCompress).
 * Parity status of each part is recorded in oracle.cpp's header comment.
 */
#ifndef CUDACOMPRESS_ORACLE_H
#define CUDACOMPRESS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_OK = 0,
    ORC_ERR_SYMBOL_OVERFLOW = 1,
    ORC_ERR_INVALID = 3,
    ORC_ERR_CODEWORD_TOO_LONG = 5,
    ORC_ERR_CORRUPT = 6
};

/* GPU dialect: src/cudaCompress/HuffmanTable.cu:470-485 (header), :507 (make_heap tree).
 * CPU dialect: src/CPU/HuffmanTableCPU.cpp:60-62 (push tree), :108-135 (14-bit header). */
enum { ORC_DIALECT_GPU = 0, ORC_DIALECT_CPU = 1 };

typedef struct orc_params {
    int dialect;      /* ORC_DIALECT_GPU or ORC_DIALECT_CPU */
    int log2;         /* log2HuffmanDistinctSymbolCountMax (ignored for the CPU dialect: 14) */
    int cbs;          /* codingBlockSize */
    int symbol_size;  /* 2 (Symbol16) or 4 (Symbol32) */
} orc_params;

void orc_free(void* p);

/* ---- entropy coder ---------------------------------------------------- */

/* Zero-run RLE, src/cudaCompress/RunLengthKernels.cui:5-44 / src/CPU/RunLengthCPU.cpp:8-32.
 * out_syms/out_zc need room for n entries.  Returns the pair count in *out_m. */
int orc_rle_encode(int symbol_size, const void* syms, uint32_t n, void* out_syms, void* out_zc, uint32_t* out_m);
/* src/cudaCompress/RunLengthKernels.cui:57-76 (out must be zeroed by the caller). */
int orc_rle_decode(int symbol_size, const void* syms, const void* zc, uint32_t m, void* out, uint32_t n);

/* Exact histogram as the reference GPU path counts it, including the pad word
 * (src/cudaCompress/Histogram.cu:90-105, HistogramKernels.cui large kernels).
 * call_symbol_max is the max symbol over all streams of the design call. */
void orc_histogram(int symbol_size, const void* syms, uint32_t n, uint32_t call_symbol_max, uint32_t* hist, uint32_t nbins);

/* Huffman table design for one stream, src/cudaCompress/HuffmanTable.cu:487-586.
 * hist has stream_symbol_max+1 bins.  Outputs:
 *   lengths/codewords indexed by symbol (length 0xFFFFFFFF = unused),
 *   count_per_length[32] (entries [0, *num_lengths) valid),
 *   symbols in codeword-index order (*nsyms entries). */
int orc_huff_design(int dialect, const uint32_t* hist, uint32_t stream_symbol_max,
                    uint32_t* lengths, uint32_t* codewords,
                    uint32_t* num_lengths, uint32_t* count_per_length,
                    uint32_t* symbols, uint32_t* nsyms);

/* Full bitstream encode, src/cudaCompress/Encode.cpp:208-471 (rl=1) and :649-790 (rl=0);
 * CPU dialect src/CPU/EncodeCPU.cpp:98-197, :275-320.  Streams are written back to back
 * starting at bit start_bit (0..31) of word 0.  *out_words is malloc'd (orc_free). */
int orc_encode(int rl, const orc_params* p, const void* const* streams, uint32_t nstreams, uint32_t count,
               uint32_t start_bit, uint32_t** out_words, uint64_t* out_words_len, uint64_t* out_bits);
/* Decode, src/cudaCompress/Encode.cpp:474-646 (rl=1), :793-917 (rl=0).
 * Reads from *bitpos and advances it.  Outputs are fully written (RLE outputs zero-filled). */
int orc_decode(int rl, const orc_params* p, const uint32_t* words, uint64_t nbits, uint64_t* bitpos,
               void* const* out_streams, uint32_t nstreams, uint32_t count);

/* packInc16CPU / unpackInc16CPU, src/cudaCompress/PackInc.cu:210-229. */
void orc_pack_inc16(const uint32_t* values, uint16_t* incs, uint32_t n);
void orc_unpack_inc16(const uint16_t* incs, uint32_t* values, uint32_t n);

/* ---- transforms -------------------------------------------------------- */

/* One DWT level in place on a dense sx*sy*sz array (sz==1: 2D), X->Y->Z forward,
 * Z->Y->X inverse.  CDF 9/7 math: src/cudaCompress/util/DWTFloatKernels.cui:14-81,433-503,
 * filters src/cudaCompress/util/DWTFloatFilters.cuh:42-49.  fp32 with fused multiply-add
 * accumulation in filter order; the f64 twin uses the same decimal coefficients. */
void orc_dwt97_forward_f32(float* data, int sx, int sy, int sz);
void orc_dwt97_inverse_f32(float* data, int sx, int sy, int sz);
void orc_dwt97_forward_f64(double* data, int sx, int sy, int sz);
void orc_dwt97_inverse_f64(double* data, int sx, int sy, int sz);

/* CDF 5/3 integer lifting on int16 with wrap, src/cudaCompress/util/DWTIntKernels.cui:15-100,597-687,
 * src/cudaCompress/util/DWTCommon.cuh:133-142. */
void orc_dwt53_forward(int16_t* data, int sx, int sy, int sz);
void orc_dwt53_inverse(int16_t* data, int sx, int sy, int sz);

/* Quantizer, src/cudaCompress/util/Quantize.cuh:55-147 (qtype 0 deadzone, 1 uniform).
 * inv_step is 1.0f/step computed in fp32 (src/cudaCompress/util/Quantize.cu:94). */
void orc_quantize(const float* data, uint32_t n, float step, int qtype, int symbol_size, void* symbols);
void orc_unquantize(const void* symbols, uint32_t n, float step, int qtype, int symbol_size, float* data);
void orc_symbolize(const int16_t* data, uint32_t n, uint16_t* symbols);
/* quantizeToShort / unquantizeFromShort: src/cudaCompress/util/QuantizeKernels.cui:136-172
 * (quantize<Q> / unquantize<Q> of src/cudaCompress/util/Quantize.cuh:76-147 with the int result
 * stored as a short: two's-complement wrap). */
void orc_quantize_to_short(const float* data, uint32_t n, float step, int qtype, int16_t* quant);
void orc_unquantize_from_short(const int16_t* quant, uint32_t n, float step, int qtype, float* data);
void orc_unsymbolize(const uint16_t* symbols, uint32_t n, int16_t* data);

#ifdef __cplusplus
}
#endif

#endif
