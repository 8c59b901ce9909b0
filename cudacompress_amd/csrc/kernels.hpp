// Internal launch interface of the gfx950 kernels (namespace cca).  Every launcher checks its
// arguments on the host before launching and throws cca::Error on failure.
#pragma once

#include "common.hpp"
#include "profile.hpp"

namespace cca {

// ---- DWT (src/cudaCompress/util/DWTFloat.cu, DWTInt.cu) -------------------------------------
// one 9/7 analysis level in FP64 (no reference counterpart; measures the FP32 transform's error)
void dwt97_forward_f64(double* dst, double* buf, const double* src, int sx, int sy, int sz, int row, int slice,
                       hipStream_t st);
void dwt97_forward(float* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int srcCh,
                   int dRow, int dSlice, int sRow, int sSlice, hipStream_t st);
void dwt97_inverse(float* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int dstCh,
                   int dRow, int dSlice, int sRow, int sSlice, hipStream_t st);
void dwt97_inverse_from_symbols(float* dst, float* buf2, float* buf1, const float* lowpass,
                                const void* const* dd_highpass, int symbolSize, float step, int sx, int sy, int sz,
                                int dstCh, int dRow, int dSlice, int lpRow, int lpSlice, hipStream_t st);
void dwt97_forward_from_byte(float* dst, float* buf2, float* buf1, const uint8_t* src, int sx, int sy, int sz,
                             int srcCh, int dRow, int dSlice, int sRow, int sSlice, hipStream_t st);
void dwt97_inverse_to_byte(uint8_t* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int dstCh,
                           int dRow, int dSlice, int sRow, int sSlice, hipStream_t st);
void dwt97_inverse_from_symbols_to_byte(uint8_t* dst, float* buf2, float* buf1, const float* lowpass,
                                        const void* const* dd_highpass, float step, int sx, int sy, int sz, int dstCh,
                                        int dRow, int dSlice, int lpRow, int lpSlice, hipStream_t st);
void dwt97_forward_lowpass_only_from_byte(float* dst, float* buf, const uint8_t* src, int sx, int sy, int srcCh,
                                          int dRow, int sRow, hipStream_t st);
void dwt53_forward(int16_t* dst, int16_t* buf, const int16_t* src, int sx, int sy, int sz, int dRow, int dSlice,
                   int sRow, int sSlice, hipStream_t st);
void dwt53_forward_lowpass_only(int16_t* dst, int16_t* buf, const int16_t* src, int sx, int sy, int dRow, int sRow,
                                hipStream_t st);
void dwt53_inverse(int16_t* dst, int16_t* buf, const int16_t* src, int sx, int sy, int sz, int dRow, int dSlice,
                   int sRow, int sSlice, hipStream_t st);

// ---- quantization (src/cudaCompress/util/Quantize.cu) ---------------------------------------
void symbolize(uint16_t* sym, const int16_t* data, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row,
               uint32_t slice, hipStream_t st);
void unsymbolize(int16_t* data, const uint16_t* sym, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row,
                 uint32_t slice, hipStream_t st);
void quantize_to_symbols(void* sym, int symSize, float* data, bool roundtrip, uint32_t sx, uint32_t sy, uint32_t sz,
                         float step, uint32_t row, uint32_t slice, int qtype, hipStream_t st);
void unquantize_from_symbols(float* data, const void* sym, int symSize, uint32_t sx, uint32_t sy, uint32_t sz,
                             float step, uint32_t row, uint32_t slice, int qtype, hipStream_t st);
void quantize_to_short(int16_t* q, const float* data, uint32_t sx, uint32_t sy, uint32_t sz, float step,
                       uint32_t row, uint32_t slice, int qtype, hipStream_t st);
void unquantize_from_short(float* data, const int16_t* q, uint32_t sx, uint32_t sy, uint32_t sz, float step,
                           uint32_t row, uint32_t slice, int qtype, hipStream_t st);
void max_abs(const float* data, uint32_t n, float* out, hipStream_t st);

// ---- image path and bit planes (src/cudaCompress/util/Quantize.cu, YCoCg.cu, Bits.cu) ---------
void float_to_byte_2d(uint8_t* out, uint32_t channelCount, uint32_t channel, const float* data, uint32_t sx,
                      uint32_t sy, uint32_t row, hipStream_t st);
void byte_to_float_2d(float* data, const uint8_t* in, uint32_t channelCount, uint32_t channel, uint32_t sx,
                      uint32_t sy, uint32_t row, hipStream_t st);
void quantize_difference_2d(uint16_t* sym, float* data, float step, const uint8_t* ref, uint32_t channelCount,
                            uint32_t channel, uint32_t sx, uint32_t sy, uint32_t row, int qtype, bool roundtrip,
                            hipStream_t st);
void unquantize_difference_2d(float* data, const uint16_t* sym, float step, const uint8_t* ref, uint32_t channelCount,
                              uint32_t channel, uint32_t sx, uint32_t sy, uint32_t row, int qtype, hipStream_t st);
void convert_ycocg(uint8_t* out, const uint8_t* in, uint32_t pixels, int channels, bool to_ycocg, hipStream_t st);
void lsb_get(const int16_t* data, uint32_t* bits, uint32_t n, hipStream_t st);
void lsb_remove(int16_t* data, uint32_t* bits, uint32_t n, int16_t shift, hipStream_t st);
void lsb_append(int16_t* data, const uint32_t* bits, uint32_t n, int16_t shift, hipStream_t st);
void shift_elems(int16_t* data, uint32_t n, int16_t shift, hipStream_t st);

// ---- entropy coder kernels -----------------------------------------------------------------
// Per-stream descriptors live in device memory; see codec.cpp for how they are filled.
struct RleStream {
    const void* in;       // n symbols
    void* out_sym;        // compacted symbols (capacity n)
    void* out_zc;         // zero counts (capacity n)
    uint32_t n;
    uint32_t tile_base;   // index of this stream's first tile in the tile arrays
};
struct RleTile {          // per 2048-symbol tile
    uint32_t nnz;
    uint32_t first_pos;   // position of the first nonzero (or 0xFFFFFFFF)
    uint32_t last_pos;    // position of the last nonzero
    uint32_t inner_fill;  // filler pairs emitted before nonzeros 2..nnz of this tile
    uint32_t out_base;    // output index of this tile's first pair (filled by the scan)
    uint32_t pad[3];
};
struct RleResult {        // per stream
    uint32_t m;           // pair count
    uint32_t sym_max;     // max compacted symbol (0 when m == 0)
    uint32_t zc_max;      // max zero count
    uint32_t pad;
};
constexpr int kRleTile = 2048;
void rle_encode(const RleStream* d_streams, uint32_t nstreams, RleTile* d_tiles, uint32_t ntiles,
                const uint32_t* d_tile_stream, RleResult* d_results, int symSize, uint64_t in_elems, hipStream_t st);

struct RleDecStream {
    const void* sym;     // m compacted symbols
    const void* zc;      // m zero counts
    void* out;           // n outputs
    uint32_t m;
    uint32_t n;
    uint32_t tile_base;  // first tile (of kRleTile pairs) of this stream
    uint32_t pad;
};
// Zero-fills every output element, then scatters the pairs.
void rle_decode(const RleDecStream* d_streams, uint32_t nstreams, uint32_t* d_tile_sums, uint32_t ntiles,
                const uint32_t* d_tile_stream, uint32_t nmax_out, int symSize, uint64_t total_pairs, uint64_t total_out,
                hipStream_t st);
// Per-stream max symbol into results[i].sym_max (results[i].m = n) for streams coded without RLE.
void stream_max(const RleStream* d_streams, uint32_t nstreams, uint32_t nmax, RleResult* d_results, int symSize,
                hipStream_t st);

struct HistStream {
    const void* data;
    uint32_t n;
    uint32_t nbins;       // stream symbol max + 1
    uint32_t* hist;       // nbins counters (zeroed by the launcher)
    uint32_t quirk;       // add the Symbol16 pad word to bin 65535 (see codec.cpp)
};
void histogram(const HistStream* d_streams, const HistStream* h_streams, uint32_t nstreams, int symSize,
               hipStream_t st);

struct HuffEncStream {
    const void* sym;       // n symbols
    const uint32_t* code;  // codeword by symbol
    const uint32_t* len;   // length by symbol
    uint32_t* words;       // output words (zeroed by the launcher)
    uint32_t* offsets;     // bit offset of each coding block (ceil(n/cbs))
    uint32_t n;
    uint32_t block_base;   // first coding block in the global block numbering
    uint32_t chunk_base;   // first pack chunk
    uint32_t pad;
};
void huffman_encode(const HuffEncStream* d_streams, const HuffEncStream* h_streams, uint32_t nstreams,
                    uint32_t cbs, int symSize, uint32_t* d_block_bits, uint32_t nblocks,
                    const uint32_t* d_block_stream, uint32_t* d_stream_bits, const uint32_t* d_chunk_stream,
                    uint32_t nchunks, uint64_t total_bits, hipStream_t st);

struct HuffDecStream {
    const uint32_t* words;     // codeword words (+1 word of padding)
    const uint32_t* offsets;   // absolute bit offset of each coding block
    const int32_t* table;      // first[32], min[32], max[32] per codeword length
    const void* symbols;       // symbol table by codeword index
    void* out;                 // n decoded symbols
    uint32_t n;
    uint32_t nsyms;            // symbol table size
    uint32_t wave_base;        // first decode wave of this stream
    uint32_t nwords;           // codeword words (readable: nwords + 2)
};
void huffman_decode(const HuffDecStream* d_streams, uint32_t nstreams, const uint32_t* d_wave_stream,
                    uint32_t nwaves, uint32_t cbs, int symSize, uint64_t total_syms, uint64_t total_bits,
                    hipStream_t st);

}  // namespace cca
