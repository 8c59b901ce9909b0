// Device-resident entropy coder: descriptors shared by the kernels (hip/coder.hip) and the host
// orchestration (host/batch.cpp).
//
// An encode job is a list of coder calls (an encodeRLHuff or encodeHuff call each, src/cudaCompress/
// Encode.cpp:208-471, :649-790); a call has one item per input stream; an item yields one coded
// sequence (Huffman only) or two (zero-run RLE: symbols, then zero counts).  Everything from the
// symbol streams to the final bitstream stays in device memory:
//   stats   one workgroup per item: zero-run RLE computed on the fly, histograms in LDS
//   design  one wave per coded sequence: the table design of huff_design.hpp
//   pack    one workgroup per item: RLE recomputed, codewords packed through an LDS window
//   layout  one workgroup: word offset of every section (sizes from the tables and bit counts)
//   gather  one workgroup per coded sequence: table header, bit count, codewords, offsets
// Intermediate data (histograms, tables, codewords, block offsets) live in a scratch arena
// carved by an atomic bump allocator; the final layout does not depend on allocation order.
#pragma once

#include "common.hpp"

namespace cca {

constexpr uint32_t kStatBins = 4096;    // LDS histogram bins of the streaming passes
constexpr uint32_t kDecLutBits = 10;    // lookup-table bits of the block decoder (codes of up to that length)
constexpr uint32_t kDecGroupBlocks = 256;  // coding blocks per block-decoder workgroup of a long sequence
constexpr uint32_t kDecWaveBlocks = 64;    // a sequence of at most this many blocks: one wave
// Block-decoder entries of a sequence of (at most) nblk coding blocks.
inline bool dec_long(uint32_t nblk) { return nblk > kDecWaveBlocks; }
inline uint32_t dec_entries(uint32_t nblk) {
    return dec_long(nblk) ? (nblk + kDecGroupBlocks - 1) / kDecGroupBlocks : (nblk ? 1u : 0u);
}
constexpr uint32_t kCoderThreads = 256;
constexpr uint32_t kPerThread = 8;
constexpr uint32_t kChunk = kCoderThreads * kPerThread;  // symbols per workgroup step

struct CodeItem {
    const void* in;      // symbols
    uint32_t n;          // symbol count
    uint32_t rl;         // 1: zero-run RLE + two coded sequences
    uint32_t seq;        // index of the first coded sequence
    uint32_t call;       // coder call
};

enum CodeErr : uint32_t {
    kErrNone = 0,
    kErrOverflow = 1,   // symbol >= 2^log2HuffmanDistinctSymbolCountMax (the reference returns false)
    kErrTooLong = 5,    // a codeword longer than 31 bits
    kErrScratch = 7,    // scratch arena exhausted
    kErrOutput = 8      // output buffer too small
};

struct SeqInfo {
    uint32_t count;      // symbols coded (pairs for RLE sequences)
    uint32_t smax;       // largest symbol (0 when count == 0)
    uint32_t nsym;       // distinct symbols
    uint32_t nlen;       // codeword lengths used
    uint32_t err;
    uint32_t quirk;      // histogram pad symbol counted in bin 65535
    uint32_t wide;       // symbols beyond the LDS histogram: global-memory path
    uint32_t item;
    uint64_t bits;       // codeword bits
    uint64_t hist_off;   // scratch word offsets ...
    uint64_t code_off;   // code[smax+1] then len[smax+1]
    uint64_t syms_off;   // symbols in codeword-index order
    uint64_t last_nz1;   // position of the stream's last nonzero + 1 (0: none), set by the stats pass
    uint64_t boff_off;   // absolute bit offset of every coding block
    uint64_t out_off;    // output word offset of the table header
    uint32_t cpl[32];    // codewords per length
};

struct CoderParams {
    uint32_t cbs;
    uint32_t log2;
    uint32_t dialect;       // CC_DIALECT_GPU / CC_DIALECT_CPU
    uint32_t count_bits;    // header field widths (log2 / 14)
    uint32_t symbits_bits;  // 4 or 5 (GPU), 14 (CPU)
    uint32_t pad;
};

struct CoderState {          // device-side counters of one job
    unsigned long long scratch_used;  // words
    unsigned long long out_words;     // total output words (written by layout)
    unsigned long long out_base;      // word offset of the job's first section in the output
    uint32_t err;                     // first error of the job (CodeErr), 0 if none
    uint32_t pad;
};

struct CallOut {              // per coder call
    uint64_t word_begin;
    uint64_t words;
    uint32_t err;
    uint32_t pad;
};

// ---- decoder ---------------------------------------------------------------------------------
// A decode job: calls, each a word-aligned bitstream segment in device memory holding the items
// of one encodeRLHuff/encodeHuff call in order.  parse (one wave per call: the section sizes
// are in-band, so a call is walked sequentially, each section head read as one window) -> tables (one workgroup per coded sequence:
// first/min/max per length, absolute block offsets) -> blocks (four waves per 256 coding blocks,
// one LDS lookup table; the zero counts' chunk spans) -> rle (one workgroup per chunk of 2048
// pairs: zero runs expanded into the output).
struct DecItem {
    void* out;           // n decoded symbols
    uint32_t n;
    uint32_t rl;
    uint32_t seq;        // first coded sequence
    uint32_t call;
};

struct DecCall {
    const uint32_t* words;  // the call's bitstream
    uint64_t nwords;        // words available from `words` on
    uint32_t first_item, nitems;
};

enum DecErr : uint32_t { kDecOk = 0, kDecCorrupt = 6, kDecScratch = 7 };

struct DecSeq {
    const uint32_t* base;   // the call's words
    uint64_t nwords;
    uint64_t hdr_bit;       // table header
    uint64_t sym_bit;       // first symbol field of the header
    uint64_t words_off;     // codewords (word offset from base)
    uint64_t incs_off;      // u16 increments (word offset from base)
    uint64_t boff_off;      // arena: absolute block offsets
    uint64_t lut_off;       // arena: 2^kDecLutBits-entry lookup table (len << 24 | symbol), then the symbol table
    void* out;              // decoded values (arena for RLE sequences, the item output otherwise)
    uint32_t nlen, nsym, sb, count;
    uint32_t nwords_cw;     // codeword words
    uint32_t err;
    uint32_t item;
    uint32_t pad;
    int32_t tab[96];        // first | min | max codeword per length
};

// One sequence of a decode bound to a resident index: the index sequence it copies, the request
// item it belongs to, and the byte offset of its RLE decode buffer (~0: a plain sequence decoding
// straight into the item's output).
struct DecBind {
    uint32_t idx_seq;
    uint32_t item;
    uint64_t rl_off;
};

struct DecCallOut {
    uint64_t words;         // words consumed by the call
    uint32_t err;
    uint32_t pad;
};

// Launchers (hip/coder.hip).  scratch: arena of `scratch_cap` words.
// Side streams for launches that may run concurrently (the table-design classes): `s` are
// kSideStreams non-blocking streams, `fork` / `join` events owned by the caller; a null `side`
// argument puts everything on st.
constexpr int kSideStreams = 4;
// Table-design work lists (device buffer of kDesignListBase + kDesignLists * nseq words): the
// counters of the design classes at [0, kDesignLists), class list c from kDesignListBase + c * nseq.
constexpr uint32_t kDesignLists = 6;   // four LDS classes, the arena class, the per-thread class
constexpr uint32_t kLaneClass = 5;     // tables designed one per thread (on class 0's stream)
constexpr uint32_t kDesignListBase = 8;
static_assert(kDesignLists <= kDesignListBase, "counters before the lists");
struct SideStreams {
    hipStream_t s[kSideStreams];
    hipEvent_t fork;
    hipEvent_t join[kSideStreams];
};
// Jobs writing one output one after the other: a job's sections start where the previous job's
// end (prev: that job's device state, null for the first); the layout waits for `wait` (the
// previous job's layout) and records `signal` after its own.
struct LayoutLink {
    const CoderState* prev;
    hipEvent_t wait;
    hipEvent_t signal;
};
// One separately laid out job of a concatenation: its bitstream words and its state (word count
// out_base + out_words).
struct ConcatGroup {
    const uint32_t* words;
    const CoderState* st;
};
// Copies the jobs' bitstreams back to back into dst (capacity dst_cap words), in group order.
void coder_concat(const ConcatGroup* d_groups, uint32_t ngroups, uint32_t* dst, uint64_t dst_cap, uint64_t bytes,
                  hipStream_t st);
// d_classes: kDesignListBase + kDesignLists * nseq words of work lists.  d_item_order: workgroup
// k of the stats and pack passes (one per item, each a serial loop over its chunks) serves item
// d_item_order[k] (null: item k); largest items first, so the longest loops start first.
void coder_encode(const CodeItem* d_items, uint32_t nitems, SeqInfo* d_seqs, uint32_t nseq, const CoderParams& p,
                  int symSize, uint32_t* scratch, uint64_t scratch_cap, CoderState* d_state, uint32_t* d_classes,
                  CallOut* d_calls,
                  uint32_t ncalls, const uint32_t* d_call_first_item, uint32_t* d_out, uint64_t out_cap,
                  uint64_t total_symbols, hipStream_t st, const SideStreams* side,
                  const LayoutLink* link = nullptr, const uint32_t* design_grid = nullptr, uint32_t lane_k = 0,
                  const uint32_t* d_item_order = nullptr);
// wave_seq: for every block-decoder workgroup the coded sequence it serves; wave_first: that
// workgroup's first block.  The first nlong entries are workgroups of kDecGroupBlocks blocks of
// the long sequences (dec_long of the block-count bound: RLE sequences have at most n pairs), the
// rest one wave for each short sequence (first block 0).
// Zero-run expansion in chunks of kRleChunk pairs, every chunk its own workgroup: item i of a
// decode job has rle_chunks(n_i, rl_i) chunks, numbered from item_first[i] (nitems + 1 entries,
// item_first[nitems] = nchunks); sums receives one word per chunk.
constexpr uint32_t kRleChunk = 2048;
inline uint32_t rle_chunks(uint32_t n, bool rl) { return rl ? (n + kRleChunk - 1) / kRleChunk : 0u; }
struct RleChunks {
    const uint32_t* item_first;
    uint32_t nchunks;
    uint32_t* sums;
};
void coder_decode(const DecCall* d_calls, uint32_t ncalls, const DecItem* d_items, uint32_t nitems, DecSeq* d_seqs,
                  uint32_t nseq, const CoderParams& p, int symSize, uint32_t* scratch, uint64_t scratch_cap,
                  CoderState* d_state, DecCallOut* d_out, const uint32_t* d_wave_seq, const uint32_t* d_wave_first,
                  uint32_t nwaves, uint32_t nlong, uint64_t total_symbols, const RleChunks& rc,
                  hipStream_t st);
// Resident decode index: parse + tables of every call once (RLE sequences get no decode buffer),
// per-call results in d_out.  scratch then holds the tables and block offsets for bound decodes.
void coder_decode_index(const DecCall* d_calls, uint32_t ncalls, const DecItem* d_items, DecSeq* d_seqs, uint32_t nseq,
                        const CoderParams& p, int symSize, uint32_t* scratch, uint64_t scratch_cap, CoderState* d_state,
                        DecCallOut* d_out, hipStream_t st);
// A decode whose sequences come from a resident index (d_bind, one entry per sequence; call c is
// index call d_call_brick[c]): bind, block decode, zero-run expansion, per-call status.
void coder_decode_bound(const DecSeq* d_idx_seqs, const uint32_t* idx_scratch, const DecCallOut* d_idx_out,
                        const DecBind* d_bind, const uint32_t* d_call_brick, const DecCall* d_calls, uint32_t ncalls,
                        const DecItem* d_items, uint32_t nitems, DecSeq* d_seqs, uint32_t nseq, const CoderParams& p,
                        int symSize, unsigned char* rl_arena, CoderState* d_state, DecCallOut* d_out,
                        const uint32_t* d_wave_seq, const uint32_t* d_wave_first, uint32_t nwaves,
                        uint32_t nlong, uint64_t total_symbols, const RleChunks& rc, hipStream_t st);
// calls[i].words = base + enc[i].word_begin, calls[i].nwords = enc[i].words (the rest unchanged);
// an empty range for a failed call or one reaching past cap words
void coder_calls_from_encode(DecCall* d_calls, const CallOut* d_enc, const uint32_t* base, uint64_t cap,
                             uint32_t ncalls, hipStream_t st);

}  // namespace cca
