// Host side of the device-resident coder (hip/coder.hip, hip/decoder.hip).  A job is staged once
// (prepare: host tables + one pinned upload), launched as kernels only (launch: capturable into a
// hipGraph), and completed by one download + synchronisation (finish).
#pragma once

#include <vector>

#include "../coder.hpp"
#include "instance.hpp"

namespace cca {

// One input stream (encode) or output stream (decode) of a job.  Items are grouped by call, in
// call order; a call is one contiguous bitstream segment (items may differ in length and rl).
struct ItemSpec {
    void* ptr;
    uint32_t n;
    bool rl;
    uint32_t call;
};

// Header-format parameters of the instance's coder (coding block size, alphabet, dialect).
CoderParams params_of(const cc_instance* I);
// Throws CC_ERR_INVALID_ARG when symSize does not fit the instance (Symbol32 iff log2 > 16; the
// CPU dialect is Symbol16 only).
void check_symbols(const cc_instance* I, int symSize);
struct CallWords {
    const uint32_t* words;  // device bitstream of a decode call
    uint64_t nwords;
};

struct EncodeJob {
    uint32_t nitems = 0, nseq = 0, ncalls = 0;
    uint64_t total_syms = 0;
    int symSize = 2;
    CoderParams p{};
    uint64_t scratch_cap = 0;  // arena words
    uint32_t design_grid[4] = {4096, 2048, 1536, 768};  // resident workgroups per design class
    uint32_t lane_k = 0;  // tables of at most this many symbols are designed one per thread (0: none)
};
void encode_prepare(cc_instance* I, DeviceCoder& D, const std::vector<ItemSpec>& items, uint32_t ncalls, int symSize,
                    EncodeJob& job, hipStream_t st);
// Sizes the arena to at least `words` (allocates: not during graph capture).
void encode_reserve(DeviceCoder& D, EncodeJob& job, uint64_t words);
void encode_launch(DeviceCoder& D, const EncodeJob& job, uint32_t* d_out, uint64_t out_cap, hipStream_t st,
                   const LayoutLink* link = nullptr);
CoderState encode_finish(DeviceCoder& D, const EncodeJob& job, std::vector<CallOut>* results, hipStream_t st);

struct DecodeJob {
    uint32_t nitems = 0, nseq = 0, ncalls = 0, nwaves = 0, nlong = 0;
    uint64_t total_syms = 0, arena = 0;
    int symSize = 2;
    CoderParams p{};
    size_t o_items = 0, o_wseq = 0, o_wfirst = 0;
    // zero-run expansion chunks: item_first table (uploaded) and chunk sums (device only)
    uint32_t nchunks = 0;
    size_t o_ifirst = 0, o_sums = 0;
};
// Calls may carry null words when decode_calls_from_encode fills them on the device.
void decode_prepare(cc_instance* I, DeviceCoder& D, const std::vector<ItemSpec>& items,
                    const std::vector<CallWords>& calls, int symSize, DecodeJob& job, hipStream_t st);
void decode_launch(DeviceCoder& D, const DecodeJob& job, hipStream_t st);
void decode_finish(DeviceCoder& D, const DecodeJob& job, std::vector<DecCallOut>* results, hipStream_t st);
// Call i of the decode reads base + enc[i].word_begin, enc[i].words words (one kernel).
void decode_calls_from_encode(DeviceCoder& D, const DecodeJob& job, const CallOut* d_enc, const uint32_t* base,
                              uint64_t cap, hipStream_t st);

// Resident decode index (serving): every call's parsed headers, decode tables and block offsets,
// built once; a bound decode copies the sequences of the calls it serves and decodes them with no
// parse and no table build.  Calls of the index are referred to by number.
struct DecodeIndex {
    DevBuf d_calls, d_seqs, d_scratch, d_out, d_state;
    PinBuf h_stage;
    std::vector<uint32_t> call_first_seq;  // first index sequence of every call
    std::vector<uint32_t> call_nseq;
    std::vector<DecCallOut> results;       // parse status of every call
    uint64_t arena = 0;
    uint32_t nseq = 0;
};
// Builds the index (one synchronisation).  `items` carry shapes only (ptr unused).
void decode_index_build(cc_instance* I, DecodeIndex& X, const std::vector<ItemSpec>& items,
                        const std::vector<CallWords>& calls, int symSize, hipStream_t st);
struct BoundJob {
    uint32_t nitems = 0, nseq = 0, ncalls = 0, nwaves = 0, nlong = 0;
    uint64_t total_syms = 0, rl_bytes = 0;
    int symSize = 2;
    CoderParams p{};
    size_t o_items = 0, o_bind = 0, o_cbrick = 0, o_wseq = 0, o_wfirst = 0;
    uint32_t nchunks = 0;
    size_t o_ifirst = 0, o_sums = 0;
};
// `items` (grouped by call) are the request's outputs; request call c decodes index call
// index_calls[c] and must have that call's item shapes.
void decode_prepare_bound(cc_instance* I, DeviceCoder& D, const DecodeIndex& X, const std::vector<ItemSpec>& items,
                          const std::vector<uint32_t>& index_calls, int symSize, BoundJob& job, hipStream_t st);
void decode_launch_bound(DeviceCoder& D, const DecodeIndex& X, const BoundJob& job, hipStream_t st);

// Eager forms: prepare + launch (+ finish with one synchronisation when `wait`; the encoder
// retries with a larger arena when it ran out).
uint64_t device_encode_items(cc_instance* I, const std::vector<ItemSpec>& items, uint32_t ncalls, int symSize,
                             uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait);
void device_decode_items(cc_instance* I, const std::vector<ItemSpec>& items, const std::vector<CallWords>& calls,
                         int symSize, std::vector<DecCallOut>* results, bool wait);

struct CoderCall {
    bool rl;                       // encodeRLHuff (true) or encodeHuff
    uint32_t first_stream;         // index into the job's stream pointer list
    uint32_t nstreams;
    uint32_t n;                    // symbols per stream
};
uint64_t device_encode(cc_instance* I, const std::vector<CoderCall>& calls, const std::vector<const void*>& streams,
                       int symSize, uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait);

struct DecodeCall {
    const uint32_t* words;         // device bitstream of the call (word aligned)
    uint64_t nwords;               // words readable from `words` on
    bool rl;
    uint32_t first_stream;         // index into the job's output pointer list
    uint32_t nstreams;
    uint32_t n;
};
void device_decode(cc_instance* I, const std::vector<DecodeCall>& calls, const std::vector<void*>& outs, int symSize,
                   std::vector<DecCallOut>* results, bool wait);

// device_decode of one call of nstreams equal streams with no allocation: every buffer beyond
// the bitstream and the outputs is carved from a per-device static array (hip/noalloc.hip).
// Synchronises the instance stream; returns the call's word count and sets *err (DecErr).
uint64_t decode_device_noalloc(cc_instance* I, bool rl, int symSize, const uint32_t* d_words, uint64_t wordCount,
                               void* const* outs, uint32_t nstreams, uint32_t n, uint32_t* err);

}  // namespace cca
