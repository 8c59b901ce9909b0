// Host side of the device-resident coder (hip/coder.hip): builds the item list of a job of coder
// calls, sizes the scratch arena, launches the pipeline and reports per-call results.
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
struct CallWords {
    const uint32_t* words;  // device bitstream of a decode call
    uint64_t nwords;
};

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

// Encodes `calls` (streams: device pointers) into d_out (capacity out_cap words) without any
// host synchronisation inside the pipeline; when `wait` is set the call synchronises once at the
// end, retries with a larger arena if it ran out, and fills `results` (one per call).
// Returns the total word count (valid when waited).
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

// Decodes `calls` into the device outputs without host synchronisation inside the pipeline;
// with `wait`, synchronises once and fills `results` (words consumed and status per call).
void device_decode(cc_instance* I, const std::vector<DecodeCall>& calls, const std::vector<void*>& outs, int symSize,
                   std::vector<DecCallOut>* results, bool wait);

}  // namespace cca
