// Entropy-coder orchestration (host side of encodeRLHuff/encodeHuff/decodeRLHuff/decodeHuff).
#pragma once

#include "instance.hpp"

namespace cca {

// Encodes into inst->out_words / seg_begin / seg_bits.  Throws cca::Error.
void encode_streams(cc_instance* inst, bool rl, int symSize, const void* const* streams, uint32_t ns, uint32_t n,
                    const uint32_t* startBits);
// Decodes; positions advanced.  Work is left enqueued on inst->stream.  Throws cca::Error.
void decode_streams(cc_instance* inst, bool rl, int symSize, const uint32_t* const* bsWords, const uint64_t* bsBits,
                    uint64_t* bsPos, uint32_t nbs, const uint32_t* s2b, void* const* outs, uint32_t ns, uint32_t n);

}  // namespace cca
