// Drop-in replacement of include/cudaCompress/Encode.h (m0bl0/cudaCompress): the sixteen
// encode/decode overloads as header-only wrappers over cc_encode_* / cc_decode_*.
// Relaxed caller obligations (Encode.h:21-25 of the reference): decodeRLHuff outputs need not be
// zeroed, decodeHuff outputs need no padding, encodeHuff does not modify its inputs.
// Return value: false on symbol overflow, as the reference; also false on any other error
// (cc_last_error() has the reason).
#ifndef CUDACOMPRESS_AMD_DROPIN_ENCODE_H
#define CUDACOMPRESS_AMD_DROPIN_ENCODE_H

#include <cudaCompress/BitStream.h>
#include <cudaCompress/EncodeCommon.h>
#include <cudaCompress/Instance.h>
#include <cudaCompress/global.h>

#include <vector>

namespace cudaCompress {

namespace detail {

typedef cc_status (*EncodeFn)(cc_instance*, int, const void* const*, uint32_t, uint32_t, const uint32_t*,
                              const uint32_t**, const uint64_t**, const uint64_t**);
typedef cc_status (*DecodeFn)(cc_instance*, int, const uint32_t* const*, const uint64_t*, uint64_t*, uint32_t,
                              const uint32_t*, void* const*, uint32_t, uint32_t);

inline bool encodeImpl(EncodeFn fn, Instance* inst, BitStream* const* bs, bool single, int symSize,
                       const void* const* streams, uint streamCount, uint count) {
    std::vector<uint32_t> start(streamCount, 0u);
    std::vector<const BitStream*> seen;
    for (uint i = 0; i < streamCount; i++) {
        const BitStream* b = bs[single ? 0 : i];
        bool first = true;
        for (const BitStream* s : seen)
            if (s == b) first = false;
        if (first) {
            seen.push_back(b);
            start[i] = b->getBitPosition() % 32;
        }
    }
    const uint32_t* words = nullptr;
    const uint64_t* begin = nullptr;
    const uint64_t* bits = nullptr;
    if (fn(raw(inst), symSize, streams, streamCount, count, start.data(), &words, &begin, &bits) != CC_OK) return false;
    for (uint i = 0; i < streamCount; i++)
        bs[single ? 0 : i]->appendEncodedSegment(words + begin[i], bits[i], start[i]);
    return true;
}

inline bool decodeImpl(DecodeFn fn, Instance* inst, BitStreamReadOnly* const* bs, bool single, int symSize,
                       void* const* streams, uint streamCount, uint count) {
    std::vector<BitStreamReadOnly*> uniq;
    std::vector<uint32_t> map(streamCount, 0u);
    for (uint i = 0; i < (single ? 1u : streamCount); i++) {
        BitStreamReadOnly* b = bs[i];
        uint k = 0;
        while (k < uniq.size() && uniq[k] != b) k++;
        if (k == uniq.size()) uniq.push_back(b);
        map[i] = k;
    }
    std::vector<const uint32_t*> words(uniq.size());
    std::vector<uint64_t> size(uniq.size()), pos(uniq.size());
    for (size_t k = 0; k < uniq.size(); k++) {
        words[k] = uniq[k]->getRaw();
        size[k] = uniq[k]->getBitSize();
        pos[k] = uniq[k]->getBitPosition();
    }
    if (fn(raw(inst), symSize, words.data(), size.data(), pos.data(), (uint32_t)uniq.size(),
           single ? nullptr : map.data(), streams, streamCount, count) != CC_OK)
        return false;
    for (size_t k = 0; k < uniq.size(); k++) uniq[k]->setBitPosition((uint)pos[k]);
    return true;
}

}  // namespace detail

// single bitstream for all blocks (Encode.h:28-39)
inline bool encodeRLHuff(Instance* p, BitStream& b, const Symbol16* const s[], uint n, uint c) {
    BitStream* pb = &b;
    return detail::encodeImpl(cc_encode_rlhuff, p, &pb, true, 2, (const void* const*)s, n, c);
}
inline bool decodeRLHuff(Instance* p, BitStreamReadOnly& b, Symbol16* s[], uint n, uint c) {
    BitStreamReadOnly* pb = &b;
    return detail::decodeImpl(cc_decode_rlhuff, p, &pb, true, 2, (void* const*)s, n, c);
}
inline bool encodeHuff(Instance* p, BitStream& b, Symbol16* const s[], uint n, uint c) {
    BitStream* pb = &b;
    return detail::encodeImpl(cc_encode_huff, p, &pb, true, 2, (const void* const*)s, n, c);
}
inline bool decodeHuff(Instance* p, BitStreamReadOnly& b, Symbol16* s[], uint n, uint c) {
    BitStreamReadOnly* pb = &b;
    return detail::decodeImpl(cc_decode_huff, p, &pb, true, 2, (void* const*)s, n, c);
}
inline bool encodeRLHuff(Instance* p, BitStream& b, const Symbol32* const s[], uint n, uint c) {
    BitStream* pb = &b;
    return detail::encodeImpl(cc_encode_rlhuff, p, &pb, true, 4, (const void* const*)s, n, c);
}
inline bool decodeRLHuff(Instance* p, BitStreamReadOnly& b, Symbol32* s[], uint n, uint c) {
    BitStreamReadOnly* pb = &b;
    return detail::decodeImpl(cc_decode_rlhuff, p, &pb, true, 4, (void* const*)s, n, c);
}
inline bool encodeHuff(Instance* p, BitStream& b, Symbol32* const s[], uint n, uint c) {
    BitStream* pb = &b;
    return detail::encodeImpl(cc_encode_huff, p, &pb, true, 4, (const void* const*)s, n, c);
}
inline bool decodeHuff(Instance* p, BitStreamReadOnly& b, Symbol32* s[], uint n, uint c) {
    BitStreamReadOnly* pb = &b;
    return detail::decodeImpl(cc_decode_huff, p, &pb, true, 4, (void* const*)s, n, c);
}

// separate bitstream for each block, possibly repeated (Encode.h:42-53)
inline bool encodeRLHuff(Instance* p, BitStream* b[], const Symbol16* const s[], uint n, uint c) {
    return detail::encodeImpl(cc_encode_rlhuff, p, b, false, 2, (const void* const*)s, n, c);
}
inline bool decodeRLHuff(Instance* p, BitStreamReadOnly* b[], Symbol16* s[], uint n, uint c) {
    return detail::decodeImpl(cc_decode_rlhuff, p, b, false, 2, (void* const*)s, n, c);
}
inline bool encodeHuff(Instance* p, BitStream* b[], Symbol16* const s[], uint n, uint c) {
    return detail::encodeImpl(cc_encode_huff, p, b, false, 2, (const void* const*)s, n, c);
}
inline bool decodeHuff(Instance* p, BitStreamReadOnly* b[], Symbol16* s[], uint n, uint c) {
    return detail::decodeImpl(cc_decode_huff, p, b, false, 2, (void* const*)s, n, c);
}
inline bool encodeRLHuff(Instance* p, BitStream* b[], const Symbol32* const s[], uint n, uint c) {
    return detail::encodeImpl(cc_encode_rlhuff, p, b, false, 4, (const void* const*)s, n, c);
}
inline bool decodeRLHuff(Instance* p, BitStreamReadOnly* b[], Symbol32* s[], uint n, uint c) {
    return detail::decodeImpl(cc_decode_rlhuff, p, b, false, 4, (void* const*)s, n, c);
}
inline bool encodeHuff(Instance* p, BitStream* b[], Symbol32* const s[], uint n, uint c) {
    return detail::encodeImpl(cc_encode_huff, p, b, false, 4, (const void* const*)s, n, c);
}
inline bool decodeHuff(Instance* p, BitStreamReadOnly* b[], Symbol32* s[], uint n, uint c) {
    return detail::decodeImpl(cc_decode_huff, p, b, false, 4, (void* const*)s, n, c);
}

}  // namespace cudaCompress

#endif
