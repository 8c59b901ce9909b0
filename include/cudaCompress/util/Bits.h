// Drop-in replacement of include/cudaCompress/util/Bits.h (m0bl0/cudaCompress): header-only
// wrappers over cc_get_lsb / cc_remove_lsb / cc_append_lsb / cc_shift_elems
// (include/cudacompress_amd.h), on the null stream as in the reference.  Any elemCount works
// (the reference needs multiples of 4096); dpBits holds ceil(elemCount / 32) words.
#ifndef CUDACOMPRESS_AMD_DROPIN_BITS_H
#define CUDACOMPRESS_AMD_DROPIN_BITS_H

#include <cudaCompress/global.h>
#include <cudacompress_amd.h>

namespace cudaCompress {
namespace util {

// least significant bit of each element into dpBits
inline bool getLSB(const short* dpData, uint* dpBits, uint elemCount) {
    return cc_get_lsb(dpData, dpBits, elemCount, nullptr) == CC_OK;
}
// v += shift, bit into dpBits (if non-null), v = floor(v / 2)
inline bool removeLSB(short* dpData, uint* dpBits, uint elemCount, short shift = 0) {
    return cc_remove_lsb(dpData, dpBits, elemCount, shift, nullptr) == CC_OK;
}
// v = 2 * v + bit, then v += shift
inline bool appendLSB(short* dpData, const uint* dpBits, uint elemCount, short shift = 0) {
    return cc_append_lsb(dpData, dpBits, elemCount, shift, nullptr) == CC_OK;
}
inline bool shiftElems(short* dpData, uint elemCount, short shift) {
    return cc_shift_elems(dpData, elemCount, shift, nullptr) == CC_OK;
}

}  // namespace util
}  // namespace cudaCompress

#endif
