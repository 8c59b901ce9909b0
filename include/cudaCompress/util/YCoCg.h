// Drop-in replacement of include/cudaCompress/util/YCoCg.h (m0bl0/cudaCompress): header-only
// wrappers over cc_convert_rgb_to_ycocg / cc_convert_ycocg_to_rgb (include/cudacompress_amd.h),
// on the null stream as in the reference.
#ifndef CUDACOMPRESS_AMD_DROPIN_YCOCG_H
#define CUDACOMPRESS_AMD_DROPIN_YCOCG_H

#include <cudaCompress/global.h>
#include <cudacompress_amd.h>

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

namespace cudaCompress {
namespace util {

inline void convertRGBtoYCoCg(uchar3* dpTarget, const uchar3* dpData, int pixelCount) {
    cc_convert_rgb_to_ycocg((uint8_t*)dpTarget, (const uint8_t*)dpData, pixelCount, 3, nullptr);
}
inline void convertYCoCgtoRGB(uchar3* dpTarget, const uchar3* dpData, int pixelCount) {
    cc_convert_ycocg_to_rgb((uint8_t*)dpTarget, (const uint8_t*)dpData, pixelCount, 3, nullptr);
}
inline void convertRGBtoYCoCg(uchar4* dpTarget, const uchar4* dpData, int pixelCount) {
    cc_convert_rgb_to_ycocg((uint8_t*)dpTarget, (const uint8_t*)dpData, pixelCount, 4, nullptr);
}
inline void convertYCoCgtoRGB(uchar4* dpTarget, const uchar4* dpData, int pixelCount) {
    cc_convert_ycocg_to_rgb((uint8_t*)dpTarget, (const uint8_t*)dpData, pixelCount, 4, nullptr);
}

}  // namespace util
}  // namespace cudaCompress

#endif
