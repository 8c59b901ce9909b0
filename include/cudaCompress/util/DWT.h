// Drop-in replacement of include/cudaCompress/util/DWT.h (m0bl0/cudaCompress): header-only
// wrappers over cc_dwt_* (include/cudacompress_amd.h).  All sizes must be even.
#ifndef CUDACOMPRESS_AMD_DROPIN_DWT_H
#define CUDACOMPRESS_AMD_DROPIN_DWT_H

#include <cudaCompress/global.h>
#include <cudacompress_amd.h>

#include <hip/hip_runtime_api.h>

namespace cudaCompress {
namespace util {

inline void dwtFloat2DForward(float* dpDest, float* dpBuffer, const float* dpSource, int sizeX, int sizeY,
                              int srcChannelCount = 1, int dstRowPitch = 0, int srcRowPitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_forward(dpDest, nullptr, dpBuffer, dpSource, sizeX, sizeY, 1, srcChannelCount, dstRowPitch, 0,
                         srcRowPitch, 0, (void*)stream);
}
inline void dwtFloat3DForward(float* dpDest, float* dpBuffer2, float* dpBuffer1, const float* dpSource, int sizeX,
                              int sizeY, int sizeZ, int srcChannelCount = 1, int dstRowPitch = 0, int dstSlicePitch = 0,
                              int srcRowPitch = 0, int srcSlicePitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_forward(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, srcChannelCount, dstRowPitch,
                         dstSlicePitch, srcRowPitch, srcSlicePitch, (void*)stream);
}
inline void dwtFloat2DInverse(float* dpDest, float* dpBuffer, const float* dpSource, int sizeX, int sizeY,
                              int dstChannelCount = 1, int dstRowPitch = 0, int srcRowPitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_inverse(dpDest, nullptr, dpBuffer, dpSource, sizeX, sizeY, 1, dstChannelCount, dstRowPitch, 0,
                         srcRowPitch, 0, (void*)stream);
}
inline void dwtFloat3DInverse(float* dpDest, float* dpBuffer2, float* dpBuffer1, const float* dpSource, int sizeX,
                              int sizeY, int sizeZ, int dstChannelCount = 1, int dstRowPitch = 0, int dstSlicePitch = 0,
                              int srcRowPitch = 0, int srcSlicePitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_inverse(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch,
                         dstSlicePitch, srcRowPitch, srcSlicePitch, (void*)stream);
}
// dppHighpass: device array of 3 (2D) or 7 (3D) device pointers to the quantized highpass bands
inline void dwtFloat2DInverseFromSymbols(float* dpDest, float* dpBuffer, const float* dpLowpass,
                                         const ushort* const* dppHighpass, float quantStep, int sizeX, int sizeY,
                                         int dstChannelCount = 1, int dstRowPitch = 0, int lowpassRowPitch = 0,
                                         hipStream_t stream = 0) {
    cc_dwt_float_inverse_from_symbols(dpDest, nullptr, dpBuffer, dpLowpass, (const void* const*)dppHighpass, 2,
                                      quantStep, sizeX, sizeY, 1, dstChannelCount, dstRowPitch, 0, lowpassRowPitch, 0,
                                      (void*)stream);
}
inline void dwtFloat3DInverseFromSymbols(float* dpDest, float* dpBuffer2, float* dpBuffer1, const float* dpLowpass,
                                         const ushort* const* dppHighpass, float quantStep, int sizeX, int sizeY,
                                         int sizeZ, int dstChannelCount = 1, int dstRowPitch = 0, int dstSlicePitch = 0,
                                         int lowpassRowPitch = 0, int lowpassSlicePitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_inverse_from_symbols(dpDest, dpBuffer2, dpBuffer1, dpLowpass, (const void* const*)dppHighpass, 2,
                                      quantStep, sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch,
                                      lowpassRowPitch, lowpassSlicePitch, (void*)stream);
}
inline void dwtFloat2DInverseFromSymbols(float* dpDest, float* dpBuffer, const float* dpLowpass,
                                         const uint* const* dppHighpass, float quantStep, int sizeX, int sizeY,
                                         int dstChannelCount = 1, int dstRowPitch = 0, int lowpassRowPitch = 0,
                                         hipStream_t stream = 0) {
    cc_dwt_float_inverse_from_symbols(dpDest, nullptr, dpBuffer, dpLowpass, (const void* const*)dppHighpass, 4,
                                      quantStep, sizeX, sizeY, 1, dstChannelCount, dstRowPitch, 0, lowpassRowPitch, 0,
                                      (void*)stream);
}
inline void dwtFloat3DInverseFromSymbols(float* dpDest, float* dpBuffer2, float* dpBuffer1, const float* dpLowpass,
                                         const uint* const* dppHighpass, float quantStep, int sizeX, int sizeY,
                                         int sizeZ, int dstChannelCount = 1, int dstRowPitch = 0, int dstSlicePitch = 0,
                                         int lowpassRowPitch = 0, int lowpassSlicePitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_inverse_from_symbols(dpDest, dpBuffer2, dpBuffer1, dpLowpass, (const void* const*)dppHighpass, 4,
                                      quantStep, sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch,
                                      lowpassRowPitch, lowpassSlicePitch, (void*)stream);
}

inline void dwtIntForward(short* dpDest, short* dpBuffer, const short* dpSource, int sizeX, int sizeY, int sizeZ,
                          int dstRowPitch = 0, int dstSlicePitch = 0, int srcRowPitch = 0, int srcSlicePitch = 0,
                          hipStream_t stream = 0) {
    cc_dwt_int_forward(dpDest, dpBuffer, dpSource, sizeX, sizeY, sizeZ, dstRowPitch, dstSlicePitch, srcRowPitch,
                       srcSlicePitch, (void*)stream);
}
inline void dwtIntInverse(short* dpDest, short* dpBuffer, const short* dpSource, int sizeX, int sizeY, int sizeZ,
                          int dstRowPitch = 0, int dstSlicePitch = 0, int srcRowPitch = 0, int srcSlicePitch = 0,
                          hipStream_t stream = 0) {
    cc_dwt_int_inverse(dpDest, dpBuffer, dpSource, sizeX, sizeY, sizeZ, dstRowPitch, dstSlicePitch, srcRowPitch,
                       srcSlicePitch, (void*)stream);
}

// byte images: dpSource / dpDest point to the channel within the first element
inline void dwtFloat2DForwardFromByte(float* dpDest, float* dpBuffer, const byte* dpSource, int sizeX, int sizeY,
                                      int srcChannelCount = 1, int dstRowPitch = 0, int srcRowPitch = 0,
                                      hipStream_t stream = 0) {
    cc_dwt_float_forward_from_byte(dpDest, nullptr, dpBuffer, dpSource, sizeX, sizeY, 1, srcChannelCount, dstRowPitch,
                                   0, srcRowPitch, 0, (void*)stream);
}
inline void dwtFloat3DForwardFromByte(float* dpDest, float* dpBuffer2, float* dpBuffer1, const byte* dpSource,
                                      int sizeX, int sizeY, int sizeZ, int srcChannelCount = 1, int dstRowPitch = 0,
                                      int dstSlicePitch = 0, int srcRowPitch = 0, int srcSlicePitch = 0,
                                      hipStream_t stream = 0) {
    cc_dwt_float_forward_from_byte(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, srcChannelCount,
                                   dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch, (void*)stream);
}
inline void dwtFloat2DInverseToByte(byte* dpDest, float* dpBuffer, const float* dpSource, int sizeX, int sizeY,
                                    int dstChannelCount = 1, int dstRowPitch = 0, int srcRowPitch = 0,
                                    hipStream_t stream = 0) {
    cc_dwt_float_inverse_to_byte(dpDest, nullptr, dpBuffer, dpSource, sizeX, sizeY, 1, dstChannelCount, dstRowPitch, 0,
                                 srcRowPitch, 0, (void*)stream);
}
inline void dwtFloat3DInverseToByte(byte* dpDest, float* dpBuffer2, float* dpBuffer1, const float* dpSource, int sizeX,
                                    int sizeY, int sizeZ, int dstChannelCount = 1, int dstRowPitch = 0,
                                    int dstSlicePitch = 0, int srcRowPitch = 0, int srcSlicePitch = 0,
                                    hipStream_t stream = 0) {
    cc_dwt_float_inverse_to_byte(dpDest, dpBuffer2, dpBuffer1, dpSource, sizeX, sizeY, sizeZ, dstChannelCount,
                                 dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch, (void*)stream);
}
inline void dwtFloat2DInverseFromSymbolsToByte(byte* dpDest, float* dpBuffer, const float* dpLowpass,
                                               const ushort* const* dppHighpass, float quantStep, int sizeX, int sizeY,
                                               int dstChannelCount = 1, int dstRowPitch = 0, int lowpassRowPitch = 0,
                                               hipStream_t stream = 0) {
    cc_dwt_float_inverse_from_symbols_to_byte(dpDest, nullptr, dpBuffer, dpLowpass, dppHighpass, quantStep, sizeX,
                                              sizeY, 1, dstChannelCount, dstRowPitch, 0, lowpassRowPitch, 0,
                                              (void*)stream);
}
inline void dwtFloat3DInverseFromSymbolsToByte(byte* dpDest, float* dpBuffer2, float* dpBuffer1,
                                               const float* dpLowpass, const ushort* const* dppHighpass,
                                               float quantStep, int sizeX, int sizeY, int sizeZ,
                                               int dstChannelCount = 1, int dstRowPitch = 0, int dstSlicePitch = 0,
                                               int lowpassRowPitch = 0, int lowpassSlicePitch = 0,
                                               hipStream_t stream = 0) {
    cc_dwt_float_inverse_from_symbols_to_byte(dpDest, dpBuffer2, dpBuffer1, dpLowpass, dppHighpass, quantStep, sizeX,
                                              sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch,
                                              lowpassRowPitch, lowpassSlicePitch, (void*)stream);
}
inline void dwtFloat2DForwardLowpassOnlyFromByte(float* dpDest, float* dpBuffer, const byte* dpSource, int sizeX,
                                                 int sizeY, int srcChannelCount = 1, int dstRowPitch = 0,
                                                 int srcRowPitch = 0, hipStream_t stream = 0) {
    cc_dwt_float_forward_lowpass_only_from_byte(dpDest, dpBuffer, dpSource, sizeX, sizeY, srcChannelCount,
                                                dstRowPitch, srcRowPitch, (void*)stream);
}
inline void dwtIntForwardLowpassOnly(short* dpDest, short* dpBuffer, const short* dpSource, int sizeX, int sizeY,
                                     int dstRowPitch = 0, int srcRowPitch = 0, hipStream_t stream = 0) {
    cc_dwt_int_forward_lowpass_only(dpDest, dpBuffer, dpSource, sizeX, sizeY, dstRowPitch, srcRowPitch,
                                    (void*)stream);
}

}  // namespace util
}  // namespace cudaCompress

#endif
