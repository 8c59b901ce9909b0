// Drop-in replacement of include/cudaCompress/util/Quantize.h (m0bl0/cudaCompress): header-only
// wrappers over cc_quantize_* / cc_symbolize (include/cudacompress_amd.h).  As in the reference,
// these run on the null stream (unquantizeDifferenceFromSymbols2D takes a stream).
#ifndef CUDACOMPRESS_AMD_DROPIN_QUANTIZE_H
#define CUDACOMPRESS_AMD_DROPIN_QUANTIZE_H

#include <cudaCompress/Instance.h>
#include <cudaCompress/global.h>
#include <cudacompress_amd.h>

namespace cudaCompress {
namespace util {

enum EQuantizeType {
    QUANTIZE_DEADZONE = 0,  // midtread quantizer with twice larger zero bin
    QUANTIZE_UNIFORM,       // standard uniform midtread quantizer
    QUANTIZE_COUNT
};

inline void symbolize(ushort* dpSymbols, const short* dpData, uint sizeX, uint sizeY, uint sizeZ, uint rowPitchSrc = 0,
                      uint slicePitchSrc = 0) {
    cc_symbolize(dpSymbols, dpData, sizeX, sizeY, sizeZ, rowPitchSrc, slicePitchSrc, nullptr);
}
inline void unsymbolize(short* dpData, const ushort* dpSymbols, uint sizeX, uint sizeY, uint sizeZ,
                        uint rowPitchDst = 0, uint slicePitchDst = 0) {
    cc_unsymbolize(dpData, dpSymbols, sizeX, sizeY, sizeZ, rowPitchDst, slicePitchDst, nullptr);
}

#define CUDACOMPRESS_AMD_QUANT_OVERLOADS(SYM, SIZE)                                                                  \
    inline void quantizeToSymbols(SYM* dpSymbols, const float* dpData, uint sizeX, uint sizeY, uint sizeZ,          \
                                  float quantizationStep, uint rowPitchSrc = 0, uint slicePitchSrc = 0,              \
                                  EQuantizeType quantType = QUANTIZE_DEADZONE) {                                     \
        cc_quantize_to_symbols(dpSymbols, SIZE, dpData, sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc,          \
                               slicePitchSrc, (int)quantType, nullptr);                                              \
    }                                                                                                                \
    inline void quantizeToSymbolsRoundtrip(SYM* dpSymbols, float* dpData, uint sizeX, uint sizeY, uint sizeZ,       \
                                           float quantizationStep, uint rowPitchSrc = 0, uint slicePitchSrc = 0,     \
                                           EQuantizeType quantType = QUANTIZE_DEADZONE) {                            \
        cc_quantize_to_symbols_roundtrip(dpSymbols, SIZE, dpData, sizeX, sizeY, sizeZ, quantizationStep,            \
                                         rowPitchSrc, slicePitchSrc, (int)quantType, nullptr);                       \
    }                                                                                                                \
    inline void unquantizeFromSymbols(float* dpData, const SYM* dpSymbols, uint sizeX, uint sizeY, uint sizeZ,      \
                                      float quantizationStep, uint rowPitchDst = 0, uint slicePitchDst = 0,          \
                                      EQuantizeType quantType = QUANTIZE_DEADZONE) {                                 \
        cc_unquantize_from_symbols(dpData, dpSymbols, SIZE, sizeX, sizeY, sizeZ, quantizationStep, rowPitchDst,      \
                                   slicePitchDst, (int)quantType, nullptr);                                          \
    }                                                                                                                \
    inline void quantizeToSymbols2D(SYM* dpSymbols, const float* dpData, uint sizeX, uint sizeY,                    \
                                    float quantizationStep, uint rowPitchSrc = 0,                                    \
                                    EQuantizeType quantType = QUANTIZE_DEADZONE) {                                   \
        quantizeToSymbols(dpSymbols, dpData, sizeX, sizeY, 1, quantizationStep, rowPitchSrc, 0, quantType);          \
    }                                                                                                                \
    inline void quantizeToSymbolsRoundtrip2D(SYM* dpSymbols, float* dpData, uint sizeX, uint sizeY,                 \
                                             float quantizationStep, uint rowPitchSrc = 0,                           \
                                             EQuantizeType quantType = QUANTIZE_DEADZONE) {                          \
        quantizeToSymbolsRoundtrip(dpSymbols, dpData, sizeX, sizeY, 1, quantizationStep, rowPitchSrc, 0, quantType); \
    }                                                                                                                \
    inline void unquantizeFromSymbols2D(float* dpData, const SYM* dpSymbols, uint sizeX, uint sizeY,                \
                                        float quantizationStep, uint rowPitchDst = 0,                                \
                                        EQuantizeType quantType = QUANTIZE_DEADZONE) {                               \
        unquantizeFromSymbols(dpData, dpSymbols, sizeX, sizeY, 1, quantizationStep, rowPitchDst, 0, quantType);      \
    }

CUDACOMPRESS_AMD_QUANT_OVERLOADS(ushort, 2)
CUDACOMPRESS_AMD_QUANT_OVERLOADS(uint, 4)
#undef CUDACOMPRESS_AMD_QUANT_OVERLOADS

inline void quantizeToShort(short* dpQuant, const float* dpData, uint sizeX, uint sizeY, uint sizeZ,
                            float quantizationStep, uint rowPitchSrc = 0, uint slicePitchSrc = 0,
                            EQuantizeType quantType = QUANTIZE_DEADZONE) {
    cc_quantize_to_short(dpQuant, dpData, sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc, slicePitchSrc,
                         (int)quantType, nullptr);
}
inline void unquantizeFromShort(float* dpData, const short* dpQuant, uint sizeX, uint sizeY, uint sizeZ,
                                float quantizationStep, uint rowPitchDst = 0, uint slicePitchDst = 0,
                                EQuantizeType quantType = QUANTIZE_DEADZONE) {
    cc_unquantize_from_short(dpData, dpQuant, sizeX, sizeY, sizeZ, quantizationStep, rowPitchDst, slicePitchDst,
                             (int)quantType, nullptr);
}
inline void floatToByte2D(byte* dpQuant, uint channelCount, uint channel, const float* dpData, uint sizeX, uint sizeY,
                          uint rowPitchSrc = 0) {
    cc_float_to_byte_2d(dpQuant, channelCount, channel, dpData, sizeX, sizeY, rowPitchSrc, nullptr);
}
inline void byteToFloat2D(float* dpData, const byte* dpQuant, uint channelCount, uint channel, uint sizeX, uint sizeY,
                          uint rowPitchDst = 0) {
    cc_byte_to_float_2d(dpData, dpQuant, channelCount, channel, sizeX, sizeY, rowPitchDst, nullptr);
}
inline void quantizeDifferenceToSymbols2D(ushort* dpSymbols, const float* dpData, float quantizationStep,
                                          const byte* dpReference, uint channelCount, uint channel, uint sizeX,
                                          uint sizeY, uint rowPitchSrc = 0,
                                          EQuantizeType quantType = QUANTIZE_DEADZONE) {
    cc_quantize_difference_to_symbols_2d(dpSymbols, dpData, quantizationStep, dpReference, channelCount, channel,
                                         sizeX, sizeY, rowPitchSrc, (int)quantType, nullptr);
}
inline void quantizeDifferenceToSymbolsRoundtrip2D(ushort* dpSymbols, float* dpData, float quantizationStep,
                                                   const byte* dpReference, uint channelCount, uint channel,
                                                   uint sizeX, uint sizeY, uint rowPitchSrc = 0,
                                                   EQuantizeType quantType = QUANTIZE_DEADZONE) {
    cc_quantize_difference_to_symbols_roundtrip_2d(dpSymbols, dpData, quantizationStep, dpReference, channelCount,
                                                   channel, sizeX, sizeY, rowPitchSrc, (int)quantType, nullptr);
}
inline void unquantizeDifferenceFromSymbols2D(float* dpData, const ushort* dpSymbols, float quantizationStep,
                                              const byte* dpReference, uint channelCount, uint channel, uint sizeX,
                                              uint sizeY, uint rowPitchDst = 0,
                                              EQuantizeType quantType = QUANTIZE_DEADZONE, hipStream_t stream = 0) {
    cc_unquantize_difference_from_symbols_2d(dpData, dpSymbols, quantizationStep, dpReference, channelCount, channel,
                                             sizeX, sizeY, rowPitchDst, (int)quantType, (void*)stream);
}
inline void getMaxAbs(Instance* pInstance, const float* dpImage, uint elemCount, float* dpValMax) {
    cc_get_max_abs(detail::raw(pInstance), dpImage, elemCount, dpValMax);
}

}  // namespace util
}  // namespace cudaCompress

#endif
