// Compile check of the drop-in headers: every wrapper of include/cudaCompress is referenced, so a
// signature that no longer matches the C ABI fails to compile.  Not linked or run.
#include <cudaCompress/BitStream.h>
#include <cudaCompress/Encode.h>
#include <cudaCompress/Instance.h>
#include <cudaCompress/Timing.h>
#include <cudaCompress/util/Bits.h>
#include <cudaCompress/util/DWT.h>
#include <cudaCompress/util/Quantize.h>
#include <cudaCompress/util/YCoCg.h>

using namespace cudaCompress;

void dropin_use(Instance* inst, float* f, short* s, byte* b, uchar3* c3, uchar4* c4, ushort* u, uint* w,
                const ushort* const* hp) {
    util::dwtFloat2DForwardFromByte(f, f, b, 64, 64, 3);
    util::dwtFloat3DForwardFromByte(f, f, f, b, 8, 8, 8);
    util::dwtFloat2DInverseToByte(b, f, f, 64, 64, 3);
    util::dwtFloat3DInverseToByte(b, f, f, f, 8, 8, 8);
    util::dwtFloat2DInverseFromSymbolsToByte(b, f, f, hp, 0.5f, 64, 64);
    util::dwtFloat3DInverseFromSymbolsToByte(b, f, f, f, hp, 0.5f, 8, 8, 8);
    util::dwtFloat2DForwardLowpassOnlyFromByte(f, f, b, 64, 64, 4);
    util::dwtIntForwardLowpassOnly(s, s, s, 64, 64);
    util::floatToByte2D(b, 3, 1, f, 64, 64);
    util::byteToFloat2D(f, b, 3, 1, 64, 64);
    util::quantizeDifferenceToSymbols2D(u, f, 0.5f, b, 3, 1, 64, 64);
    util::quantizeDifferenceToSymbolsRoundtrip2D(u, f, 0.5f, b, 3, 1, 64, 64, 0, util::QUANTIZE_UNIFORM);
    util::unquantizeDifferenceFromSymbols2D(f, u, 0.5f, b, 3, 1, 64, 64);
    util::convertRGBtoYCoCg(c3, c3, 16);
    util::convertYCoCgtoRGB(c4, c4, 16);
    util::getLSB(s, w, 4096);
    util::removeLSB(s, w, 4096, 1);
    util::appendLSB(s, w, 4096, -1);
    util::shiftElems(s, 4096, 3);
    setTimingDetail(inst, TIMING_DETAIL_MEDIUM);
    std::vector<std::string> names;
    std::vector<float> times;
    getTimings(inst, names, times);
    printTimings(inst);
    resetTimings(inst);
}
