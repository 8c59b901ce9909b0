// Volume codecs (caller layouts of src/examples/CompressVolume.cpp); output in inst->vol_bits.
#pragma once

#include "instance.hpp"

namespace cca {

void compress_volume_float(cc_instance* I, const float* image, uint32_t sx, uint32_t sy, uint32_t sz,
                           uint32_t levels, float q0, bool rle_only_lvl0, int symSize);
void decompress_volume_float(cc_instance* I, float* image, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t levels,
                             const uint32_t* words, uint64_t nbits, float q0, bool rle_only_lvl0, int symSize);
void compress_volume_lossless(cc_instance* I, const int16_t* image, uint32_t sx, uint32_t sy, uint32_t sz,
                              uint32_t levels);
void decompress_volume_lossless(cc_instance* I, int16_t* image, uint32_t sx, uint32_t sy, uint32_t sz,
                                uint32_t levels, const uint32_t* words, uint64_t nbits);
// One channel of a multi-channel decode (VolumeChannel, src/examples/CompressVolume.h:36-42).
struct VolumeChannelIn {
    float* image;
    const uint32_t* words;
    uint64_t nbits;
    float step;  // quantizationStepLevel0 (float codec) or quantizationStep (QuantFirst)
};

void decompress_volume_float_multichannel(cc_instance* I, const VolumeChannelIn* ch, uint32_t nch, uint32_t sx,
                                          uint32_t sy, uint32_t sz, uint32_t levels, bool rle_only_lvl0,
                                          int symSize);
void decompress_volume_float_quant_first_multichannel(cc_instance* I, const VolumeChannelIn* ch, uint32_t nch,
                                                      uint32_t sx, uint32_t sy, uint32_t sz, uint32_t levels,
                                                      bool rle_only_lvl0);
void compress_volume_float_quant_first(cc_instance* I, const float* image, uint32_t sx, uint32_t sy, uint32_t sz,
                                       uint32_t levels, float step, bool rle_only_lvl0);
void decompress_volume_float_quant_first(cc_instance* I, float* image, uint32_t sx, uint32_t sy, uint32_t sz,
                                         uint32_t levels, const uint32_t* words, uint64_t nbits, float step,
                                         bool rle_only_lvl0);

}  // namespace cca
