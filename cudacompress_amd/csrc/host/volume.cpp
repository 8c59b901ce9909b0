// Volume codecs built on the library API, restating the caller layouts of
// src/examples/CompressVolume.cpp (compressVolumeFloat :313-409, decompressVolumeFloat :412-516,
// compressVolumeLossless :234-277, decompressVolumeLossless :279-309,
// compressVolumeFloatQuantFirst :651-757, decompressVolumeFloatQuantFirst :759-856, the
// multi-channel decoders decompressVolumeFloatMultiChannel :518-648 and
// decompressVolumeFloatQuantFirstMultiChannel :863-985) and, for 2D lossless input, the three-band
// layout of src/examples/CompressHeightfield.cpp:116-126.
//
// Differences from the reference callers, none of which changes the bitstream:
//  * every kernel runs on the instance stream (the reference quantizes on the default stream);
//  * decode outputs are not cleared first (decodeRLHuff here writes every element);
//  * 2D lossless input (sizeZ == 1) uses 3 highpass bands per level -- compressVolumeLossless
//    itself only handles 3D (it would compute a zero blockSizeZ);
//  * decompressVolumeFloatQuantFirst starts its inverse transforms from the buffer holding the
//    coarsest level's coefficients; the reference starts from buffer numLevels % 2, the same
//    buffer for an even level count (an odd count reads a buffer its decode did not fill); its
//    multi-channel form also carries that index from one channel to the next;
//  * the multi-channel decoders give each channel exactly its single-channel decode.
#include "volume.hpp"

#include <cstring>

#include "../kernels.hpp"
#include "codec.hpp"

namespace cca {
namespace {

void append_segments(cc_instance* I, std::vector<uint32_t>& bits) {
    // every encode call here starts on a word boundary, so segments concatenate as words
    if (I->out_words.empty()) return;
    bits.insert(bits.end(), I->out_words.begin(), I->out_words.end());
}

struct Band {
    uint32_t x, y, z;
};
Band band_of(uint32_t block, bool is3d) {  // CompressVolume.cpp:375-379, CompressHeightfield.cpp:119-121
    if (is3d) return Band{block % 2, block / 2 % 2, block / 4};
    return Band{block % 2, block / 2, 0};
}

}  // namespace

void compress_volume_float(cc_instance* I, const float* image, uint32_t sx, uint32_t sy, uint32_t sz,
                           uint32_t levels, float q0, bool rle_only_lvl0, int symSize) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(levels >= 1, CC_ERR_INVALID_ARG, "numLevels must be >= 1");
    CCA_REQUIRE(sz > 1, CC_ERR_INVALID_ARG, "compressVolumeFloat is 3D");
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && sz % f_all == 0, CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint32_t nb = 7;
    float* dwt[2] = {I->d_vol[0].as<float>(elems), I->d_vol[1].as<float>(elems)};
    const uint64_t per_block_max = elems / 8;
    unsigned char* sym = (unsigned char*)I->d_vol[2].get(nb * per_block_max * symSize);
    std::vector<const void*> streams(nb);
    for (uint32_t i = 0; i < nb; i++) streams[i] = sym + i * per_block_max * symSize;

    int d = 0;
    dwt97_forward(dwt[d], dwt[1 - d], dwt[d], image, (int)sx, (int)sy, (int)sz, 1, 0, 0, 0, 0, st);
    uint32_t f = 2;
    for (uint32_t level = 1; level < levels; level++) {
        dwt97_forward(dwt[1 - d], dwt[d], dwt[1 - d], dwt[d], (int)(sx / f), (int)(sy / f), (int)(sz / f), 1, (int)sx,
                      (int)(sx * sy), (int)sx, (int)(sx * sy), st);
        d = 1 - d;
        f *= 2;
    }
    std::vector<uint32_t>& bits = I->vol_bits;
    bits.clear();
    // lowpass band (CompressVolume.cpp:353-362)
    float q = q0 / float(f / 2);
    quantize_to_symbols(sym, symSize, dwt[d], false, sx / f, sy / f, sz / f, q, sx, sx * sy, 0, st);
    encode_streams(I, false, symSize, streams.data(), 1, (uint32_t)(elems / ((uint64_t)f * f * f)), nullptr);
    append_segments(I, bits);
    // highpass bands, coarsest level first (CompressVolume.cpp:365-402)
    for (int level = (int)levels - 1; level >= 0; level--) {
        q = q0 / float(f / 2);
        const uint32_t bx = sx / f, by = sy / f, bz = sz / f;
        for (uint32_t b = 1; b <= nb; b++) {
            Band bd = band_of(b, true);
            uint64_t off = bd.x * bx + (uint64_t)bd.y * by * sx + (uint64_t)bd.z * bz * sx * sy;
            quantize_to_symbols(sym + (b - 1) * per_block_max * symSize, symSize, dwt[d] + off, false, bx, by, bz, q,
                                sx, sx * sy, 0, st);
        }
        d = 1 - d;
        const uint32_t count = bx * by * bz;
        encode_streams(I, !(level > 0 && rle_only_lvl0), symSize, streams.data(), nb, count, nullptr);
        append_segments(I, bits);
        f /= 2;
    }
}

void decompress_volume_float(cc_instance* I, float* image, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t levels,
                             const uint32_t* words, uint64_t nbits, float q0, bool rle_only_lvl0, int symSize) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(levels >= 1 && sz > 1, CC_ERR_INVALID_ARG, "bad volume parameters");
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && sz % f_all == 0, CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint32_t nb = 7;
    float* dwt[2] = {I->d_vol[0].as<float>(elems), I->d_vol[1].as<float>(elems)};
    const uint64_t per_block_max = elems / 8;
    unsigned char* sym = (unsigned char*)I->d_vol[2].get(nb * per_block_max * symSize);
    std::vector<void*> streams(nb);
    for (uint32_t i = 0; i < nb; i++) streams[i] = sym + i * per_block_max * symSize;
    // device table of the 7 band pointers for the fused dequantize+IDWT
    void** hptr = I->h_vol_ptrs.as<void*>(nb);
    for (uint32_t i = 0; i < nb; i++) hptr[i] = streams[i];
    void** dptr = I->d_vol_ptrs.as<void*>(nb);
    CCA_HIP(hipMemcpyAsync(dptr, hptr, sizeof(void*) * nb, hipMemcpyHostToDevice, st));

    uint64_t pos = 0;
    uint32_t f = f_all;
    // lowpass (CompressVolume.cpp:466-475)
    const uint32_t nlp = (uint32_t)(elems / ((uint64_t)f * f * f));
    decode_streams(I, false, symSize, &words, &nbits, &pos, 1, nullptr, streams.data(), 1, nlp);
    unquantize_from_symbols(image, sym, symSize, sx / f, sy / f, sz / f, q0 / float(f / 2), sx, sx * sy, 0, st);
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint32_t count = (sx / f) * (sy / f) * (sz / f);
        decode_streams(I, !(level > 0 && rle_only_lvl0), symSize, &words, &nbits, &pos, 1, nullptr, streams.data(), nb,
                       count);
        f /= 2;
        dwt97_inverse_from_symbols(image, dwt[1], dwt[0], image, (const void* const*)dptr, symSize, q0 / float(f),
                                   (int)(sx / f), (int)(sy / f), (int)(sz / f), 1, (int)sx, (int)(sx * sy), (int)sx,
                                   (int)(sx * sy), st);
    }
    // the pointer table upload must finish before the pinned copy is rewritten
    CCA_HIP(hipStreamSynchronize(st));
}

namespace {

void copy3d(int16_t* dst, uint64_t dRow, uint64_t dSlice, const int16_t* src, uint64_t sRow, uint64_t sSlice,
            uint32_t w, uint32_t h, uint32_t dpt, hipStream_t st) {
    for (uint32_t z = 0; z < dpt; z++)
        CCA_HIP(hipMemcpy2DAsync(dst + z * dSlice, dRow * 2, src + z * sSlice, sRow * 2, (size_t)w * 2, h,
                                 hipMemcpyDeviceToDevice, st));
}

// compressVolumeLosslessOneLevel, CompressVolume.cpp:84-150 (3D) / compressHeightfieldOneLevel's
// band layout (2D)
void lossless_one_level(cc_instance* I, const int16_t* src, uint32_t sx, uint32_t sy, uint32_t sz, int16_t* lowpass,
                        std::vector<uint32_t>& out) {
    hipStream_t st = I->stream;
    const bool is3d = sz > 1;
    const uint32_t nb = is3d ? 7 : 3;
    const uint32_t bx = sx / 2, by = sy / 2, bz = is3d ? sz / 2 : 1;
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint64_t per_block = (uint64_t)bx * by * bz;
    int16_t* dwt0 = I->d_vol[0].as<int16_t>(elems);
    int16_t* dwt1 = I->d_vol[1].as<int16_t>(elems);
    uint16_t* sym = I->d_vol[2].as<uint16_t>(nb * per_block);
    dwt53_forward(dwt1, dwt0, src, (int)sx, (int)sy, (int)sz, 0, 0, 0, 0, st);
    copy3d(lowpass, bx, (uint64_t)bx * by, dwt1, sx, (uint64_t)sx * sy, bx, by, bz, st);
    std::vector<const void*> streams(nb);
    for (uint32_t b = 1; b <= nb; b++) {
        Band bd = band_of(b, is3d);
        uint64_t off = bd.x * bx + (uint64_t)bd.y * by * sx + (uint64_t)bd.z * bz * sx * sy;
        streams[b - 1] = sym + (b - 1) * per_block;
        symbolize(sym + (b - 1) * per_block, dwt1 + off, bx, by, bz, sx, sx * sy, st);
    }
    encode_streams(I, true, 2, streams.data(), nb, (uint32_t)per_block, nullptr);
    out.assign(I->out_words.begin(), I->out_words.end());
}

}  // namespace

void compress_volume_lossless(cc_instance* I, const int16_t* image, uint32_t sx, uint32_t sy, uint32_t sz,
                              uint32_t levels) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(I->log2 <= 16, CC_ERR_INVALID_ARG, "lossless volumes use Symbol16 (log2 <= 16)");
    CCA_REQUIRE(levels >= 1, CC_ERR_INVALID_ARG, "numLevels must be >= 1");
    const bool is3d = sz > 1;
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && (!is3d || sz % f_all == 0), CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    const uint64_t elems = (uint64_t)sx * sy * sz;
    int16_t* buf0 = I->d_vol[3].as<int16_t>(elems + elems / 4 + 64);
    int16_t* buf1 = buf0 + align_up(elems, 64);  // lowpass of the current level (<= elems/4)
    std::vector<std::vector<uint32_t>> level_bits(levels);
    uint32_t f = 1;
    for (uint32_t level = 0; level < levels; level++) {
        const int16_t* src = level == 0 ? image : buf0;
        lossless_one_level(I, src, sx / f, sy / f, is3d ? sz / f : 1, buf1, level_bits[level]);
        f *= 2;
        const uint64_t lp = (uint64_t)(sx / f) * (sy / f) * (is3d ? sz / f : 1);
        CCA_HIP(hipMemcpyAsync(buf0, buf1, lp * 2, hipMemcpyDeviceToDevice, st));
    }
    const uint32_t lx = sx / f, ly = sy / f, lz = is3d ? sz / f : 1;
    uint16_t* lsym = (uint16_t*)buf1;
    symbolize(lsym, buf0, lx, ly, lz, lx, lx * ly, st);
    const void* s0 = lsym;
    encode_streams(I, false, 2, &s0, 1, lx * ly * lz, nullptr);
    std::vector<uint32_t>& bits = I->vol_bits;
    bits.assign(I->out_words.begin(), I->out_words.end());
    for (int level = (int)levels - 1; level >= 0; level--)
        bits.insert(bits.end(), level_bits[level].begin(), level_bits[level].end());
}

void decompress_volume_lossless(cc_instance* I, int16_t* image, uint32_t sx, uint32_t sy, uint32_t sz,
                                uint32_t levels, const uint32_t* words, uint64_t nbits) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(I->log2 <= 16, CC_ERR_INVALID_ARG, "lossless volumes use Symbol16 (log2 <= 16)");
    CCA_REQUIRE(levels >= 1, CC_ERR_INVALID_ARG, "numLevels must be >= 1");
    const bool is3d = sz > 1;
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && (!is3d || sz % f_all == 0), CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    const uint64_t elems = (uint64_t)sx * sy * sz;
    int16_t* buf0 = I->d_vol[3].as<int16_t>(elems + elems / 4 + 64);
    int16_t* buf1 = buf0 + align_up(elems, 64);
    uint32_t f = f_all;
    uint32_t lx = sx / f, ly = sy / f, lz = is3d ? sz / f : 1;
    uint64_t pos = 0;
    void* s0 = buf1;
    decode_streams(I, false, 2, &words, &nbits, &pos, 1, nullptr, &s0, 1, lx * ly * lz);
    unsymbolize(buf0, (const uint16_t*)buf1, lx, ly, lz, lx, lx * ly, st);
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint64_t lp = (uint64_t)(sx / f) * (sy / f) * (is3d ? sz / f : 1);
        CCA_HIP(hipMemcpyAsync(buf1, buf0, lp * 2, hipMemcpyDeviceToDevice, st));
        f /= 2;
        int16_t* target = level == 0 ? image : buf0;
        // decompressVolumeLosslessOneLevel, CompressVolume.cpp:152-226
        const uint32_t cx = sx / f, cy = sy / f, cz = is3d ? sz / f : 1;
        const uint32_t nb = is3d ? 7 : 3;
        const uint32_t bx = cx / 2, by = cy / 2, bz = is3d ? cz / 2 : 1;
        const uint64_t celems = (uint64_t)cx * cy * cz, per_block = (uint64_t)bx * by * bz;
        int16_t* dwt0 = I->d_vol[0].as<int16_t>(celems);
        int16_t* dwt1 = I->d_vol[1].as<int16_t>(celems);
        uint16_t* sym = I->d_vol[2].as<uint16_t>(nb * per_block);
        copy3d(dwt1, cx, (uint64_t)cx * cy, buf1, bx, (uint64_t)bx * by, bx, by, bz, st);
        std::vector<void*> streams(nb);
        for (uint32_t b = 0; b < nb; b++) streams[b] = sym + b * per_block;
        decode_streams(I, true, 2, &words, &nbits, &pos, 1, nullptr, streams.data(), nb, (uint32_t)per_block);
        for (uint32_t b = 1; b <= nb; b++) {
            Band bd = band_of(b, is3d);
            uint64_t off = bd.x * bx + (uint64_t)bd.y * by * cx + (uint64_t)bd.z * bz * cx * cy;
            unsymbolize(dwt1 + off, sym + (b - 1) * per_block, bx, by, bz, cx, cx * cy, st);
        }
        dwt53_inverse(target, dwt0, dwt1, (int)cx, (int)cy, (int)cz, 0, 0, 0, 0, st);
    }
}

void compress_volume_float_quant_first(cc_instance* I, const float* image, uint32_t sx, uint32_t sy, uint32_t sz,
                                       uint32_t levels, float step, bool rle_only_lvl0) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(I->log2 <= 16, CC_ERR_INVALID_ARG, "QuantFirst volumes use Symbol16 (log2 <= 16)");
    CCA_REQUIRE(levels >= 1 && sz > 1, CC_ERR_INVALID_ARG, "QuantFirst volumes are 3D with numLevels >= 1");
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && sz % f_all == 0, CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint32_t nb = 7;
    int16_t* dwt[2] = {I->d_vol[0].as<int16_t>(elems), I->d_vol[1].as<int16_t>(elems)};
    const uint64_t per_block_max = elems / 8;
    uint16_t* sym = I->d_vol[2].as<uint16_t>(nb * per_block_max);
    std::vector<const void*> streams(nb);
    for (uint32_t i = 0; i < nb; i++) streams[i] = sym + i * per_block_max;
    // uniform quantization of the whole volume (CompressVolume.cpp:685-687)
    quantize_to_short(dwt[0], image, sx, sy, sz, step, sx, sx * sy, 1, st);
    // every 5/3 level: data path src -> dst -> buf -> dst with src and buf the same buffer, so the
    // coarsest level's coefficients end in dwt[d], the next finer level's in dwt[1 - d], ...
    int d = 0;
    uint32_t f = 1;
    for (uint32_t level = 0; level < levels; level++) {
        dwt53_forward(dwt[1 - d], dwt[d], dwt[d], (int)(sx / f), (int)(sy / f), (int)(sz / f), (int)sx, (int)(sx * sy),
                      (int)sx, (int)(sx * sy), st);
        d = 1 - d;
        f *= 2;
    }
    std::vector<uint32_t>& bits = I->vol_bits;
    bits.clear();
    // lowpass band (CompressVolume.cpp:704-712)
    symbolize(sym, dwt[d], sx / f, sy / f, sz / f, sx, sx * sy, st);
    encode_streams(I, false, 2, streams.data(), 1, (uint32_t)(elems / ((uint64_t)f * f * f)), nullptr);
    append_segments(I, bits);
    // highpass bands, coarsest level first (CompressVolume.cpp:715-749)
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint32_t bx = sx / f, by = sy / f, bz = sz / f;
        for (uint32_t b = 1; b <= nb; b++) {
            const Band bd = band_of(b, true);
            const uint64_t off = bd.x * bx + (uint64_t)bd.y * by * sx + (uint64_t)bd.z * bz * sx * sy;
            symbolize(sym + (b - 1) * per_block_max, dwt[d] + off, bx, by, bz, sx, sx * sy, st);
        }
        d = 1 - d;
        encode_streams(I, !(level > 0 && rle_only_lvl0), 2, streams.data(), nb, bx * by * bz, nullptr);
        append_segments(I, bits);
        f /= 2;
    }
}

void decompress_volume_float_quant_first(cc_instance* I, float* image, uint32_t sx, uint32_t sy, uint32_t sz,
                                         uint32_t levels, const uint32_t* words, uint64_t nbits, float step,
                                         bool rle_only_lvl0) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(I->log2 <= 16, CC_ERR_INVALID_ARG, "QuantFirst volumes use Symbol16 (log2 <= 16)");
    CCA_REQUIRE(levels >= 1 && sz > 1, CC_ERR_INVALID_ARG, "QuantFirst volumes are 3D with numLevels >= 1");
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && sz % f_all == 0, CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint32_t nb = 7;
    int16_t* dwt[2] = {I->d_vol[0].as<int16_t>(elems), I->d_vol[1].as<int16_t>(elems)};
    const uint64_t per_block_max = elems / 8;
    uint16_t* sym = I->d_vol[2].as<uint16_t>(nb * per_block_max);
    std::vector<void*> streams(nb);
    for (uint32_t i = 0; i < nb; i++) streams[i] = sym + i * per_block_max;
    uint64_t pos = 0;
    uint32_t f = f_all;
    // lowpass band into dwt[0] (CompressVolume.cpp:794-801)
    decode_streams(I, false, 2, &words, &nbits, &pos, 1, nullptr, streams.data(), 1,
                   (uint32_t)(elems / ((uint64_t)f * f * f)));
    unsymbolize(dwt[0], sym, sx / f, sy / f, sz / f, sx, sx * sy, st);
    // highpass bands, coarsest level first, alternating buffers from dwt[0] (:804-830)
    int d = 0;
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint32_t bx = sx / f, by = sy / f, bz = sz / f;
        decode_streams(I, !(level > 0 && rle_only_lvl0), 2, &words, &nbits, &pos, 1, nullptr, streams.data(), nb,
                       bx * by * bz);
        f /= 2;
        for (uint32_t b = 1; b <= nb; b++) {
            const Band bd = band_of(b, true);
            const uint64_t off = bd.x * bx + (uint64_t)bd.y * by * sx + (uint64_t)bd.z * bz * sx * sy;
            unsymbolize(dwt[d] + off, sym + (b - 1) * per_block_max, bx, by, bz, sx, sx * sy, st);
        }
        d = 1 - d;
    }
    // inverse levels from the coarsest, whose coefficients are in dwt[0] (:833-838)
    d = 0;
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint32_t fl = 1u << level;
        dwt53_inverse(dwt[1 - d], dwt[d], dwt[d], (int)(sx / fl), (int)(sy / fl), (int)(sz / fl), (int)sx,
                      (int)(sx * sy), (int)sx, (int)(sx * sy), st);
        d = 1 - d;
    }
    unquantize_from_short(image, dwt[d], sx, sy, sz, step, sx, sx * sy, 1, st);
}

namespace {

void check_channels(cc_instance* I, const VolumeChannelIn* ch, uint32_t nch, uint32_t sx, uint32_t sy, uint32_t sz,
                    uint32_t levels) {
    CCA_REQUIRE(nch >= 1 && ch, CC_ERR_INVALID_ARG, "no channels");
    CCA_REQUIRE(levels >= 1 && sz > 1, CC_ERR_INVALID_ARG, "bad volume parameters");
    const uint32_t f_all = 1u << levels;
    CCA_REQUIRE(sx % f_all == 0 && sy % f_all == 0 && sz % f_all == 0, CC_ERR_INVALID_ARG,
                "sizes must be divisible by 2^numLevels");
    for (uint32_t c = 0; c < nch; c++)
        CCA_REQUIRE(ch[c].image && ch[c].words, CC_ERR_INVALID_ARG, "null channel buffer");
}

// Bitstream cursors of the channels; stream s of a call with `per` streams per channel reads
// channel s / per.
struct ChannelStreams {
    std::vector<const uint32_t*> words;
    std::vector<uint64_t> nbits, pos;
    std::vector<uint32_t> s2b_lp, s2b_hp;
    ChannelStreams(const VolumeChannelIn* ch, uint32_t nch, uint32_t per) : nbits(nch), pos(nch, 0) {
        for (uint32_t c = 0; c < nch; c++) {
            words.push_back(ch[c].words);
            nbits[c] = ch[c].nbits;
            s2b_lp.push_back(c);
            for (uint32_t b = 0; b < per; b++) s2b_hp.push_back(c);
        }
    }
};

}  // namespace

void decompress_volume_float_multichannel(cc_instance* I, const VolumeChannelIn* ch, uint32_t nch, uint32_t sx,
                                          uint32_t sy, uint32_t sz, uint32_t levels, bool rle_only_lvl0,
                                          int symSize) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    check_channels(I, ch, nch, sx, sy, sz, levels);
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint32_t nb = 7, ns = nb * nch;
    float* dwt[2] = {I->d_vol[0].as<float>(elems), I->d_vol[1].as<float>(elems)};
    const uint64_t per_block_max = elems / 8;
    unsigned char* sym = (unsigned char*)I->d_vol[2].get((uint64_t)ns * per_block_max * symSize);
    std::vector<void*> streams(ns);
    for (uint32_t i = 0; i < ns; i++) streams[i] = sym + i * per_block_max * symSize;
    // device table of every channel's 7 band pointers for the fused dequantize + IDWT
    void** hptr = I->h_vol_ptrs.as<void*>(ns);
    for (uint32_t i = 0; i < ns; i++) hptr[i] = streams[i];
    void** dptr = I->d_vol_ptrs.as<void*>(ns);
    CCA_HIP(hipMemcpyAsync(dptr, hptr, sizeof(void*) * ns, hipMemcpyHostToDevice, st));
    ChannelStreams cs(ch, nch, nb);
    uint32_t f = 1u << levels;
    // every channel's lowpass band in one decode (CompressVolume.cpp:588-599)
    decode_streams(I, false, symSize, cs.words.data(), cs.nbits.data(), cs.pos.data(), nch, cs.s2b_lp.data(),
                   streams.data(), nch, (uint32_t)(elems / ((uint64_t)f * f * f)));
    for (uint32_t c = 0; c < nch; c++)
        unquantize_from_symbols(ch[c].image, streams[c], symSize, sx / f, sy / f, sz / f, ch[c].step / float(f / 2),
                                sx, sx * sy, 0, st);
    // per level every channel's 7 bands in one decode, then each channel's IDWT (:602-641)
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint32_t count = (sx / f) * (sy / f) * (sz / f);
        decode_streams(I, !(level > 0 && rle_only_lvl0), symSize, cs.words.data(), cs.nbits.data(), cs.pos.data(), nch,
                       cs.s2b_hp.data(), streams.data(), ns, count);
        f /= 2;
        for (uint32_t c = 0; c < nch; c++)
            dwt97_inverse_from_symbols(ch[c].image, dwt[1], dwt[0], ch[c].image, (const void* const*)(dptr + nb * c),
                                       symSize, ch[c].step / float(f), (int)(sx / f), (int)(sy / f), (int)(sz / f), 1,
                                       (int)sx, (int)(sx * sy), (int)sx, (int)(sx * sy), st);
    }
    // the pointer table upload must finish before the pinned copy is rewritten
    CCA_HIP(hipStreamSynchronize(st));
}

void decompress_volume_float_quant_first_multichannel(cc_instance* I, const VolumeChannelIn* ch, uint32_t nch,
                                                      uint32_t sx, uint32_t sy, uint32_t sz, uint32_t levels,
                                                      bool rle_only_lvl0) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    CCA_REQUIRE(I->log2 <= 16, CC_ERR_INVALID_ARG, "QuantFirst volumes use Symbol16 (log2 <= 16)");
    check_channels(I, ch, nch, sx, sy, sz, levels);
    const uint64_t elems = (uint64_t)sx * sy * sz;
    const uint32_t nb = 7, ns = nb * nch;
    // two coefficient buffers per channel (:880-883)
    int16_t* b0 = I->d_vol[0].as<int16_t>(elems * nch);
    int16_t* b1 = I->d_vol[1].as<int16_t>(elems * nch);
    auto buf = [&](uint32_t c, int k) { return (k == 0 ? b0 : b1) + elems * c; };
    const uint64_t per_block_max = elems / 8;
    uint16_t* sym = I->d_vol[2].as<uint16_t>((uint64_t)ns * per_block_max);
    std::vector<void*> streams(ns);
    for (uint32_t i = 0; i < ns; i++) streams[i] = sym + i * per_block_max;
    ChannelStreams cs(ch, nch, nb);
    uint32_t f = 1u << levels;
    decode_streams(I, false, 2, cs.words.data(), cs.nbits.data(), cs.pos.data(), nch, cs.s2b_lp.data(), streams.data(),
                   nch, (uint32_t)(elems / ((uint64_t)f * f * f)));
    for (uint32_t c = 0; c < nch; c++) unsymbolize(buf(c, 0), (const uint16_t*)streams[c], sx / f, sy / f, sz / f, sx, sx * sy, st);
    int d = 0;
    for (int level = (int)levels - 1; level >= 0; level--) {
        const uint32_t bx = sx / f, by = sy / f, bz = sz / f;
        decode_streams(I, !(level > 0 && rle_only_lvl0), 2, cs.words.data(), cs.nbits.data(), cs.pos.data(), nch,
                       cs.s2b_hp.data(), streams.data(), ns, bx * by * bz);
        f /= 2;
        for (uint32_t c = 0; c < nch; c++)
            for (uint32_t b = 1; b <= nb; b++) {
                const Band bd = band_of(b, true);
                const uint64_t off = bd.x * bx + (uint64_t)bd.y * by * sx + (uint64_t)bd.z * bz * sx * sy;
                unsymbolize(buf(c, d) + off, (const uint16_t*)streams[nb * c + b - 1], bx, by, bz, sx, sx * sy, st);
            }
        d = 1 - d;
    }
    for (uint32_t c = 0; c < nch; c++) {
        int dc = 0;  // each channel's coarsest level is in its first buffer
        for (int level = (int)levels - 1; level >= 0; level--) {
            const uint32_t fl = 1u << level;
            dwt53_inverse(buf(c, 1 - dc), buf(c, dc), buf(c, dc), (int)(sx / fl), (int)(sy / fl), (int)(sz / fl),
                          (int)sx, (int)(sx * sy), (int)sx, (int)(sx * sy), st);
            dc = 1 - dc;
        }
        unquantize_from_short(ch[c].image, buf(c, dc), sx, sy, sz, ch[c].step, sx, sx * sy, 1, st);
    }
}

}  // namespace cca
