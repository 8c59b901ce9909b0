// Image-path and bit-plane utilities for gfx950: byte <-> float conversion of one channel of an
// interleaved 2D image (src/cudaCompress/util/Quantize.cu:19-43, QuantizeKernels.cui:11-35),
// quantization of the difference to a byte reference image (QuantizeKernels.cui:228-283,
// Quantize.cu:276-352), RGB <-> YCoCg (src/cudaCompress/util/YCoCg.cu:23-114) and the
// least-significant-bit planes of int16 arrays (src/cudaCompress/util/Bits.cu:11-223).
//
// All of these are elementwise and HBM-bound: one coalesced read and one coalesced write per
// element, 64x4 workgroups over 2D arrays.  The bit-plane kernels give each wave 2048 consecutive
// elements: 32 ballots of 64 lanes produce the wave's 64 bit words, which lane j keeps for word j
// and stores in one coalesced 256-byte write.
#include "../kernels.hpp"
#include "quant_ops.hpp"

namespace cca {
namespace {

// floatToByte / byteToFloat, src/cudaCompress/util/Quantize.cuh:15-23
__device__ __forceinline__ uint8_t float_to_byte(float v) { return (uint8_t)fminf(fmaxf(v + 128.5f, 0.0f), 255.0f); }
__device__ __forceinline__ float byte_to_float(uint8_t v) { return float(v) - 128.0f; }

struct Idx2 {
    uint32_t x, y;
    bool ok;
};
__device__ __forceinline__ Idx2 idx2(uint32_t sx, uint32_t sy) {
    Idx2 r;
    r.x = blockIdx.x * 64 + (threadIdx.x & 63);
    r.y = blockIdx.y * 4 + (threadIdx.x >> 6);
    r.ok = r.x < sx && r.y < sy;
    return r;
}

__global__ __launch_bounds__(256) void k_float_to_byte(uint8_t* __restrict__ out, uint32_t stride,
                                                       const float* __restrict__ data, uint32_t sx, uint32_t sy,
                                                       uint32_t row) {
    Idx2 i = idx2(sx, sy);
    if (!i.ok) return;
    out[((size_t)i.y * sx + i.x) * stride] = float_to_byte(data[(size_t)i.y * row + i.x]);
}
__global__ __launch_bounds__(256) void k_byte_to_float(float* __restrict__ data, const uint8_t* __restrict__ in,
                                                       uint32_t stride, uint32_t sx, uint32_t sy, uint32_t row) {
    Idx2 i = idx2(sx, sy);
    if (!i.ok) return;
    data[(size_t)i.y * row + i.x] = byte_to_float(in[((size_t)i.y * sx + i.x) * stride]);
}

__global__ __launch_bounds__(256) void k_quantize_diff(uint16_t* __restrict__ sym, float* __restrict__ data,
                                                       float step, float inv, const uint8_t* __restrict__ ref,
                                                       uint32_t stride, uint32_t sx, uint32_t sy, uint32_t row,
                                                       int qtype, int roundtrip) {
    Idx2 i = idx2(sx, sy);
    if (!i.ok) return;
    const size_t di = (size_t)i.y * sx + i.x, si = (size_t)i.y * row + i.x;
    const float r = byte_to_float(ref[di * stride]);
    const int q = quantize_f(data[si] - r, inv, qtype);
    sym[di] = (uint16_t)symbolize_i(q);
    // the roundtrip writes the symbol's reconstruction: unquantize(uint16 symbol) + reference
    if (roundtrip) data[si] = unquantize_i(unsymbolize_u((uint16_t)symbolize_i(q)), step, qtype) + r;
}
__global__ __launch_bounds__(256) void k_unquantize_diff(float* __restrict__ data, const uint16_t* __restrict__ sym,
                                                         float step, const uint8_t* __restrict__ ref, uint32_t stride,
                                                         uint32_t sx, uint32_t sy, uint32_t row, int qtype) {
    Idx2 i = idx2(sx, sy);
    if (!i.ok) return;
    const size_t di = (size_t)i.y * sx + i.x;
    data[(size_t)i.y * row + i.x] = unquantize_i(unsymbolize_u(sym[di]), step, qtype) + byte_to_float(ref[di * stride]);
}

// RGB -> YCoCg without the reversible lift: chroma keeps 8 bits by dropping its lowest bit
// (YCoCg.cu:23-34).  C integer division truncates toward zero, as here.
template <int C>
__global__ __launch_bounds__(256) void k_rgb_to_ycocg(uint8_t* __restrict__ out, const uint8_t* __restrict__ in,
                                                      uint32_t n) {
    for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const uint8_t* s = in + (size_t)p * C;
        const int r = s[0], g = s[1], b = s[2];
        uint8_t* d = out + (size_t)p * C;
        d[0] = (uint8_t)((r + 2 * g + b + 2) / 4);
        d[1] = (uint8_t)((r - b + 1) / 2 + 127);
        d[2] = (uint8_t)((-r + 2 * g - b + 2) / 4 + 127);
        if (C == 4) d[3] = s[3];
    }
}
template <int C>
__global__ __launch_bounds__(256) void k_ycocg_to_rgb(uint8_t* __restrict__ out, const uint8_t* __restrict__ in,
                                                      uint32_t n) {
    for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const uint8_t* s = in + (size_t)p * C;
        const int y = s[0], co = s[1], cg = s[2];
        uint8_t* d = out + (size_t)p * C;
        d[0] = (uint8_t)min(max(y + co - cg, 0), 255);
        d[1] = (uint8_t)min(max(y + cg - 127, 0), 255);
        d[2] = (uint8_t)min(max(y - co - cg + 254, 0), 255);
        if (C == 4) d[3] = s[3];
    }
}

// ---- bit planes ------------------------------------------------------------------------------
// Word w of the bit array holds bit i = LSB of element 32*w + i (Bits.cu:18-50).  Elements past
// n read as 0; words past ceil(n/32) are not written.
constexpr uint32_t kBitsPerWave = 2048;

// mode 0: getLSB; mode 1: removeLSB (bits optional); shift is added before the bit is taken.
template <int MODE>
__global__ __launch_bounds__(256) void k_lsb_planes(int16_t* __restrict__ data, uint32_t* __restrict__ bits, uint32_t n,
                                                    int shift) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t base = ((uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kBitsPerWave;
    if (base >= n) return;  // wave-uniform
    uint32_t word = 0;
#pragma unroll 4
    for (uint32_t it = 0; it < 32; it++) {
        const uint64_t e = base + it * 64 + lane;
        int16_t v = 0;
        if (e < n) v = (int16_t)(data[e] + shift);
        const uint64_t m = __ballot(v & 1);
        if ((lane >> 1) == it) word = (lane & 1) ? (uint32_t)(m >> 32) : (uint32_t)m;
        if (MODE == 1 && e < n) data[e] = (int16_t)(v >> 1);  // (v + (v >> 15)) / 2 == floor(v / 2)
    }
    const uint64_t w = base / 32 + lane;
    if (bits && w < ((uint64_t)n + 31) / 32) bits[w] = word;
}

__global__ __launch_bounds__(256) void k_append_lsb(int16_t* __restrict__ data, const uint32_t* __restrict__ bits,
                                                    uint32_t n, int shift) {
    for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const int b = (bits[e >> 5] >> (e & 31)) & 1;
        data[e] = (int16_t)((int16_t)((data[e] * 2) | b) + shift);
    }
}
__global__ __launch_bounds__(256) void k_shift_elems(int16_t* __restrict__ data, uint32_t n, int shift) {
    for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) data[e] = (int16_t)(data[e] + shift);
}

dim3 grid2(uint32_t sx, uint32_t sy) { return dim3(div_up(sx, 64), div_up(sy, 4)); }
void check2(uint32_t sx, uint32_t sy, uint32_t channelCount, uint32_t channel) {
    CCA_REQUIRE(div_up(sy, 4) <= 65535u, CC_ERR_INVALID_ARG, "image too large for the launch grid");
    CCA_REQUIRE(channelCount >= 1 && channel < channelCount, CC_ERR_INVALID_ARG, "channel out of range");
}
uint32_t grid_stride_blocks(uint32_t n) { return std::min<uint32_t>(div_up(n, 256), 8192u); }

}  // namespace

void float_to_byte_2d(uint8_t* out, uint32_t channelCount, uint32_t channel, const float* data, uint32_t sx,
                      uint32_t sy, uint32_t row, hipStream_t st) {
    if (row == 0) row = sx;
    check2(sx, sy, channelCount, channel);
    if ((uint64_t)sx * sy == 0) return;
    CCA_REQUIRE(out && data, CC_ERR_INVALID_ARG, "null image");
    CCA_LAUNCH("float_to_byte", (uint64_t)sx * sy * 5, k_float_to_byte, grid2(sx, sy), dim3(256), 0, st, out + channel,
               channelCount, data, sx, sy, row);
    CCA_LAUNCH_CHECK();
}

void byte_to_float_2d(float* data, const uint8_t* in, uint32_t channelCount, uint32_t channel, uint32_t sx,
                      uint32_t sy, uint32_t row, hipStream_t st) {
    if (row == 0) row = sx;
    check2(sx, sy, channelCount, channel);
    if ((uint64_t)sx * sy == 0) return;
    CCA_REQUIRE(in && data, CC_ERR_INVALID_ARG, "null image");
    CCA_LAUNCH("byte_to_float", (uint64_t)sx * sy * 5, k_byte_to_float, grid2(sx, sy), dim3(256), 0, st, data,
               in + channel, channelCount, sx, sy, row);
    CCA_LAUNCH_CHECK();
}

void quantize_difference_2d(uint16_t* sym, float* data, float step, const uint8_t* ref, uint32_t channelCount,
                            uint32_t channel, uint32_t sx, uint32_t sy, uint32_t row, int qtype, bool roundtrip,
                            hipStream_t st) {
    if (row == 0) row = sx;
    CCA_REQUIRE(qtype == 0 || qtype == 1, CC_ERR_INVALID_ARG, "unknown quantizer type");
    check2(sx, sy, channelCount, channel);
    if ((uint64_t)sx * sy == 0) return;
    CCA_REQUIRE(sym && data && ref, CC_ERR_INVALID_ARG, "null image");
    const float inv = 1.0f / step;  // host fp32 division, Quantize.cu:291
    CCA_LAUNCH("quantize_difference", (uint64_t)sx * sy * (roundtrip ? 11 : 7), k_quantize_diff, grid2(sx, sy),
               dim3(256), 0, st, sym, data, step, inv, ref + channel, channelCount, sx, sy, row, qtype,
               roundtrip ? 1 : 0);
    CCA_LAUNCH_CHECK();
}

void unquantize_difference_2d(float* data, const uint16_t* sym, float step, const uint8_t* ref, uint32_t channelCount,
                              uint32_t channel, uint32_t sx, uint32_t sy, uint32_t row, int qtype, hipStream_t st) {
    if (row == 0) row = sx;
    CCA_REQUIRE(qtype == 0 || qtype == 1, CC_ERR_INVALID_ARG, "unknown quantizer type");
    check2(sx, sy, channelCount, channel);
    if ((uint64_t)sx * sy == 0) return;
    CCA_REQUIRE(sym && data && ref, CC_ERR_INVALID_ARG, "null image");
    CCA_LAUNCH("unquantize_difference", (uint64_t)sx * sy * 7, k_unquantize_diff, grid2(sx, sy), dim3(256), 0, st,
               data, sym, step, ref + channel, channelCount, sx, sy, row, qtype);
    CCA_LAUNCH_CHECK();
}

void convert_ycocg(uint8_t* out, const uint8_t* in, uint32_t pixels, int channels, bool to_ycocg, hipStream_t st) {
    CCA_REQUIRE(channels == 3 || channels == 4, CC_ERR_INVALID_ARG, "pixels have 3 or 4 channels");
    if (pixels == 0) return;
    CCA_REQUIRE(out && in, CC_ERR_INVALID_ARG, "null image");
    const dim3 g(grid_stride_blocks(pixels));
    const uint64_t bytes = (uint64_t)pixels * channels * 2;
    if (to_ycocg) {
        if (channels == 3) CCA_LAUNCH("rgb_to_ycocg", bytes, k_rgb_to_ycocg<3>, g, dim3(256), 0, st, out, in, pixels);
        else CCA_LAUNCH("rgb_to_ycocg", bytes, k_rgb_to_ycocg<4>, g, dim3(256), 0, st, out, in, pixels);
    } else {
        if (channels == 3) CCA_LAUNCH("ycocg_to_rgb", bytes, k_ycocg_to_rgb<3>, g, dim3(256), 0, st, out, in, pixels);
        else CCA_LAUNCH("ycocg_to_rgb", bytes, k_ycocg_to_rgb<4>, g, dim3(256), 0, st, out, in, pixels);
    }
    CCA_LAUNCH_CHECK();
}

void lsb_get(const int16_t* data, uint32_t* bits, uint32_t n, hipStream_t st) {
    if (n == 0) return;
    CCA_REQUIRE(data && bits, CC_ERR_INVALID_ARG, "null array");
    const uint32_t waves = div_up(n, kBitsPerWave);
    CCA_LAUNCH("get_lsb", (uint64_t)n * 2 + (uint64_t)div_up(n, 32) * 4, k_lsb_planes<0>, dim3(div_up(waves, 4)),
               dim3(256), 0, st, const_cast<int16_t*>(data), bits, n, 0);
    CCA_LAUNCH_CHECK();
}

void lsb_remove(int16_t* data, uint32_t* bits, uint32_t n, int16_t shift, hipStream_t st) {
    if (n == 0) return;
    CCA_REQUIRE(data, CC_ERR_INVALID_ARG, "null array");
    const uint32_t waves = div_up(n, kBitsPerWave);
    CCA_LAUNCH("remove_lsb", (uint64_t)n * 4 + (bits ? (uint64_t)div_up(n, 32) * 4 : 0), k_lsb_planes<1>,
               dim3(div_up(waves, 4)), dim3(256), 0, st, data, bits, n, (int)shift);
    CCA_LAUNCH_CHECK();
}

void lsb_append(int16_t* data, const uint32_t* bits, uint32_t n, int16_t shift, hipStream_t st) {
    if (n == 0) return;
    CCA_REQUIRE(data && bits, CC_ERR_INVALID_ARG, "null array");
    CCA_LAUNCH("append_lsb", (uint64_t)n * 4 + (uint64_t)div_up(n, 32) * 4, k_append_lsb, dim3(grid_stride_blocks(n)),
               dim3(256), 0, st, data, bits, n, (int)shift);
    CCA_LAUNCH_CHECK();
}

void shift_elems(int16_t* data, uint32_t n, int16_t shift, hipStream_t st) {
    if (n == 0) return;
    CCA_REQUIRE(data, CC_ERR_INVALID_ARG, "null array");
    CCA_LAUNCH("shift_elems", (uint64_t)n * 4, k_shift_elems, dim3(grid_stride_blocks(n)), dim3(256), 0, st, data, n,
               (int)shift);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
