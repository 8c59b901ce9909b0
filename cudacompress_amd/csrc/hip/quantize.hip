// Quantizer and symbol mapping for gfx950, src/cudaCompress/util/Quantize.cuh:55-147 and
// src/cudaCompress/util/QuantizeKernels.cui:39-172.  Elementwise, so each kernel is one coalesced
// read and one coalesced write per element: a 64x4 workgroup covers 64 consecutive x of 4 rows.
// The library is built with -ffp-contract=off, so `v * inv + s * 0.5f` rounds the product
// before the add exactly as written.
#include "../kernels.hpp"
#include "quant_ops.hpp"

namespace cca {
namespace {

struct Idx {
    uint32_t x, y, z;
    bool ok;
};
__device__ __forceinline__ Idx idx3(uint32_t sx, uint32_t sy, uint32_t sz) {
    Idx r;
    r.x = blockIdx.x * 64 + (threadIdx.x & 63);
    r.y = blockIdx.y * 4 + (threadIdx.x >> 6);
    r.z = blockIdx.z;
    r.ok = r.x < sx && r.y < sy && r.z < sz;
    return r;
}

__global__ __launch_bounds__(256) void k_symbolize(uint16_t* __restrict__ sym, const int16_t* __restrict__ data,
                                                   uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row, uint32_t slice) {
    Idx i = idx3(sx, sy, sz);
    if (!i.ok) return;
    sym[i.x + (size_t)i.y * sx + (size_t)i.z * sx * sy] = (uint16_t)symbolize_i(data[i.x + (size_t)i.y * row + (size_t)i.z * slice]);
}
__global__ __launch_bounds__(256) void k_unsymbolize(int16_t* __restrict__ data, const uint16_t* __restrict__ sym,
                                                     uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row, uint32_t slice) {
    Idx i = idx3(sx, sy, sz);
    if (!i.ok) return;
    data[i.x + (size_t)i.y * row + (size_t)i.z * slice] = (int16_t)unsymbolize_u(sym[i.x + (size_t)i.y * sx + (size_t)i.z * sx * sy]);
}
template <typename S>
__global__ __launch_bounds__(256) void k_quantize(S* __restrict__ sym, float* __restrict__ data, float step, float inv,
                                                  uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row, uint32_t slice,
                                                  int qtype, int roundtrip) {
    Idx i = idx3(sx, sy, sz);
    if (!i.ok) return;
    size_t si = i.x + (size_t)i.y * row + (size_t)i.z * slice;
    int q = quantize_f(data[si], inv, qtype);
    sym[i.x + (size_t)i.y * sx + (size_t)i.z * sx * sy] = (S)symbolize_i(q);
    if (roundtrip) data[si] = unquantize_i(q, step, qtype);
}
template <typename S>
__global__ __launch_bounds__(256) void k_unquantize(float* __restrict__ data, const S* __restrict__ sym, float step,
                                                    uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row, uint32_t slice,
                                                    int qtype) {
    Idx i = idx3(sx, sy, sz);
    if (!i.ok) return;
    data[i.x + (size_t)i.y * row + (size_t)i.z * slice] =
        unquantize_i(unsymbolize_u((uint32_t)sym[i.x + (size_t)i.y * sx + (size_t)i.z * sx * sy]), step, qtype);
}
__global__ __launch_bounds__(256) void k_quantize_short(int16_t* __restrict__ q, const float* __restrict__ data, float inv,
                                                        uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row, uint32_t slice,
                                                        int qtype) {
    Idx i = idx3(sx, sy, sz);
    if (!i.ok) return;
    q[i.x + (size_t)i.y * sx + (size_t)i.z * sx * sy] = (int16_t)quantize_f(data[i.x + (size_t)i.y * row + (size_t)i.z * slice], inv, qtype);
}
__global__ __launch_bounds__(256) void k_unquantize_short(float* __restrict__ data, const int16_t* __restrict__ q, float step,
                                                          uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row, uint32_t slice,
                                                          int qtype) {
    Idx i = idx3(sx, sy, sz);
    if (!i.ok) return;
    data[i.x + (size_t)i.y * row + (size_t)i.z * slice] = unquantize_i(q[i.x + (size_t)i.y * sx + (size_t)i.z * sx * sy], step, qtype);
}

// max |v| over n floats: grid-stride, wave64 reduction through lane shuffles, one
// atomicMax on the float's bit pattern per workgroup (non-negative floats order as integers).
__global__ __launch_bounds__(256) void k_max_abs(const float* __restrict__ v, uint32_t n, uint32_t* __restrict__ out) {
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) m = fmaxf(m, fabsf(v[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    __shared__ float s[4];
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(s[0], s[1]), fmaxf(s[2], s[3]));
        atomicMax(out, __float_as_uint(m));
    }
}
__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

dim3 grid3(uint32_t sx, uint32_t sy, uint32_t sz) { return dim3(div_up(sx, 64), div_up(sy, 4), sz); }
void check3(uint32_t sx, uint32_t sy, uint32_t sz) {
    CCA_REQUIRE(sz <= 65535u && div_up(sy, 4) <= 65535u, CC_ERR_INVALID_ARG, "array too large for the launch grid");
}

}  // namespace

void symbolize(uint16_t* sym, const int16_t* data, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row,
               uint32_t slice, hipStream_t st) {
    if (row == 0) row = sx;
    if (slice == 0) slice = row * sy;
    if ((uint64_t)sx * sy * sz == 0) return;
    check3(sx, sy, sz);
    CCA_LAUNCH("symbolize", (uint64_t)sx * sy * sz * 4, k_symbolize, grid3(sx, sy, sz), dim3(256), 0, st, sym, data, sx, sy, sz, row, slice);
    CCA_LAUNCH_CHECK();
}

void unsymbolize(int16_t* data, const uint16_t* sym, uint32_t sx, uint32_t sy, uint32_t sz, uint32_t row,
                 uint32_t slice, hipStream_t st) {
    if (row == 0) row = sx;
    // The reference leaves slicePitchDst at 0 when not given (src/cudaCompress/util/Quantize.cu:64-65),
    // which writes every z slice onto slice 0; the dense default is used here instead.
    if (slice == 0) slice = row * sy;
    if ((uint64_t)sx * sy * sz == 0) return;
    check3(sx, sy, sz);
    CCA_LAUNCH("unsymbolize", (uint64_t)sx * sy * sz * 4, k_unsymbolize, grid3(sx, sy, sz), dim3(256), 0, st, data, sym, sx, sy, sz, row, slice);
    CCA_LAUNCH_CHECK();
}

void quantize_to_symbols(void* sym, int symSize, float* data, bool roundtrip, uint32_t sx, uint32_t sy, uint32_t sz,
                         float step, uint32_t row, uint32_t slice, int qtype, hipStream_t st) {
    if (row == 0) row = sx;
    if (slice == 0) slice = row * sy;
    CCA_REQUIRE(qtype == 0 || qtype == 1, CC_ERR_INVALID_ARG, "unknown quantizer type");
    CCA_REQUIRE(symSize == 2 || symSize == 4, CC_ERR_INVALID_ARG, "symbolSize must be 2 or 4");
    if ((uint64_t)sx * sy * sz == 0) return;
    check3(sx, sy, sz);
    const float inv = 1.0f / step;  // host fp32 division, Quantize.cu:94
    if (symSize == 2)
        CCA_LAUNCH("quantize", (uint64_t)sx * sy * sz * (roundtrip ? 10 : 6), k_quantize<uint16_t>, grid3(sx, sy, sz), dim3(256), 0, st, (uint16_t*)sym, data, step, inv,
                           sx, sy, sz, row, slice, qtype, roundtrip ? 1 : 0);
    else
        CCA_LAUNCH("quantize", (uint64_t)sx * sy * sz * (roundtrip ? 12 : 8), k_quantize<uint32_t>, grid3(sx, sy, sz), dim3(256), 0, st, (uint32_t*)sym, data, step, inv,
                           sx, sy, sz, row, slice, qtype, roundtrip ? 1 : 0);
    CCA_LAUNCH_CHECK();
}

void unquantize_from_symbols(float* data, const void* sym, int symSize, uint32_t sx, uint32_t sy, uint32_t sz,
                             float step, uint32_t row, uint32_t slice, int qtype, hipStream_t st) {
    if (row == 0) row = sx;
    if (slice == 0) slice = row * sy;
    CCA_REQUIRE(qtype == 0 || qtype == 1, CC_ERR_INVALID_ARG, "unknown quantizer type");
    CCA_REQUIRE(symSize == 2 || symSize == 4, CC_ERR_INVALID_ARG, "symbolSize must be 2 or 4");
    if ((uint64_t)sx * sy * sz == 0) return;
    check3(sx, sy, sz);
    if (symSize == 2)
        CCA_LAUNCH("unquantize", (uint64_t)sx * sy * sz * 6, k_unquantize<uint16_t>, grid3(sx, sy, sz), dim3(256), 0, st, data, (const uint16_t*)sym,
                           step, sx, sy, sz, row, slice, qtype);
    else
        CCA_LAUNCH("unquantize", (uint64_t)sx * sy * sz * 8, k_unquantize<uint32_t>, grid3(sx, sy, sz), dim3(256), 0, st, data, (const uint32_t*)sym,
                           step, sx, sy, sz, row, slice, qtype);
    CCA_LAUNCH_CHECK();
}

void quantize_to_short(int16_t* q, const float* data, uint32_t sx, uint32_t sy, uint32_t sz, float step, uint32_t row,
                       uint32_t slice, int qtype, hipStream_t st) {
    if (row == 0) row = sx;
    if (slice == 0) slice = row * sy;
    CCA_REQUIRE(qtype == 0 || qtype == 1, CC_ERR_INVALID_ARG, "unknown quantizer type");
    if ((uint64_t)sx * sy * sz == 0) return;
    check3(sx, sy, sz);
    CCA_LAUNCH("quantize_short", (uint64_t)sx * sy * sz * 6, k_quantize_short, grid3(sx, sy, sz), dim3(256), 0, st, q, data, 1.0f / step, sx, sy, sz, row,
                       slice, qtype);
    CCA_LAUNCH_CHECK();
}

void unquantize_from_short(float* data, const int16_t* q, uint32_t sx, uint32_t sy, uint32_t sz, float step,
                           uint32_t row, uint32_t slice, int qtype, hipStream_t st) {
    if (row == 0) row = sx;
    if (slice == 0) slice = row * sy;
    CCA_REQUIRE(qtype == 0 || qtype == 1, CC_ERR_INVALID_ARG, "unknown quantizer type");
    if ((uint64_t)sx * sy * sz == 0) return;
    check3(sx, sy, sz);
    CCA_LAUNCH("unquantize_short", (uint64_t)sx * sy * sz * 6, k_unquantize_short, grid3(sx, sy, sz), dim3(256), 0, st, data, q, step, sx, sy, sz, row, slice,
                       qtype);
    CCA_LAUNCH_CHECK();
}

void max_abs(const float* data, uint32_t n, float* out, hipStream_t st) {
    CCA_LAUNCH("max_abs_init", 4, k_zero_u32, dim3(1), dim3(1), 0, st, (uint32_t*)out);
    CCA_LAUNCH_CHECK();
    if (n == 0) return;
    uint32_t blocks = std::min<uint32_t>(div_up(n, 256 * 8), 2048u);
    CCA_LAUNCH("max_abs", (uint64_t)n * 4, k_max_abs, dim3(blocks), dim3(256), 0, st, data, n, (uint32_t*)out);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
