// Scalar quantizer and symbol mapping, src/cudaCompress/util/Quantize.cuh:55-147; shared by
// quantize.hip and the fused brick passes (bricks.hip).  Built with -ffp-contract=off so
// `v * inv + s * 0.5f` rounds the product before the add exactly as written.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cca {
namespace {

__device__ __forceinline__ int to_int_sat(float v) {  // v_cvt_i32_f32 semantics, NaN -> 0
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int)0x80000000u;
    return (int)v;
}
__device__ __forceinline__ float sign_f(float v) { return v > 0.0f ? 1.0f : (v < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ uint32_t symbolize_i(int v) {  // Quantize.cuh:55-60
    return 2u * (uint32_t)abs(v) - (v < 0 ? 1u : 0u);
}
__device__ __forceinline__ int unsymbolize_u(uint32_t s) {  // Quantize.cuh:62-67
    uint32_t neg = s & 1u;
    return (int)((1u - 2u * neg) * ((s + neg) / 2u));
}
__device__ __forceinline__ int quantize_f(float v, float inv, int qtype) {  // Quantize.cuh:76-93
    if (qtype == 0) return to_int_sat(v * inv);
    float t = v * inv;
    return to_int_sat(t + sign_f(v) * 0.5f);
}
__device__ __forceinline__ float unquantize_i(int q, float step, int qtype) {  // Quantize.cuh:102-120
    if (qtype == 0) {
        float sg = q < 0 ? -1.0f : (q > 0 ? 1.0f : 0.0f);
        return (float(q) + sg * 0.5f) * step;
    }
    return float(q) * step;
}

}  // namespace
}  // namespace cca
