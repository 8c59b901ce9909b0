// Per-line CDF 9/7 and CDF 5/3 operations shared by the single-volume passes (dwt.hip) and the
// brick-batched passes (bricks.hip).  Math: src/cudaCompress/util/DWTFloatKernels.cui:14-81
// (9/7 analysis), :433-503 (synthesis), src/cudaCompress/util/DWTFloatFilters.cuh:42-49
// (coefficients), src/cudaCompress/util/DWTIntKernels.cui:15-100, :597-687 (5/3 lifting),
// src/cudaCompress/util/DWTCommon.cuh:54-142 (extensions, div2/div4).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cca {
namespace {

__constant__ float c_aL[9] = {0.0267488f, -0.0168641f, -0.0782233f, 0.2668641f, 0.6029490f,
                              0.2668641f, -0.0782233f, -0.0168641f, 0.0267488f};
__constant__ float c_aH[9] = {0.0f, 0.0456359f, -0.0287718f, -0.2956359f, 0.5575435f,
                              -0.2956359f, -0.0287718f, 0.0456359f, 0.0f};
__constant__ float c_gE[9] = {0.0f, 0.0337282f, -0.0575435f, -0.5337281f, 1.1150871f,
                              -0.5337281f, -0.0575435f, 0.0337282f, 0.0f};
__constant__ float c_gO[9] = {0.0534975f, -0.0912718f, -0.1564465f, 0.5912718f, 1.2058980f,
                              0.5912718f, -0.1564465f, -0.0912718f, 0.0534975f};

// Whole-sample symmetric extension (mirrorLeft/mirrorRight, DWTCommon.cuh:55-56), repeated so
// that lines shorter than the filter stay in bounds.
__device__ __forceinline__ int mirror(int j, int n) {
    if (n <= 1) return 0;
    while (j < 0 || j >= n) {
        if (j < 0) j = -j;
        if (j >= n) j = 2 * (n - 1) - j;
    }
    return j;
}
// Extensions of the lowpass / highpass halves on synthesis (mirrorLeft/mirrorRightRepeat and
// mirrorLeftRepeat/mirrorRight as used at DWTFloatKernels.cui:456-476).
__device__ __forceinline__ int ext_low(int g, int m) {
    while (g < 0 || g >= m) {
        if (g < 0) g = -g;
        if (g >= m) g = 2 * m - 1 - g;
    }
    return g;
}
__device__ __forceinline__ int ext_high(int g, int m) {
    while (g < 0 || g >= m) {
        if (g < 0) g = -g - 1;
        if (g >= m) g = 2 * m - 2 - g;
    }
    return g;
}

__device__ __forceinline__ int16_t div2(int16_t v) {
    v = (int16_t)(v + ((v >= 0) ? 1 : -1));
    return (int16_t)(v / 2);
}
__device__ __forceinline__ int16_t div4(int16_t v) {
    v = (int16_t)(v + ((v >= 0) ? 2 : -2));
    return (int16_t)(v / 4);
}

// ---- per-axis line operations ----------------------------------------------------------------
// Forward ops read a window of extended input positions [2*k0 - H, 2*k0 + 2P + H) and produce
// pairs (low[k], high[k]).  Inverse ops read interleaved pairs [k0 - G, k0 + P + G) and produce
// outputs [2*k0, 2*k0 + 2P).
struct Fwd97 {
    using T = float;
    static constexpr int H = 4;
    __device__ static void pair(const T* w, int q, T& lo, T& hi) {  // w indexed from 2*k0 - H
        const T* x = w + 2 * q;
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < 9; i++) s = __builtin_fmaf(c_aL[i], x[i], s);
        lo = s;
        s = 0.0f;
#pragma unroll
        for (int i = 0; i < 9; i++) s = __builtin_fmaf(c_aH[i], x[i + 1], s);
        hi = s;
    }
};
// FP64 twin of Fwd97 (no reference counterpart): the decimal filter coefficients of
// DWTFloatFilters.cuh:42-49 at double precision, fused multiply-adds in the same filter order.
// It measures the error of the FP32 transform; nothing on the coding path uses it.
__constant__ double c_aLd[9] = {0.0267488, -0.0168641, -0.0782233, 0.2668641, 0.6029490,
                                0.2668641, -0.0782233, -0.0168641, 0.0267488};
__constant__ double c_aHd[9] = {0.0, 0.0456359, -0.0287718, -0.2956359, 0.5575435,
                                -0.2956359, -0.0287718, 0.0456359, 0.0};
struct Fwd97d {
    using T = double;
    static constexpr int H = 4;
    __device__ static void pair(const T* w, int q, T& lo, T& hi) {
        const T* x = w + 2 * q;
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++) s = __builtin_fma(c_aLd[i], x[i], s);
        lo = s;
        s = 0.0;
#pragma unroll
        for (int i = 0; i < 9; i++) s = __builtin_fma(c_aHd[i], x[i + 1], s);
        hi = s;
    }
};
struct Inv97 {
    using T = float;
    static constexpr int G = 2;
    // x: the 9 interleaved taps centred on the output (x[0] = v[j-4]); odd: output index parity
    __device__ static T out(const T* x, int odd) {
        const float* g = odd ? c_gO : c_gE;
        float s = 0.0f;
#pragma unroll
        for (int i = 0; i < 9; i++) s = __builtin_fmaf(g[i], x[i], s);
        return s;
    }
};
struct Fwd53 {
    using T = int16_t;
    static constexpr int H = 2;
    __device__ static void pair(const T* w, int q, T& lo, T& hi) {
        const T* x = w + 2 * q;  // x[0] = in[2k-2] ... x[4] = in[2k+2]
        int16_t dl = (int16_t)(x[1] - div2((int16_t)(x[0] + x[2])));
        int16_t d = (int16_t)(x[3] - div2((int16_t)(x[2] + x[4])));
        lo = (int16_t)(x[2] + div4((int16_t)(dl + d)));
        hi = d;
    }
};
struct Inv53 {
    using T = int16_t;
    static constexpr int G = 1;
    // v[0..1] = (s, d) of pair k-1, v[2..3] pair k, v[4..5] pair k+1 (extended);
    // s'(g) = s[EL(g)] - div4(d[EH(g-1)] + d[EH(g)])
    __device__ static void pair_out(const T* v, T& even, T& odd) {
        int16_t sp0 = (int16_t)(v[2] - div4((int16_t)(v[1] + v[3])));
        int16_t sp1 = (int16_t)(v[4] - div4((int16_t)(v[3] + v[5])));
        even = sp0;
        odd = (int16_t)(v[3] + div2((int16_t)(sp0 + sp1)));
    }
};

}  // namespace
}  // namespace cca
