// wave64 / workgroup primitives for gfx950 (scan, reduce, ballot helpers).
// Written from the standard Kogge-Stone formulation on DPP lane moves.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cca {

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ uint64_t lanes_below(int lane) { return (1ull << lane) - 1ull; }

// Inclusive wave scans on DPP moves (GFX9 data-parallel primitives): Kogge-Stone inside each row
// of 16 lanes (row_shr 1, 2, 4, 8), then row_bcast:15 carries row 0 into row 1 and row 2 into
// row 3, and row_bcast:31 carries lanes 0-31 into rows 2-3.  Lanes that receive nothing in a step
// combine with the identity (`old` operand, bound_ctrl off).
namespace dpp {
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143, kWaveShr1 = 0x138;

template <int Ctl, int RowMask = 0xf>
__device__ __forceinline__ uint32_t mov(uint32_t identity, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, Ctl, RowMask, 0xf, false);
}
template <int Ctl, int RowMask = 0xf>
__device__ __forceinline__ uint64_t mov64(uint64_t identity, uint64_t v) {
    const uint32_t lo = mov<Ctl, RowMask>((uint32_t)identity, (uint32_t)v);
    const uint32_t hi = mov<Ctl, RowMask>((uint32_t)(identity >> 32), (uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
template <class T, class Op>
__device__ __forceinline__ T incl_scan(T v, T identity, Op op) {
    if constexpr (sizeof(T) == 8) {
        v = op(v, (T)mov64<kRowShr1>((uint64_t)identity, (uint64_t)v));
        v = op(v, (T)mov64<kRowShr2>((uint64_t)identity, (uint64_t)v));
        v = op(v, (T)mov64<kRowShr4>((uint64_t)identity, (uint64_t)v));
        v = op(v, (T)mov64<kRowShr8>((uint64_t)identity, (uint64_t)v));
        v = op(v, (T)mov64<kRowBcast15, 0xa>((uint64_t)identity, (uint64_t)v));
        v = op(v, (T)mov64<kRowBcast31, 0xc>((uint64_t)identity, (uint64_t)v));
    } else {
        v = op(v, (T)mov<kRowShr1>((uint32_t)identity, (uint32_t)v));
        v = op(v, (T)mov<kRowShr2>((uint32_t)identity, (uint32_t)v));
        v = op(v, (T)mov<kRowShr4>((uint32_t)identity, (uint32_t)v));
        v = op(v, (T)mov<kRowShr8>((uint32_t)identity, (uint32_t)v));
        v = op(v, (T)mov<kRowBcast15, 0xa>((uint32_t)identity, (uint32_t)v));
        v = op(v, (T)mov<kRowBcast31, 0xc>((uint32_t)identity, (uint32_t)v));
    }
    return v;
}
// the value of lane 63 on every lane
__device__ __forceinline__ uint32_t last_lane(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 63); }
__device__ __forceinline__ uint64_t last_lane64(uint64_t v) {
    return ((uint64_t)last_lane((uint32_t)(v >> 32)) << 32) | last_lane((uint32_t)v);
}
}  // namespace dpp

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    return dpp::incl_scan<uint32_t>(v, 0u, [](uint32_t a, uint32_t b) { return a + b; });
}
__device__ __forceinline__ uint64_t wave_incl_scan_u64(uint64_t v) {
    return dpp::incl_scan<uint64_t>(v, 0ull, [](uint64_t a, uint64_t b) { return a + b; });
}
__device__ __forceinline__ long long wave_incl_max_i64(long long v) {
    return dpp::incl_scan<long long>(v, -1ll, [](long long a, long long b) { return a > b ? a : b; });
}
// the previous lane's value (lane 0: identity)
__device__ __forceinline__ long long wave_shift_up_i64(long long v, long long identity) {
    return (long long)dpp::mov64<dpp::kWaveShr1>((uint64_t)identity, (uint64_t)v);
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return dpp::last_lane(wave_incl_scan(v)); }
__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) { return dpp::last_lane64(wave_incl_scan_u64(v)); }
__device__ __forceinline__ uint32_t wave_max(uint32_t v) {
    return dpp::last_lane(dpp::incl_scan<uint32_t>(v, 0u, [](uint32_t a, uint32_t b) { return a > b ? a : b; }));
}

// Exclusive scan over a workgroup of NT threads (NT multiple of 64).  `total` receives the sum.
// Uses NT/64 words of LDS scratch; contains barriers, so every thread must call it.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* scratch, uint32_t& total) {
    constexpr int NW = NT / 64;
    const int lane = lane_id(), wave = threadIdx.x >> 6;
    uint32_t incl = wave_incl_scan(v);
    if (lane == 63) scratch[wave] = incl;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        uint32_t s = scratch[w];
        if (w < wave) base += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + incl - v;
}

// The dispatcher hands workgroups to the 8 XCDs round robin (workgroup b runs on XCD b % 8), and
// every XCD has its own L2.  xcd_order(b, n) renumbers the n workgroups of a launch so that each
// XCD runs one contiguous range of logical workgroups: neighbours in logical order, which read the
// same tables or lines, then share an L2 instead of fetching them once per XCD.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t n) {
    const uint32_t x = b % kXcds, k = b / kXcds;
    const uint32_t q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + k;
}

// Sum over a workgroup of NT threads (NT multiple of 64), every thread gets it; NT/64 words of
// LDS scratch; contains barriers, so every thread must call it.
template <int NT>
__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* scratch) {
    constexpr int NW = NT / 64;
    const uint64_t w = wave_sum_u64(v);
    if (lane_id() == 0) scratch[threadIdx.x >> 6] = w;
    __syncthreads();
    uint64_t tot = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) tot += scratch[k];
    __syncthreads();
    return tot;
}

}  // namespace cca
