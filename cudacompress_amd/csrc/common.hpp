// Shared definitions for the cudacompress_amd library (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/cudacompress_amd.h"

namespace cca {

// Thread-local error text behind cc_last_error().
void set_error(const std::string& msg);
void clear_error();

struct Error {
    cc_status status;
    std::string msg;
};

#define CCA_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            throw ::cca::Error{CC_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                  " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
        }                                                                                      \
    } while (0)

#define CCA_REQUIRE(cond, status, text)                                  \
    do {                                                                 \
        if (!(cond)) throw ::cca::Error{(status), std::string(text)};    \
    } while (0)

#define CCA_LAUNCH_CHECK() CCA_HIP(hipGetLastError())

// Launch bracketed by the kernel timer (see profile.hpp): name and algorithmic bytes of the launch.
#define CCA_LAUNCH(name, bytes, kern, grid, block, shm, st, ...)                 \
    do {                                                                         \
        ::cca::KernelTimer kt_((name), (st), (uint64_t)(bytes));                 \
        hipLaunchKernelGGL(kern, grid, block, shm, st, __VA_ARGS__);             \
    } while (0)

__host__ __device__ inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
__host__ __device__ inline uint32_t div_up_dev(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
inline uint64_t align_up(uint64_t a, uint64_t b) { return (a + b - 1) / b * b; }

// Number of bits needed to store maxValue, src/cudaCompress/util.h:23-32.
inline uint32_t required_bits(uint32_t v) {
    uint32_t b = 0;
    while (v > 0) { b++; v >>= 1; }
    return b;
}

constexpr uint32_t kMaxCodewordBits = 32;  // include/cudaCompress/EncodeCommon.h:24-25
constexpr uint32_t kZeroCountMax = 255;    // include/cudaCompress/EncodeCommon.h:31

}  // namespace cca
