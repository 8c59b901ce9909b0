// Per-kernel hipEvent timing (off unless cc_profile_enable(1)).  Each launch site wraps its
// launch in a KernelTimer carrying the kernel name and its algorithmic byte count (one read of
// its inputs plus one write of its outputs); cc_profile_collect() resolves the events.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace cca {

bool profiling_enabled();

class KernelTimer {
public:
    KernelTimer(const char* name, hipStream_t st, uint64_t bytes);
    ~KernelTimer();

private:
    const char* name_;
    hipStream_t st_;
    uint64_t bytes_;
    hipEvent_t start_ = nullptr;
};

}  // namespace cca
