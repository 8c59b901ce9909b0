// Per-instance section timing (the Timing API, include/cudaCompress/Timing.h:15-27): named,
// consecutive sections bracketed by hipEvents on the instance stream and accumulated by name in
// the order they first appear (CudaTimerResources / CudaScopedTimer,
// src/cudaCompress/util/CudaTimer.cpp).  Detail levels: 1 whole coder calls, 2 their stages,
// 3 additionally every kernel launch (through the per-kernel event timer of profile.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

struct cc_instance;

namespace cca {

struct Timings {
    struct Pending {
        std::string name;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<std::string> order;
    std::map<std::string, double> ms;
    std::vector<std::pair<std::string, float>> view;  // last cc_get_timings result
    void collect();  // waits for the pending events and adds their times
    void reset();
    ~Timings();
};

class SectionTimer {
public:
    SectionTimer(cc_instance* I, int level);
    ~SectionTimer();
    void operator()(const char* name);  // closes the open section, opens `name`

private:
    void close();
    Timings* t_ = nullptr;
    hipStream_t st_ = nullptr;
    std::string name_;
    hipEvent_t start_ = nullptr;
};

}  // namespace cca
