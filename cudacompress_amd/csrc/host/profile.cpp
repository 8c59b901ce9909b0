// Kernel event profiler behind cc_profile_* (no counterpart in the reference library, whose
// CudaTimer sections -- src/cudaCompress/util/CudaTimer.cpp:65-169 -- time host-side phases).
#include "../profile.hpp"

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../common.hpp"

namespace cca {
namespace {

struct Pending {
    const char* name;
    hipEvent_t a, b;
    uint64_t bytes;
};
struct Entry {
    std::string name;
    uint64_t launches = 0;
    double ms = 0.0;
    uint64_t bytes = 0;
};

std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
std::map<std::string, Entry> g_acc;
std::vector<Entry> g_snapshot;

hipEvent_t take_event() {
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

bool profiling_enabled() { return g_on.load(std::memory_order_relaxed); }

KernelTimer::KernelTimer(const char* name, hipStream_t st, uint64_t bytes) : name_(name), st_(st), bytes_(bytes) {
    if (!profiling_enabled()) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return;  // launches recorded into a graph are timed by the graph's replay, not per kernel
    }
    std::lock_guard<std::mutex> lk(g_mu);
    start_ = take_event();
    if (start_) (void)hipEventRecord(start_, st_);
}

KernelTimer::~KernelTimer() {
    if (!start_) return;
    std::lock_guard<std::mutex> lk(g_mu);
    hipEvent_t stop = take_event();
    if (!stop) {
        g_pool.push_back(start_);
        return;
    }
    (void)hipEventRecord(stop, st_);
    g_pending.push_back(Pending{name_, start_, stop, bytes_});
}

}  // namespace cca

extern "C" {

void cc_profile_enable(int enable) { cca::g_on.store(enable != 0); }

void cc_profile_reset(void) {
    using namespace cca;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& p : g_pending) {
        (void)hipEventSynchronize(p.b);
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
    g_acc.clear();
    g_snapshot.clear();
}

uint32_t cc_profile_collect(void) {
    using namespace cca;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& p : g_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Entry& e = g_acc[p.name];
            e.name = p.name;
            e.launches++;
            e.ms += ms;
            e.bytes += p.bytes;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
    g_snapshot.clear();
    for (auto& kv : g_acc) g_snapshot.push_back(kv.second);
    return (uint32_t)g_snapshot.size();
}

cc_status cc_profile_entry(uint32_t i, const char** name, uint64_t* launches, double* total_ms,
                           uint64_t* algorithmic_bytes) {
    using namespace cca;
    std::lock_guard<std::mutex> lk(g_mu);
    if (i >= g_snapshot.size() || !name || !launches || !total_ms || !algorithmic_bytes) return CC_ERR_INVALID_ARG;
    *name = g_snapshot[i].name.c_str();
    *launches = g_snapshot[i].launches;
    *total_ms = g_snapshot[i].ms;
    *algorithmic_bytes = g_snapshot[i].bytes;
    return CC_OK;
}

}  // extern "C"
