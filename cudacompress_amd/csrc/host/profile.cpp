// Kernel event profiler behind cc_profile_* and the per-instance kernel entries of timing detail 3
// (no counterpart in the reference library, whose CudaTimer sections --
// src/cudaCompress/util/CudaTimer.cpp:65-169 -- time host-side phases per instance,
// src/cudaCompress/InstanceImpl.cpp:164-190).
#include "../profile.hpp"

#include <atomic>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../common.hpp"

namespace cca {
namespace {

struct Pending {
    const char* name;
    hipEvent_t a, b;
    uint64_t bytes;
    const void* owner;  // null: not timed for an owner
    bool global;        // counted in the process-wide accumulator
};
struct Entry {
    std::string name;
    uint64_t launches = 0;
    double ms = 0.0;
    uint64_t bytes = 0;
};
using Acc = std::map<std::string, Entry>;

std::atomic<bool> g_direct{false};  // cc_profile_enable
std::atomic<int> g_votes{0};        // owners that voted (size of g_voters)
std::mutex g_mu;
std::set<const void*> g_voters;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_pool;
Acc g_global;
std::map<const void*, Acc> g_owned;
std::vector<Entry> g_snapshot;
thread_local const void* t_owner = nullptr;

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

void add(Acc& acc, const Pending& p, float ms) {
    Entry& e = acc[p.name];
    e.name = p.name;
    e.launches++;
    e.ms += ms;
    e.bytes += p.bytes;
}

// Resolves every pending launch into the accumulators it belongs to (g_mu held).
void drain_locked() {
    for (auto& p : g_pending) {
        float ms = 0.0f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            if (p.global) add(g_global, p, ms);
            if (p.owner && g_voters.count(p.owner)) add(g_owned[p.owner], p, ms);
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

}  // namespace

bool profiling_enabled() {
    return g_direct.load(std::memory_order_relaxed) || g_votes.load(std::memory_order_relaxed) > 0;
}

ProfileOwner::ProfileOwner(const void* owner) : prev_(t_owner) { t_owner = owner; }
ProfileOwner::~ProfileOwner() { t_owner = prev_; }

void profile_vote(const void* owner, bool on) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (on) g_voters.insert(owner);
    else g_voters.erase(owner);
    g_votes.store((int)g_voters.size());
}

std::vector<std::pair<std::string, double>> profile_owner_entries(const void* owner) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    std::vector<std::pair<std::string, double>> out;
    auto it = g_owned.find(owner);
    if (it != g_owned.end())
        for (auto& kv : it->second) out.emplace_back(kv.first, kv.second.ms);
    return out;
}

void profile_owner_reset(const void* owner) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_owned.erase(owner);
}

void profile_owner_forget(const void* owner) {
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_owned.erase(owner);
    g_voters.erase(owner);
    g_votes.store((int)g_voters.size());
}

KernelTimer::KernelTimer(const char* name, hipStream_t st, uint64_t bytes) : name_(name), st_(st), bytes_(bytes) {
    if (!profiling_enabled()) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return;  // launches recorded into a graph are timed by the graph's replay, not per kernel
    }
    std::lock_guard<std::mutex> lk(g_mu);
    global_ = g_direct.load(std::memory_order_relaxed);
    if (t_owner && g_voters.count(t_owner)) owner_ = t_owner;
    if (!global_ && !owner_) return;
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
    g_pending.push_back(Pending{name_, start_, stop, bytes_, owner_, global_});
}

}  // namespace cca

extern "C" {

void cc_profile_enable(int enable) { cca::g_direct.store(enable != 0); }

void cc_profile_reset(void) {
    using namespace cca;
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();  // owners keep their share of the pending launches
    g_global.clear();
    g_snapshot.clear();
}

uint32_t cc_profile_collect(void) {
    using namespace cca;
    std::lock_guard<std::mutex> lk(g_mu);
    drain_locked();
    g_snapshot.clear();
    for (auto& kv : g_global) g_snapshot.push_back(kv.second);
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
