// Per-kernel hipEvent timing.  Each launch site wraps its launch in a KernelTimer carrying the
// kernel name and its algorithmic byte count (one read of its inputs plus one write of its
// outputs).  A launch is timed when the process-wide switch (cc_profile_enable) is on, or when the
// API call that made it belongs to an owner (an instance at timing detail 3) that voted for kernel
// timing.  Records go to the process-wide accumulator (cc_profile_*) when the switch was on at
// launch, and to their owner's accumulator when the owner had voted; each side reads and clears
// only its own records.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace cca {

bool profiling_enabled();

// Launches made by this thread while a ProfileOwner is alive are tagged with its owner (the
// instance whose API call made them).  Scopes nest; the innermost wins.
class ProfileOwner {
public:
    explicit ProfileOwner(const void* owner);
    ~ProfileOwner();
    ProfileOwner(const ProfileOwner&) = delete;
    ProfileOwner& operator=(const ProfileOwner&) = delete;

private:
    const void* prev_;
};

// An owner's vote for kernel timing (an instance entering or leaving timing detail 3).  Votes
// of different owners and the process-wide switch are independent.
void profile_vote(const void* owner, bool on);
// Drains finished launches and returns the owner's per-kernel totals (name, ms) in name order.
std::vector<std::pair<std::string, double>> profile_owner_entries(const void* owner);
// Clears the owner's accumulated records (other owners and the process-wide records stay).
void profile_owner_reset(const void* owner);
// Withdraws the owner's vote and drops its records (instance destruction).
void profile_owner_forget(const void* owner);

class KernelTimer {
public:
    KernelTimer(const char* name, hipStream_t st, uint64_t bytes);
    ~KernelTimer();

private:
    const char* name_;
    hipStream_t st_;
    uint64_t bytes_;
    hipEvent_t start_ = nullptr;
    const void* owner_ = nullptr;  // set when the owner voted
    bool global_ = false;          // the process-wide switch was on
};

}  // namespace cca
