#include "timing.hpp"

#include "instance.hpp"

namespace cca {

void Timings::collect() {
    for (Pending& p : pending) {
        CCA_HIP(hipEventSynchronize(p.stop));
        float t = 0.0f;
        CCA_HIP(hipEventElapsedTime(&t, p.start, p.stop));
        if (!ms.count(p.name)) order.push_back(p.name);
        ms[p.name] += t;
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    pending.clear();
}

void Timings::reset() {
    collect();
    order.clear();
    ms.clear();
}

Timings::~Timings() {
    for (Pending& p : pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
}

SectionTimer::SectionTimer(cc_instance* I, int level) {
    if (I->timing_detail >= level) {
        t_ = &I->timings;
        st_ = I->stream;
    }
}

SectionTimer::~SectionTimer() {
    try {
        close();
    } catch (...) {
    }
}

void SectionTimer::close() {
    if (!start_) return;
    hipEvent_t stop = nullptr;
    CCA_HIP(hipEventCreate(&stop));
    CCA_HIP(hipEventRecord(stop, st_));
    t_->pending.push_back({name_, start_, stop});
    start_ = nullptr;
}

void SectionTimer::operator()(const char* name) {
    if (!t_) return;
    close();
    CCA_HIP(hipEventCreate(&start_));
    CCA_HIP(hipEventRecord(start_, st_));
    name_ = name;
}

}  // namespace cca
