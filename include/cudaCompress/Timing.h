// Drop-in replacement of include/cudaCompress/Timing.h (m0bl0/cudaCompress): header-only
// wrappers over cc_set_timing_detail / cc_get_timings (include/cudacompress_amd.h).  Sections are
// hipEvent pairs on the instance stream; TIMING_DETAIL_HIGH adds every kernel launch
// (process-wide).
#ifndef CUDACOMPRESS_AMD_DROPIN_TIMING_H
#define CUDACOMPRESS_AMD_DROPIN_TIMING_H

#include <cudaCompress/Instance.h>
#include <cudaCompress/global.h>
#include <cudacompress_amd.h>

#include <string>
#include <vector>

namespace cudaCompress {

enum ETimingDetail {
    TIMING_DETAIL_NONE,
    TIMING_DETAIL_LOW,
    TIMING_DETAIL_MEDIUM,
    TIMING_DETAIL_HIGH,
};

inline void setTimingDetail(Instance* pInstance, ETimingDetail detail) {
    cc_set_timing_detail(detail::raw(pInstance), (int)detail);
}
inline void getTimings(Instance* pInstance, std::vector<std::string>& names, std::vector<float>& times) {
    uint32_t n = 0;
    if (cc_get_timings(detail::raw(pInstance), &n) != CC_OK) return;
    for (uint32_t i = 0; i < n; i++) {
        const char* name = nullptr;
        float ms = 0.0f;
        if (cc_get_timing_entry(detail::raw(pInstance), i, &name, &ms) != CC_OK) continue;
        names.push_back(name);
        times.push_back(ms);
    }
}
inline void printTimings(Instance* pInstance) { cc_print_timings(detail::raw(pInstance)); }
inline void resetTimings(Instance* pInstance) { cc_reset_timings(detail::raw(pInstance)); }

}  // namespace cudaCompress

#endif
