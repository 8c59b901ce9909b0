// Drop-in replacement of include/cudaCompress/Instance.h (m0bl0/cudaCompress): header-only
// wrappers over the C ABI of include/cudacompress_amd.h.  Instance is an opaque handle; streams
// are hipStream_t.  Both setInstanceCudaStream (declared by the reference header) and
// setInstanceStream (defined by the reference library, Instance.cpp:59-67) are provided.
#ifndef CUDACOMPRESS_AMD_DROPIN_INSTANCE_H
#define CUDACOMPRESS_AMD_DROPIN_INSTANCE_H

#include <cudaCompress/global.h>
#include <cudacompress_amd.h>

#include <hip/hip_runtime_api.h>

namespace cudaCompress {

class Instance;

namespace detail {
inline cc_instance* raw(Instance* p) { return reinterpret_cast<cc_instance*>(p); }
inline const cc_instance* raw(const Instance* p) { return reinterpret_cast<const cc_instance*>(p); }
}  // namespace detail

// createInstance, Instance.h:26-31: nullptr on failure (cc_last_error() has the reason)
inline Instance* createInstance(int cudaDevice, uint streamCountMax, uint elemCountPerStreamMax,
                                uint codingBlockSize = 0, uint log2HuffmanDistinctSymbolCountMax = 0) {
    cc_instance* p = nullptr;
    if (cc_create_instance(cudaDevice, streamCountMax, elemCountPerStreamMax, codingBlockSize,
                           log2HuffmanDistinctSymbolCountMax, &p) != CC_OK)
        return nullptr;
    return reinterpret_cast<Instance*>(p);
}
inline void destroyInstance(Instance* pInstance) { cc_destroy_instance(detail::raw(pInstance)); }

inline int getInstanceCudaDevice(const Instance* p) { return cc_get_instance_device(detail::raw(p)); }
inline uint getInstanceStreamCountMax(const Instance* p) { return cc_get_instance_stream_count_max(detail::raw(p)); }
inline uint getInstanceElemCountPerStreamMax(const Instance* p) {
    return cc_get_instance_elem_count_per_stream_max(detail::raw(p));
}
inline uint getInstanceCodingBlockSize(const Instance* p) { return cc_get_instance_coding_block_size(detail::raw(p)); }
inline uint getInstanceLog2HuffmanDistinctSymbolCountMax(const Instance* p) {
    return cc_get_instance_log2_huffman_distinct_symbol_count_max(detail::raw(p));
}
inline bool getInstanceUseLongSymbols(const Instance* p) { return cc_get_instance_use_long_symbols(detail::raw(p)) != 0; }

inline void setInstanceCudaStream(Instance* p, hipStream_t str) { cc_set_instance_stream(detail::raw(p), (void*)str); }
inline hipStream_t getInstanceCudaStream(const Instance* p) { return (hipStream_t)cc_get_instance_stream(detail::raw(p)); }
inline void setInstanceStream(Instance* p, hipStream_t str) { setInstanceCudaStream(p, str); }
inline hipStream_t getInstanceStream(const Instance* p) { return getInstanceCudaStream(p); }

}  // namespace cudaCompress

#endif
