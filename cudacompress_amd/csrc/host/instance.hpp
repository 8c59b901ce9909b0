// cc_instance: device binding, parameters and the reusable device/host buffers of the coder
// (the role of src/cudaCompress/InstanceImpl.h:27-189).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../coder.hpp"
#include "../common.hpp"
#include "timing.hpp"

namespace cca {

// Grow-only device buffer (contents are not preserved across growth).
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t bytes);
    template <class T>
    T* as(size_t count) { return static_cast<T*>(get(count * sizeof(T))); }
    void release();
};

// Grow-only pinned host buffer.
struct PinBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t bytes);
    template <class T>
    T* as(size_t count) { return static_cast<T*>(get(count * sizeof(T))); }
    void release();
};

// Scoped device selection: makes `dev` current and restores the previous device.
struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev);
    ~DeviceGuard();
};

constexpr int kDecodeRing = 4;  // decode staging slots (the reference keeps 8, InstanceImpl.h:108)

// Buffers of the device-resident coder (host/batch.cpp).
struct DeviceCoder {
    DevBuf d_items, d_seqs, d_scratch, d_state, d_classes, d_calls, d_call_first;
    PinBuf h_stage, h_result;
    DevBuf d_dcalls, d_ditems, d_dseqs, d_dscratch, d_dstate, d_dout, d_waves;
    PinBuf h_dstage, h_dresult;
    // side streams for independent launches (default priority, non-blocking) and their events
    hipStream_t side[kSideStreams] = {};
    hipEvent_t fork = nullptr, join[kSideStreams] = {};
    uint64_t scratch_cap = 0;  // words, excluding the item-offset tail
    void release();  // frees buffers, streams and events (device must be current)
};

// Buffers of one brick-batch pipeline: coder context, brick work buffers and symbol streams.
struct BrickCtx {
    DeviceCoder coder;
    DevBuf d_bwork, d_bsym, d_blevels;
    PinBuf h_blevels;
    // stream and events of the context when it runs one group of a grouped batch
    hipStream_t st = nullptr;
    hipEvent_t ev_layout = nullptr, ev_done = nullptr;
    void ensure_stream();
    void release();
};

// Tuning of the kernel shapes (host/tuning.cpp).
cc_tuning default_tuning();
cc_tuning load_tuning_file(const std::string& path);  // defaults for keys the file lacks or a missing file
const cc_tuning& library_tuning();                    // the library's table, read once
void validate_tuning(const cc_tuning& t);

}  // namespace cca

struct cc_instance {
    int device = 0;
    uint32_t stream_count_max = 0;
    uint32_t elem_count_per_stream_max = 0;
    uint32_t cbs = 128;
    uint32_t log2 = 14;
    hipStream_t stream = nullptr;
    int dialect = CC_DIALECT_GPU;

    // encode scratch
    cca::DevBuf d_rle_sym, d_rle_zc;            // compacted streams, [stream][elem capacity]
    cca::DevBuf d_desc, d_tiles, d_map, d_map2, d_results;
    cca::DevBuf d_hist, d_tables, d_words, d_offsets, d_block_bits, d_stream_bits;
    cca::PinBuf h_up, h_down, h_hist, h_words;
    std::vector<uint32_t> out_words;
    std::vector<uint64_t> seg_begin, seg_bits;

    // decode scratch
    cca::DevBuf d_dec_stage, d_dec_sym, d_dec_zc, d_dec_sums, d_dec_desc;
    cca::PinBuf h_dec_stage[cca::kDecodeRing];
    hipEvent_t dec_event[cca::kDecodeRing] = {};
    int dec_slot = 0;

    // caller-level (volume codec) scratch
    cca::DevBuf d_vol[4];
    cca::DevBuf d_vol_ptrs;
    cca::PinBuf h_vol_ptrs;
    std::vector<uint32_t> vol_bits;

    // device-resident coder
    cca::DeviceCoder coder;

    // brick batches of the eager API (host/bricks.cpp): one context per group of a grouped batch
    cca::BrickCtx batch;
    std::vector<std::unique_ptr<cca::BrickCtx>> groups;
    hipEvent_t group_fork = nullptr;
    uint32_t brick_groups = 0;  // groups per brick batch, 0: chosen from the batch size
    cc_tuning tuning{};         // kernel shapes of brick batches and the device coder (cc_set_instance_tuning)

    // section timing (cc_set_timing_detail)
    int timing_detail = 0;
    cca::Timings timings;

    ~cc_instance();
};
