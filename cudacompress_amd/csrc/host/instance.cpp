// Device/host buffer helpers and instance teardown.
#include "instance.hpp"

#include "../profile.hpp"

namespace cca {

void* DevBuf::get(size_t bytes) {
    if (bytes <= cap && p) return p;
    release();
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        cap = 0;
        throw Error{CC_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(want) + " bytes failed"};
    }
    cap = want;
    return p;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

void* PinBuf::get(size_t bytes) {
    if (bytes <= cap && p) return p;
    if (bytes == 0 && p) return p;
    release();
    size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        cap = 0;
        throw Error{CC_ERR_OUT_OF_MEMORY, "hipHostMalloc of " + std::to_string(want) + " bytes failed"};
    }
    cap = want;
    return p;
}

void PinBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

DeviceGuard::DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) CCA_HIP(hipSetDevice(dev));
}

DeviceGuard::~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
}

void DeviceCoder::release() {
    DevBuf* devs[] = {&d_items, &d_seqs, &d_scratch, &d_state, &d_classes, &d_calls, &d_call_first, &d_dcalls,
                      &d_ditems, &d_dseqs, &d_dscratch, &d_dstate, &d_dout, &d_waves};
    for (auto* b : devs) b->release();
    h_stage.release();
    h_result.release();
    h_dstage.release();
    h_dresult.release();
    for (auto& e : join) {
        if (e) (void)hipEventDestroy(e);
        e = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
    for (auto& st : side) {
        if (st) (void)hipStreamDestroy(st);
        st = nullptr;
    }
    scratch_cap = 0;
}

void BrickCtx::ensure_stream() {
    if (!st) CCA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    if (!ev_layout) CCA_HIP(hipEventCreateWithFlags(&ev_layout, hipEventDisableTiming));
    if (!ev_done) CCA_HIP(hipEventCreateWithFlags(&ev_done, hipEventDisableTiming));
}

void BrickCtx::release() {
    if (st) (void)hipStreamSynchronize(st);
    coder.release();
    d_bwork.release();
    d_bsym.release();
    d_blevels.release();
    h_blevels.release();
    if (ev_layout) (void)hipEventDestroy(ev_layout);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (st) (void)hipStreamDestroy(st);
    st = nullptr;
    ev_layout = ev_done = nullptr;
}

}  // namespace cca

cc_instance::~cc_instance() {
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    else (void)hipDeviceSynchronize();
    cca::profile_owner_forget(this);  // its kernel timing vote and records
    for (auto& e : dec_event)
        if (e) (void)hipEventDestroy(e);
    cca::DevBuf* devs[] = {&d_rle_sym, &d_rle_zc, &d_desc, &d_tiles, &d_map, &d_map2, &d_results, &d_hist,
                           &d_tables, &d_words, &d_offsets, &d_block_bits, &d_stream_bits, &d_dec_stage,
                           &d_dec_sym, &d_dec_zc, &d_dec_sums, &d_dec_desc, &d_vol[0], &d_vol[1], &d_vol[2],
                           &d_vol[3], &d_vol_ptrs};
    for (auto* b : devs) b->release();
    cca::PinBuf* pins[] = {&h_up, &h_down, &h_hist, &h_words, &h_vol_ptrs};
    for (auto* b : pins) b->release();
    for (auto& b : h_dec_stage) b.release();
    coder.release();
    batch.release();
    for (auto& g : groups) g->release();
    if (group_fork) (void)hipEventDestroy(group_fork);
    if (prev >= 0) (void)hipSetDevice(prev);
}
