// extern "C" entry points declared in include/cudacompress_amd.h.  Every function converts
// cca::Error / std::bad_alloc into a status code and keeps the text for cc_last_error().
#include <algorithm>
#include <cstdio>
#include <new>
#include <string>

#include "../kernels.hpp"
#include "batch.hpp"
#include "bricks.hpp"
#include "server.hpp"
#include "codec.hpp"
#include "instance.hpp"
#include "volume.hpp"

namespace cca {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
void clear_error() { g_error.clear(); }
}  // namespace cca

namespace {

template <class F>
cc_status guarded(F&& f) {
    try {
        f();
        cca::clear_error();
        return CC_OK;
    } catch (const cca::Error& e) {
        cca::set_error(e.msg);
        return e.status;
    } catch (const std::bad_alloc&) {
        cca::set_error("host allocation failed");
        return CC_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        cca::set_error(e.what());
        return CC_ERR_DEVICE;
    }
}

// guarded() with the launches of the call tagged as the instance's (its kernel timing entries).
template <class F>
cc_status guarded_for(const void* owner, F&& f) {
    cca::ProfileOwner tag(owner);
    return guarded(std::forward<F>(f));
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

cc_status encode_entry(cc_instance* inst, bool rl, int symbolSize, const void* const* streams, uint32_t ns,
                       uint32_t n, const uint32_t* startBits, const uint32_t** outWords,
                       const uint64_t** outSegBegin, const uint64_t** outSegBits) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && outWords && outSegBegin && outSegBits, CC_ERR_INVALID_ARG, "null argument");
        cca::encode_streams(inst, rl, symbolSize, streams, ns, n, startBits);
        *outWords = inst->out_words.data();
        *outSegBegin = inst->seg_begin.data();
        *outSegBits = inst->seg_bits.data();
    });
}

cc_status decode_entry(cc_instance* inst, bool rl, int symbolSize, const uint32_t* const* words, const uint64_t* bits,
                       uint64_t* pos, uint32_t nbs, const uint32_t* s2b, void* const* outs, uint32_t ns, uint32_t n) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst, CC_ERR_INVALID_ARG, "null instance");
        cca::decode_streams(inst, rl, symbolSize, words, bits, pos, nbs, s2b, outs, ns, n);
    });
}

}  // namespace

extern "C" {

const char* cc_status_string(cc_status s) {
    switch (s) {
        case CC_OK: return "ok";
        case CC_ERR_SYMBOL_OVERFLOW: return "symbol exceeds the Huffman alphabet";
        case CC_ERR_OUT_OF_MEMORY: return "out of memory";
        case CC_ERR_INVALID_ARG: return "invalid argument";
        case CC_ERR_DEVICE: return "device error";
        case CC_ERR_CODEWORD_TOO_LONG: return "Huffman codeword longer than 31 bits";
        case CC_ERR_CORRUPT_STREAM: return "corrupt bitstream";
    }
    return "unknown status";
}

const char* cc_last_error(void) { return cca::g_error.c_str(); }

cc_status cc_create_instance(int device, uint32_t streamCountMax, uint32_t elemCountPerStreamMax,
                             uint32_t codingBlockSize, uint32_t log2HuffmanDistinctSymbolCountMax, cc_instance** out) {
    return guarded([&] {
        CCA_REQUIRE(out, CC_ERR_INVALID_ARG, "null output");
        *out = nullptr;
        if (codingBlockSize == 0) codingBlockSize = 128;
        if (log2HuffmanDistinctSymbolCountMax == 0) log2HuffmanDistinctSymbolCountMax = 14;
        CCA_REQUIRE(log2HuffmanDistinctSymbolCountMax <= 24, CC_ERR_INVALID_ARG,
                    "log2HuffmanDistinctSymbolCountMax must be <= 24");
        CCA_REQUIRE(codingBlockSize >= 32 && codingBlockSize <= 256 && (codingBlockSize & (codingBlockSize - 1)) == 0,
                    CC_ERR_INVALID_ARG, "codingBlockSize must be a power of two in [32, 256]");
        if (device < 0) CCA_HIP(hipGetDevice(&device));
        int count = 0;
        CCA_HIP(hipGetDeviceCount(&count));
        CCA_REQUIRE(device < count, CC_ERR_INVALID_ARG, "no such device");
        auto* I = new cc_instance();
        I->device = device;
        I->stream_count_max = streamCountMax;
        I->elem_count_per_stream_max = elemCountPerStreamMax;
        I->cbs = codingBlockSize;
        I->log2 = log2HuffmanDistinctSymbolCountMax;
        I->tuning = cca::library_tuning();
        try {
            cca::DeviceGuard g(device);
            for (auto& e : I->dec_event) {
                CCA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
                CCA_HIP(hipEventRecord(e, nullptr));
            }
            // reserve the per-stream buffers up front, as the reference sizes its arena
            // (InstanceImpl.cpp:85-96), so running out of memory shows up here
            const uint64_t symSize = I->log2 > 16 ? 4 : 2;
            const uint64_t per = (uint64_t)streamCountMax * cca::align_up(std::max<uint32_t>(elemCountPerStreamMax, 1), 64) * symSize;
            if (per > 0) {
                I->d_rle_sym.get(per);
                I->d_rle_zc.get(per);
                I->d_dec_sym.get(per);
                I->d_dec_zc.get(per);
            }
            CCA_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete I;
            throw;
        }
        *out = I;
    });
}

void cc_destroy_instance(cc_instance* inst) { delete inst; }

int cc_get_instance_device(const cc_instance* inst) { return inst ? inst->device : -1; }
uint32_t cc_get_instance_stream_count_max(const cc_instance* inst) { return inst ? inst->stream_count_max : 0; }
uint32_t cc_get_instance_elem_count_per_stream_max(const cc_instance* inst) {
    return inst ? inst->elem_count_per_stream_max : 0;
}
uint32_t cc_get_instance_coding_block_size(const cc_instance* inst) { return inst ? inst->cbs : 0; }
uint32_t cc_get_instance_log2_huffman_distinct_symbol_count_max(const cc_instance* inst) {
    return inst ? inst->log2 : 0;
}
int cc_get_instance_use_long_symbols(const cc_instance* inst) { return inst && inst->log2 > 16 ? 1 : 0; }
void cc_set_instance_stream(cc_instance* inst, void* stream) {
    if (inst) inst->stream = as_stream(stream);
}
void* cc_get_instance_stream(const cc_instance* inst) { return inst ? (void*)inst->stream : nullptr; }
cc_status cc_set_instance_dialect(cc_instance* inst, int dialect) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && (dialect == CC_DIALECT_GPU || dialect == CC_DIALECT_CPU), CC_ERR_INVALID_ARG,
                    "bad dialect");
        inst->dialect = dialect;
    });
}
cc_status cc_set_instance_brick_groups(cc_instance* inst, uint32_t groups) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && groups <= 64, CC_ERR_INVALID_ARG, "brick groups must be in [0, 64]");
        inst->brick_groups = groups;
    });
}

cc_status cc_encode_device(cc_instance* inst, int rl, int symbolSize, const void* const* d_symbolStreams,
                           uint32_t streamCount, uint32_t symbolCountPerStream, uint32_t* d_out,
                           uint64_t outCapacityWords, uint64_t* outWords) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && outWords && (d_out || outCapacityWords == 0), CC_ERR_INVALID_ARG, "null argument");
        CCA_REQUIRE(streamCount == 0 || d_symbolStreams, CC_ERR_INVALID_ARG, "null stream array");
        std::vector<const void*> streams(d_symbolStreams, d_symbolStreams + streamCount);
        std::vector<cca::CoderCall> calls{cca::CoderCall{rl != 0, 0, streamCount, symbolCountPerStream}};
        std::vector<cca::CallOut> res;
        *outWords = cca::device_encode(inst, calls, streams, symbolSize, d_out, outCapacityWords, &res, true);
        const uint32_t err = res.empty() ? 0u : res[0].err;
        CCA_REQUIRE(err != cca::kErrOverflow, CC_ERR_SYMBOL_OVERFLOW, "symbol exceeds 2^log2HuffmanDistinctSymbolCountMax");
        CCA_REQUIRE(err != cca::kErrTooLong, CC_ERR_CODEWORD_TOO_LONG, "Huffman codeword longer than 31 bits");
        CCA_REQUIRE(err != cca::kErrOutput, CC_ERR_INVALID_ARG, "output buffer too small");
        CCA_REQUIRE(err == 0, CC_ERR_OUT_OF_MEMORY, "device coder scratch exhausted");
    });
}

cc_status cc_decode_device(cc_instance* inst, int rl, int symbolSize, const uint32_t* d_words, uint64_t wordCount,
                           void* const* d_symbolStreams, uint32_t streamCount, uint32_t symbolCountPerStream,
                           uint64_t* outWordsConsumed) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && outWordsConsumed, CC_ERR_INVALID_ARG, "null argument");
        CCA_REQUIRE(streamCount == 0 || d_symbolStreams, CC_ERR_INVALID_ARG, "null stream array");
        std::vector<void*> outs(d_symbolStreams, d_symbolStreams + streamCount);
        std::vector<cca::DecodeCall> calls{
            cca::DecodeCall{d_words, wordCount, rl != 0, 0, streamCount, symbolCountPerStream}};
        std::vector<cca::DecCallOut> res;
        cca::device_decode(inst, calls, outs, symbolSize, &res, true);
        *outWordsConsumed = res.empty() ? 0 : res[0].words;
        const uint32_t err = res.empty() ? 0u : res[0].err;
        CCA_REQUIRE(err != cca::kDecCorrupt, CC_ERR_CORRUPT_STREAM, "malformed bitstream");
        CCA_REQUIRE(err == 0, CC_ERR_OUT_OF_MEMORY, "device decoder scratch exhausted");
    });
}

cc_status cc_decode_device_noalloc(cc_instance* inst, int rl, int symbolSize, const uint32_t* d_words,
                                   uint64_t wordCount, void* const* d_symbolStreams, uint32_t streamCount,
                                   uint32_t symbolCountPerStream, uint64_t* outWordsConsumed) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && outWordsConsumed, CC_ERR_INVALID_ARG, "null argument");
        CCA_REQUIRE(streamCount == 0 || d_symbolStreams, CC_ERR_INVALID_ARG, "null stream array");
        uint32_t err = 0;
        *outWordsConsumed = cca::decode_device_noalloc(inst, rl != 0, symbolSize, d_words, wordCount, d_symbolStreams,
                                                       streamCount, symbolCountPerStream, &err);
        CCA_REQUIRE(err != cca::kDecCorrupt, CC_ERR_CORRUPT_STREAM, "malformed bitstream");
        CCA_REQUIRE(err == 0, CC_ERR_OUT_OF_MEMORY, "no-allocation decode scratch exhausted");
    });
}

cc_status cc_compress_bricks_float(cc_instance* inst, const cc_brick* bricks, uint32_t count, uint32_t numLevels,
                                   int symbolSize, uint32_t* d_out, uint64_t outCapacityWords,
                                   uint64_t* brickWordOffsets, uint32_t* brickStatus) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && brickWordOffsets && (d_out || outCapacityWords == 0), CC_ERR_INVALID_ARG, "null argument");
        cca::compress_bricks(inst, bricks, count, numLevels, symbolSize, d_out, outCapacityWords, brickWordOffsets,
                             brickStatus);
    });
}

cc_status cc_decompress_bricks_float(cc_instance* inst, const cc_brick* bricks, uint32_t count, uint32_t numLevels,
                                     int symbolSize, const uint32_t* d_in, const uint64_t* brickWordOffsets,
                                     uint32_t* brickStatus) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && brickWordOffsets && (d_in || count == 0), CC_ERR_INVALID_ARG, "null argument");
        cca::decompress_bricks(inst, bricks, count, numLevels, symbolSize, d_in, brickWordOffsets, brickStatus);
    });
}

cc_status cc_bricks_plan_create(cc_instance* inst, const cc_brick* bricks, uint32_t count, uint32_t numLevels,
                                int symbolSize, uint32_t* d_out, uint64_t outCapacityWords, cc_bricks_plan** plan) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && plan && d_out, CC_ERR_INVALID_ARG, "null argument");
        *plan = nullptr;
        *plan = cca::plan_create(inst, bricks, count, numLevels, symbolSize, d_out, outCapacityWords);
    });
}

cc_status cc_bricks_plan_compress(cc_bricks_plan* plan, uint64_t* brickWordOffsets, uint32_t* brickStatus) {
    return guarded_for(cca::plan_owner(plan), [&] {
        CCA_REQUIRE(plan && brickWordOffsets, CC_ERR_INVALID_ARG, "null argument");
        cca::plan_compress(plan, brickWordOffsets, brickStatus);
    });
}

cc_status cc_bricks_plan_decompress(cc_bricks_plan* plan, uint32_t* brickStatus) {
    return guarded_for(cca::plan_owner(plan), [&] {
        CCA_REQUIRE(plan, CC_ERR_INVALID_ARG, "null plan");
        cca::plan_decompress(plan, brickStatus);
    });
}

cc_status cc_bricks_plan_run(cc_bricks_plan* plan, uint64_t* brickWordOffsets, uint32_t* compressStatus,
                             uint32_t* decompressStatus) {
    return guarded_for(cca::plan_owner(plan), [&] {
        CCA_REQUIRE(plan && brickWordOffsets, CC_ERR_INVALID_ARG, "null argument");
        cca::plan_run(plan, brickWordOffsets, compressStatus, decompressStatus);
    });
}

cc_status cc_bricks_plan_launch(cc_bricks_plan* plan) {
    return guarded_for(cca::plan_owner(plan), [&] {
        CCA_REQUIRE(plan, CC_ERR_INVALID_ARG, "null argument");
        cca::plan_launch(plan);
    });
}

cc_status cc_bricks_plan_wait(cc_bricks_plan* plan, uint64_t* brickWordOffsets, uint32_t* compressStatus,
                              uint32_t* decompressStatus) {
    return guarded_for(cca::plan_owner(plan), [&] {
        CCA_REQUIRE(plan && brickWordOffsets, CC_ERR_INVALID_ARG, "null argument");
        cca::plan_wait(plan, brickWordOffsets, compressStatus, decompressStatus);
    });
}

void cc_bricks_plan_destroy(cc_bricks_plan* plan) { cca::plan_destroy(plan); }

cc_status cc_bricks_plan_coding_stats(cc_bricks_plan* plan, uint64_t* stats) {
    return guarded_for(cca::plan_owner(plan), [&] {
        CCA_REQUIRE(plan && stats, CC_ERR_INVALID_ARG, "null argument");
        cca::plan_coding_stats(plan, stats);
    });
}

cc_status cc_server_create(cc_instance* inst, const uint32_t* d_words, uint32_t brickCount,
                           const uint64_t* brickWordRanges, const uint32_t* brickSizes, const float* quantSteps,
                           uint32_t numLevels, int symbolSize, uint32_t slots, uint64_t workspaceBytes,
                           cc_server** out) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && out, CC_ERR_INVALID_ARG, "null argument");
        *out = nullptr;
        *out = cca::server_create(inst, d_words, brickCount, brickWordRanges, brickSizes, quantSteps, numLevels,
                                  symbolSize, slots, workspaceBytes);
    });
}

cc_status cc_server_submit(cc_server* server, uint64_t requestId, const uint32_t* brickIds, uint32_t count,
                           float* const* d_outputs, double arrivalTime) {
    return guarded_for(cca::server_owner(server), [&] {
        CCA_REQUIRE(server, CC_ERR_INVALID_ARG, "null server");
        cca::server_submit(server, requestId, brickIds, count, d_outputs, arrivalTime);
    });
}

cc_status cc_server_poll(cc_server* server, double now, uint32_t* batchesLaunched) {
    return guarded_for(cca::server_owner(server), [&] {
        CCA_REQUIRE(server, CC_ERR_INVALID_ARG, "null server");
        const uint32_t n = cca::server_poll(server, now);
        if (batchesLaunched) *batchesLaunched = n;
    });
}

uint32_t cc_server_pending(const cc_server* server) { return server ? cca::server_busy(server) : 0u; }

uint32_t cc_server_completed(cc_server* server, uint64_t* requestIds, double* times, uint32_t max) {
    return server ? cca::server_take_completed(server, requestIds, times, max) : 0u;
}

uint32_t cc_server_failed(cc_server* server, uint64_t* requestIds, uint32_t* status, uint32_t max) {
    return server ? cca::server_take_failed(server, requestIds, status, max) : 0u;
}

cc_status cc_server_run(cc_server* server, uint32_t requestCount, const double* arrivals, const uint32_t* counts,
                        const uint32_t* brickIds, float* const* d_outputs, double* latencies, double* elapsed) {
    return guarded_for(cca::server_owner(server), [&] {
        CCA_REQUIRE(server, CC_ERR_INVALID_ARG, "null server");
        cca::server_run(server, requestCount, arrivals, counts, brickIds, d_outputs, latencies, elapsed);
    });
}

void cc_server_destroy(cc_server* server) { cca::server_destroy(server); }

cc_status cc_encode_rlhuff(cc_instance* inst, int symbolSize, const void* const* d_symbolStreams, uint32_t streamCount,
                           uint32_t symbolCountPerStream, const uint32_t* startBitOffsets, const uint32_t** outWords,
                           const uint64_t** outSegmentWordBegin, const uint64_t** outSegmentBits) {
    return encode_entry(inst, true, symbolSize, d_symbolStreams, streamCount, symbolCountPerStream, startBitOffsets,
                        outWords, outSegmentWordBegin, outSegmentBits);
}
cc_status cc_encode_huff(cc_instance* inst, int symbolSize, const void* const* d_symbolStreams, uint32_t streamCount,
                         uint32_t symbolCountPerStream, const uint32_t* startBitOffsets, const uint32_t** outWords,
                         const uint64_t** outSegmentWordBegin, const uint64_t** outSegmentBits) {
    return encode_entry(inst, false, symbolSize, d_symbolStreams, streamCount, symbolCountPerStream, startBitOffsets,
                        outWords, outSegmentWordBegin, outSegmentBits);
}
cc_status cc_decode_rlhuff(cc_instance* inst, int symbolSize, const uint32_t* const* bitstreamWords,
                           const uint64_t* bitstreamBits, uint64_t* bitstreamPos, uint32_t bitstreamCount,
                           const uint32_t* streamToBitstream, void* const* d_symbolStreams, uint32_t streamCount,
                           uint32_t symbolCountPerStream) {
    return decode_entry(inst, true, symbolSize, bitstreamWords, bitstreamBits, bitstreamPos, bitstreamCount,
                        streamToBitstream, d_symbolStreams, streamCount, symbolCountPerStream);
}
cc_status cc_decode_huff(cc_instance* inst, int symbolSize, const uint32_t* const* bitstreamWords,
                         const uint64_t* bitstreamBits, uint64_t* bitstreamPos, uint32_t bitstreamCount,
                         const uint32_t* streamToBitstream, void* const* d_symbolStreams, uint32_t streamCount,
                         uint32_t symbolCountPerStream) {
    return decode_entry(inst, false, symbolSize, bitstreamWords, bitstreamBits, bitstreamPos, bitstreamCount,
                        streamToBitstream, d_symbolStreams, streamCount, symbolCountPerStream);
}

cc_status cc_dwt_float_forward(float* d_dst, float* d_buffer2, float* d_buffer1, const float* d_src, int sizeX,
                               int sizeY, int sizeZ, int srcChannelCount, int dstRowPitch, int dstSlicePitch,
                               int srcRowPitch, int srcSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt97_forward(d_dst, d_buffer2, d_buffer1, d_src, sizeX, sizeY, sizeZ, srcChannelCount, dstRowPitch,
                           dstSlicePitch, srcRowPitch, srcSlicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_double_forward(double* d_dst, double* d_buffer, const double* d_src, int sizeX, int sizeY, int sizeZ,
                                int rowPitch, int slicePitch, void* stream) {
    return guarded([&] {
        CCA_REQUIRE(d_dst != d_src, CC_ERR_INVALID_ARG, "d_dst and d_src must not overlap");
        cca::dwt97_forward_f64(d_dst, d_buffer, d_src, sizeX, sizeY, sizeZ, rowPitch, slicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_float_inverse(float* d_dst, float* d_buffer2, float* d_buffer1, const float* d_src, int sizeX,
                               int sizeY, int sizeZ, int dstChannelCount, int dstRowPitch, int dstSlicePitch,
                               int srcRowPitch, int srcSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt97_inverse(d_dst, d_buffer2, d_buffer1, d_src, sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch,
                           dstSlicePitch, srcRowPitch, srcSlicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_float_inverse_from_symbols(float* d_dst, float* d_buffer2, float* d_buffer1, const float* d_lowpass,
                                            const void* const* dd_highpass, int symbolSize, float quantStep,
                                            int sizeX, int sizeY, int sizeZ, int dstChannelCount, int dstRowPitch,
                                            int dstSlicePitch, int lowpassRowPitch, int lowpassSlicePitch,
                                            void* stream) {
    return guarded([&] {
        cca::dwt97_inverse_from_symbols(d_dst, d_buffer2, d_buffer1, d_lowpass, dd_highpass, symbolSize, quantStep,
                                        sizeX, sizeY, sizeZ, dstChannelCount, dstRowPitch, dstSlicePitch,
                                        lowpassRowPitch, lowpassSlicePitch, as_stream(stream));
    });
}
cc_status cc_default_tuning(cc_tuning* out) {
    return guarded([&] {
        CCA_REQUIRE(out, CC_ERR_INVALID_ARG, "null argument");
        *out = cca::default_tuning();
    });
}
cc_status cc_load_tuning(const char* path, cc_tuning* out) {
    return guarded([&] {
        CCA_REQUIRE(out, CC_ERR_INVALID_ARG, "null argument");
        *out = path ? cca::load_tuning_file(path) : cca::library_tuning();
    });
}
cc_status cc_set_instance_tuning(cc_instance* inst, const cc_tuning* tuning) {
    return guarded([&] {
        CCA_REQUIRE(inst, CC_ERR_INVALID_ARG, "null instance");
        if (tuning) cca::validate_tuning(*tuning);
        inst->tuning = tuning ? *tuning : cca::library_tuning();
    });
}
cc_status cc_get_instance_tuning(const cc_instance* inst, cc_tuning* out) {
    return guarded([&] {
        CCA_REQUIRE(inst && out, CC_ERR_INVALID_ARG, "null argument");
        *out = inst->tuning;
    });
}
cc_status cc_set_timing_detail(cc_instance* inst, int detail) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && detail >= 0 && detail <= 3, CC_ERR_INVALID_ARG, "bad timing detail");
        inst->timing_detail = detail;
        // level 3: this instance's kernel launches are timed into its own records, whatever other
        // instances or the process-wide switch do
        cca::profile_vote(inst, detail >= 3);
    });
}
cc_status cc_get_timings(cc_instance* inst, uint32_t* count) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && count, CC_ERR_INVALID_ARG, "null argument");
        cca::DeviceGuard guard(inst->device);
        cca::Timings& t = inst->timings;
        t.collect();
        t.view.clear();
        for (const std::string& n : t.order) t.view.push_back({n, (float)t.ms[n]});
        if (inst->timing_detail >= 3)
            for (auto& e : cca::profile_owner_entries(inst)) t.view.push_back({"Kernel: " + e.first, (float)e.second});
        *count = (uint32_t)t.view.size();
    });
}
cc_status cc_get_timing_entry(cc_instance* inst, uint32_t i, const char** name, float* ms) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && name && ms, CC_ERR_INVALID_ARG, "null argument");
        CCA_REQUIRE(i < inst->timings.view.size(), CC_ERR_INVALID_ARG, "timing index out of range");
        *name = inst->timings.view[i].first.c_str();
        *ms = inst->timings.view[i].second;
    });
}
cc_status cc_print_timings(cc_instance* inst) {
    uint32_t n = 0;
    cc_status rc = cc_get_timings(inst, &n);
    if (rc != CC_OK) return rc;
    size_t w = 0;
    for (auto& e : inst->timings.view) w = std::max(w, e.first.size());
    for (auto& e : inst->timings.view) std::printf("%-*s: %7.2f ms\n", (int)w + 1, e.first.c_str(), e.second);
    return CC_OK;
}
cc_status cc_reset_timings(cc_instance* inst) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst, CC_ERR_INVALID_ARG, "null argument");
        cca::DeviceGuard guard(inst->device);
        inst->timings.reset();
        inst->timings.view.clear();
        cca::profile_owner_reset(inst);
    });
}
cc_status cc_dwt_float_forward_from_byte(float* d_dst, float* d_buffer2, float* d_buffer1, const uint8_t* d_src,
                                         int sizeX, int sizeY, int sizeZ, int srcChannelCount, int dstRowPitch,
                                         int dstSlicePitch, int srcRowPitch, int srcSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt97_forward_from_byte(d_dst, d_buffer2, d_buffer1, d_src, sizeX, sizeY, sizeZ, srcChannelCount,
                                     dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_float_inverse_to_byte(uint8_t* d_dst, float* d_buffer2, float* d_buffer1, const float* d_src,
                                       int sizeX, int sizeY, int sizeZ, int dstChannelCount, int dstRowPitch,
                                       int dstSlicePitch, int srcRowPitch, int srcSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt97_inverse_to_byte(d_dst, d_buffer2, d_buffer1, d_src, sizeX, sizeY, sizeZ, dstChannelCount,
                                   dstRowPitch, dstSlicePitch, srcRowPitch, srcSlicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_float_inverse_from_symbols_to_byte(uint8_t* d_dst, float* d_buffer2, float* d_buffer1,
                                                    const float* d_lowpass, const uint16_t* const* dd_highpass,
                                                    float quantStep, int sizeX, int sizeY, int sizeZ,
                                                    int dstChannelCount, int dstRowPitch, int dstSlicePitch,
                                                    int lowpassRowPitch, int lowpassSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt97_inverse_from_symbols_to_byte(d_dst, d_buffer2, d_buffer1, d_lowpass,
                                                (const void* const*)dd_highpass, quantStep, sizeX, sizeY, sizeZ,
                                                dstChannelCount, dstRowPitch, dstSlicePitch, lowpassRowPitch,
                                                lowpassSlicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_float_forward_lowpass_only_from_byte(float* d_dst, float* d_buffer, const uint8_t* d_src, int sizeX,
                                                      int sizeY, int srcChannelCount, int dstRowPitch, int srcRowPitch,
                                                      void* stream) {
    return guarded([&] {
        cca::dwt97_forward_lowpass_only_from_byte(d_dst, d_buffer, d_src, sizeX, sizeY, srcChannelCount, dstRowPitch,
                                                  srcRowPitch, as_stream(stream));
    });
}
cc_status cc_dwt_int_forward_lowpass_only(int16_t* d_dst, int16_t* d_buffer, const int16_t* d_src, int sizeX,
                                          int sizeY, int dstRowPitch, int srcRowPitch, void* stream) {
    return guarded([&] {
        cca::dwt53_forward_lowpass_only(d_dst, d_buffer, d_src, sizeX, sizeY, dstRowPitch, srcRowPitch,
                                        as_stream(stream));
    });
}
cc_status cc_float_to_byte_2d(uint8_t* d_quant, uint32_t channelCount, uint32_t channel, const float* d_data,
                              uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchSrc, void* stream) {
    return guarded([&] {
        cca::float_to_byte_2d(d_quant, channelCount, channel, d_data, sizeX, sizeY, rowPitchSrc, as_stream(stream));
    });
}
cc_status cc_byte_to_float_2d(float* d_data, const uint8_t* d_quant, uint32_t channelCount, uint32_t channel,
                              uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchDst, void* stream) {
    return guarded([&] {
        cca::byte_to_float_2d(d_data, d_quant, channelCount, channel, sizeX, sizeY, rowPitchDst, as_stream(stream));
    });
}
cc_status cc_quantize_difference_to_symbols_2d(uint16_t* d_symbols, const float* d_data, float quantizationStep,
                                               const uint8_t* d_reference, uint32_t channelCount, uint32_t channel,
                                               uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchSrc, int quantType,
                                               void* stream) {
    return guarded([&] {
        cca::quantize_difference_2d(d_symbols, const_cast<float*>(d_data), quantizationStep, d_reference, channelCount,
                                    channel, sizeX, sizeY, rowPitchSrc, quantType, false, as_stream(stream));
    });
}
cc_status cc_quantize_difference_to_symbols_roundtrip_2d(uint16_t* d_symbols, float* d_data, float quantizationStep,
                                                         const uint8_t* d_reference, uint32_t channelCount,
                                                         uint32_t channel, uint32_t sizeX, uint32_t sizeY,
                                                         uint32_t rowPitchSrc, int quantType, void* stream) {
    return guarded([&] {
        cca::quantize_difference_2d(d_symbols, d_data, quantizationStep, d_reference, channelCount, channel, sizeX,
                                    sizeY, rowPitchSrc, quantType, true, as_stream(stream));
    });
}
cc_status cc_unquantize_difference_from_symbols_2d(float* d_data, const uint16_t* d_symbols, float quantizationStep,
                                                   const uint8_t* d_reference, uint32_t channelCount, uint32_t channel,
                                                   uint32_t sizeX, uint32_t sizeY, uint32_t rowPitchDst,
                                                   int quantType, void* stream) {
    return guarded([&] {
        cca::unquantize_difference_2d(d_data, d_symbols, quantizationStep, d_reference, channelCount, channel, sizeX,
                                      sizeY, rowPitchDst, quantType, as_stream(stream));
    });
}
cc_status cc_convert_rgb_to_ycocg(uint8_t* d_target, const uint8_t* d_data, int pixelCount, int channels,
                                  void* stream) {
    return guarded([&] {
        CCA_REQUIRE(pixelCount >= 0, CC_ERR_INVALID_ARG, "negative pixel count");
        cca::convert_ycocg(d_target, d_data, (uint32_t)pixelCount, channels, true, as_stream(stream));
    });
}
cc_status cc_convert_ycocg_to_rgb(uint8_t* d_target, const uint8_t* d_data, int pixelCount, int channels,
                                  void* stream) {
    return guarded([&] {
        CCA_REQUIRE(pixelCount >= 0, CC_ERR_INVALID_ARG, "negative pixel count");
        cca::convert_ycocg(d_target, d_data, (uint32_t)pixelCount, channels, false, as_stream(stream));
    });
}
cc_status cc_get_lsb(const int16_t* d_data, uint32_t* d_bits, uint32_t elemCount, void* stream) {
    return guarded([&] { cca::lsb_get(d_data, d_bits, elemCount, as_stream(stream)); });
}
cc_status cc_remove_lsb(int16_t* d_data, uint32_t* d_bits, uint32_t elemCount, int16_t shift, void* stream) {
    return guarded([&] { cca::lsb_remove(d_data, d_bits, elemCount, shift, as_stream(stream)); });
}
cc_status cc_append_lsb(int16_t* d_data, const uint32_t* d_bits, uint32_t elemCount, int16_t shift, void* stream) {
    return guarded([&] { cca::lsb_append(d_data, d_bits, elemCount, shift, as_stream(stream)); });
}
cc_status cc_shift_elems(int16_t* d_data, uint32_t elemCount, int16_t shift, void* stream) {
    return guarded([&] { cca::shift_elems(d_data, elemCount, shift, as_stream(stream)); });
}
cc_status cc_dwt_int_forward(int16_t* d_dst, int16_t* d_buffer, const int16_t* d_src, int sizeX, int sizeY, int sizeZ,
                             int dstRowPitch, int dstSlicePitch, int srcRowPitch, int srcSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt53_forward(d_dst, d_buffer, d_src, sizeX, sizeY, sizeZ, dstRowPitch, dstSlicePitch, srcRowPitch,
                           srcSlicePitch, as_stream(stream));
    });
}
cc_status cc_dwt_int_inverse(int16_t* d_dst, int16_t* d_buffer, const int16_t* d_src, int sizeX, int sizeY, int sizeZ,
                             int dstRowPitch, int dstSlicePitch, int srcRowPitch, int srcSlicePitch, void* stream) {
    return guarded([&] {
        cca::dwt53_inverse(d_dst, d_buffer, d_src, sizeX, sizeY, sizeZ, dstRowPitch, dstSlicePitch, srcRowPitch,
                           srcSlicePitch, as_stream(stream));
    });
}

cc_status cc_symbolize(uint16_t* d_symbols, const int16_t* d_data, uint32_t sizeX, uint32_t sizeY, uint32_t sizeZ,
                       uint32_t rowPitchSrc, uint32_t slicePitchSrc, void* stream) {
    return guarded([&] { cca::symbolize(d_symbols, d_data, sizeX, sizeY, sizeZ, rowPitchSrc, slicePitchSrc, as_stream(stream)); });
}
cc_status cc_unsymbolize(int16_t* d_data, const uint16_t* d_symbols, uint32_t sizeX, uint32_t sizeY, uint32_t sizeZ,
                         uint32_t rowPitchDst, uint32_t slicePitchDst, void* stream) {
    return guarded([&] { cca::unsymbolize(d_data, d_symbols, sizeX, sizeY, sizeZ, rowPitchDst, slicePitchDst, as_stream(stream)); });
}
cc_status cc_quantize_to_symbols(void* d_symbols, int symbolSize, const float* d_data, uint32_t sizeX, uint32_t sizeY,
                                 uint32_t sizeZ, float quantizationStep, uint32_t rowPitchSrc, uint32_t slicePitchSrc,
                                 int quantType, void* stream) {
    return guarded([&] {
        cca::quantize_to_symbols(d_symbols, symbolSize, const_cast<float*>(d_data), false, sizeX, sizeY, sizeZ,
                                 quantizationStep, rowPitchSrc, slicePitchSrc, quantType, as_stream(stream));
    });
}
cc_status cc_quantize_to_symbols_roundtrip(void* d_symbols, int symbolSize, float* d_data, uint32_t sizeX,
                                           uint32_t sizeY, uint32_t sizeZ, float quantizationStep,
                                           uint32_t rowPitchSrc, uint32_t slicePitchSrc, int quantType, void* stream) {
    return guarded([&] {
        cca::quantize_to_symbols(d_symbols, symbolSize, d_data, true, sizeX, sizeY, sizeZ, quantizationStep,
                                 rowPitchSrc, slicePitchSrc, quantType, as_stream(stream));
    });
}
cc_status cc_unquantize_from_symbols(float* d_data, const void* d_symbols, int symbolSize, uint32_t sizeX,
                                     uint32_t sizeY, uint32_t sizeZ, float quantizationStep, uint32_t rowPitchDst,
                                     uint32_t slicePitchDst, int quantType, void* stream) {
    return guarded([&] {
        cca::unquantize_from_symbols(d_data, d_symbols, symbolSize, sizeX, sizeY, sizeZ, quantizationStep,
                                     rowPitchDst, slicePitchDst, quantType, as_stream(stream));
    });
}
cc_status cc_quantize_to_short(int16_t* d_quant, const float* d_data, uint32_t sizeX, uint32_t sizeY, uint32_t sizeZ,
                               float quantizationStep, uint32_t rowPitchSrc, uint32_t slicePitchSrc, int quantType,
                               void* stream) {
    return guarded([&] {
        cca::quantize_to_short(d_quant, d_data, sizeX, sizeY, sizeZ, quantizationStep, rowPitchSrc, slicePitchSrc,
                               quantType, as_stream(stream));
    });
}
cc_status cc_unquantize_from_short(float* d_data, const int16_t* d_quant, uint32_t sizeX, uint32_t sizeY,
                                   uint32_t sizeZ, float quantizationStep, uint32_t rowPitchDst, uint32_t slicePitchDst,
                                   int quantType, void* stream) {
    return guarded([&] {
        cca::unquantize_from_short(d_data, d_quant, sizeX, sizeY, sizeZ, quantizationStep, rowPitchDst, slicePitchDst,
                                   quantType, as_stream(stream));
    });
}
cc_status cc_get_max_abs(cc_instance* inst, const float* d_image, uint32_t elemCount, float* d_valMax) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && d_valMax, CC_ERR_INVALID_ARG, "null argument");
        cca::DeviceGuard g(inst->device);
        cca::max_abs(d_image, elemCount, d_valMax, inst->stream);
    });
}

cc_status cc_compress_volume_float(cc_instance* inst, const float* d_image, uint32_t sizeX, uint32_t sizeY,
                                   uint32_t sizeZ, uint32_t numLevels, float quantStepLevel0, int doRLEOnlyOnLvl0,
                                   int symbolSize, const uint32_t** outWords, uint64_t* outBits) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && outWords && outBits, CC_ERR_INVALID_ARG, "null argument");
        cca::compress_volume_float(inst, d_image, sizeX, sizeY, sizeZ, numLevels, quantStepLevel0,
                                   doRLEOnlyOnLvl0 != 0, symbolSize);
        *outWords = inst->vol_bits.data();
        *outBits = 32ull * inst->vol_bits.size();
    });
}
cc_status cc_decompress_volume_float(cc_instance* inst, float* d_image, uint32_t sizeX, uint32_t sizeY, uint32_t sizeZ,
                                     uint32_t numLevels, const uint32_t* bits, uint64_t bitCount,
                                     float quantStepLevel0, int doRLEOnlyOnLvl0, int symbolSize) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && bits, CC_ERR_INVALID_ARG, "null argument");
        cca::decompress_volume_float(inst, d_image, sizeX, sizeY, sizeZ, numLevels, bits, bitCount, quantStepLevel0,
                                     doRLEOnlyOnLvl0 != 0, symbolSize);
    });
}
cc_status cc_compress_volume_float_quant_first(cc_instance* inst, const float* d_image, uint32_t sizeX, uint32_t sizeY,
                                               uint32_t sizeZ, uint32_t numLevels, float quantizationStep,
                                               int doRLEOnlyOnLvl0, const uint32_t** outWords, uint64_t* outBits) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && d_image && outWords && outBits, CC_ERR_INVALID_ARG, "null argument");
        cca::compress_volume_float_quant_first(inst, d_image, sizeX, sizeY, sizeZ, numLevels, quantizationStep,
                                               doRLEOnlyOnLvl0 != 0);
        *outWords = inst->vol_bits.data();
        *outBits = 32ull * inst->vol_bits.size();
    });
}
cc_status cc_decompress_volume_float_quant_first(cc_instance* inst, float* d_image, uint32_t sizeX, uint32_t sizeY,
                                                 uint32_t sizeZ, uint32_t numLevels, const uint32_t* bits,
                                                 uint64_t bitCount, float quantizationStep, int doRLEOnlyOnLvl0) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && d_image && bits, CC_ERR_INVALID_ARG, "null argument");
        cca::decompress_volume_float_quant_first(inst, d_image, sizeX, sizeY, sizeZ, numLevels, bits, bitCount,
                                                 quantizationStep, doRLEOnlyOnLvl0 != 0);
    });
}
namespace {
std::vector<cca::VolumeChannelIn> channels_in(const cc_volume_channel* ch, uint32_t n) {
    std::vector<cca::VolumeChannelIn> v(n);
    for (uint32_t c = 0; c < n; c++) v[c] = {ch[c].d_image, ch[c].bits, ch[c].bitCount, ch[c].quantStep};
    return v;
}
}  // namespace
cc_status cc_decompress_volume_float_multichannel(cc_instance* inst, const cc_volume_channel* channels,
                                                  uint32_t channelCount, uint32_t sizeX, uint32_t sizeY, uint32_t sizeZ,
                                                  uint32_t numLevels, int doRLEOnlyOnLvl0, int symbolSize) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && channels && channelCount, CC_ERR_INVALID_ARG, "null argument");
        auto v = channels_in(channels, channelCount);
        cca::decompress_volume_float_multichannel(inst, v.data(), channelCount, sizeX, sizeY, sizeZ, numLevels,
                                                  doRLEOnlyOnLvl0 != 0, symbolSize);
    });
}
cc_status cc_decompress_volume_float_quant_first_multichannel(cc_instance* inst, const cc_volume_channel* channels,
                                                              uint32_t channelCount, uint32_t sizeX, uint32_t sizeY,
                                                              uint32_t sizeZ, uint32_t numLevels,
                                                              int doRLEOnlyOnLvl0) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && channels && channelCount, CC_ERR_INVALID_ARG, "null argument");
        auto v = channels_in(channels, channelCount);
        cca::decompress_volume_float_quant_first_multichannel(inst, v.data(), channelCount, sizeX, sizeY, sizeZ,
                                                              numLevels, doRLEOnlyOnLvl0 != 0);
    });
}
cc_status cc_compress_volume_lossless(cc_instance* inst, const int16_t* d_image, uint32_t sizeX, uint32_t sizeY,
                                      uint32_t sizeZ, uint32_t numLevels, const uint32_t** outWords,
                                      uint64_t* outBits) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && outWords && outBits, CC_ERR_INVALID_ARG, "null argument");
        cca::compress_volume_lossless(inst, d_image, sizeX, sizeY, sizeZ, numLevels);
        *outWords = inst->vol_bits.data();
        *outBits = 32ull * inst->vol_bits.size();
    });
}
cc_status cc_decompress_volume_lossless(cc_instance* inst, int16_t* d_image, uint32_t sizeX, uint32_t sizeY,
                                        uint32_t sizeZ, uint32_t numLevels, const uint32_t* bits, uint64_t bitCount) {
    return guarded_for(inst, [&] {
        CCA_REQUIRE(inst && bits, CC_ERR_INVALID_ARG, "null argument");
        cca::decompress_volume_lossless(inst, d_image, sizeX, sizeY, sizeZ, numLevels, bits, bitCount);
    });
}

}  // extern "C"
