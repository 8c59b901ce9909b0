// Host side of the device-resident coder: one pinned upload of the item list, the kernel
// pipeline of hip/coder.hip, one download of the per-call results.
#include "batch.hpp"

#include <algorithm>
#include <cstring>

#include "huffman_table.hpp"

namespace cca {

namespace {

// Arena estimate: histograms, tables and symbol lists of the usual (LDS-class) size, codewords
// at up to one word per symbol, block offsets.  Too small is detected on the device and retried.
uint64_t arena_estimate(const std::vector<CoderCall>& calls, uint32_t log2, uint32_t cbs) {
    const uint64_t alpha = std::min<uint64_t>(1ull << log2, 8192);
    uint64_t words = 4096;
    for (const CoderCall& c : calls) {
        const uint64_t nseq = c.rl ? 2 : 1;
        const uint64_t per = (uint64_t)c.n + c.n / cbs + 4 * std::min<uint64_t>(alpha, 2ull * c.n + 64) + 128;
        words += (uint64_t)c.nstreams * nseq * per;
    }
    return words;
}

}  // namespace

uint64_t device_encode(cc_instance* I, const std::vector<CoderCall>& calls, const std::vector<const void*>& streams,
                       int symSize, uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    DeviceCoder& D = I->coder;
    const bool longSymbols = I->log2 > 16;
    CCA_REQUIRE(symSize == (longSymbols ? 4 : 2), CC_ERR_INVALID_ARG,
                "symbol size does not match the instance (log2HuffmanDistinctSymbolCountMax > 16 uses Symbol32)");
    CCA_REQUIRE(I->dialect == CC_DIALECT_GPU || symSize == 2, CC_ERR_INVALID_ARG, "the CPU dialect is Symbol16 only");
    const uint32_t ncalls = (uint32_t)calls.size();
    uint32_t nitems = 0, nseq = 0;
    uint64_t total_syms = 0;
    for (const CoderCall& c : calls) {
        CCA_REQUIRE((uint64_t)c.first_stream + c.nstreams <= streams.size(), CC_ERR_INVALID_ARG, "bad stream range");
        nitems += c.nstreams;
        nseq += c.nstreams * (c.rl ? 2 : 1);
        total_syms += (uint64_t)c.nstreams * c.n;
    }
    // stage: items, call_first
    const size_t o_items = 0;
    const size_t o_first = align_up(sizeof(CodeItem) * std::max<uint32_t>(nitems, 1), 256);
    const size_t stage_bytes = o_first + sizeof(uint32_t) * (ncalls + 1);
    unsigned char* h = (unsigned char*)D.h_stage.get(stage_bytes);
    CodeItem* hi = (CodeItem*)(h + o_items);
    uint32_t* hf = (uint32_t*)(h + o_first);
    uint32_t item = 0, seq = 0;
    for (uint32_t c = 0; c < ncalls; c++) {
        hf[c] = item;
        for (uint32_t s = 0; s < calls[c].nstreams; s++) {
            const void* p = streams[calls[c].first_stream + s];
            CCA_REQUIRE(p != nullptr || calls[c].n == 0, CC_ERR_INVALID_ARG, "null stream");
            hi[item] = CodeItem{p, calls[c].n, calls[c].rl ? 1u : 0u, seq, c};
            item++;
            seq += calls[c].rl ? 2 : 1;
        }
    }
    hf[ncalls] = item;
    CodeItem* d_items = D.d_items.as<CodeItem>(std::max<uint32_t>(nitems, 1));
    uint32_t* d_first = D.d_call_first.as<uint32_t>(ncalls + 1);
    SeqInfo* d_seqs = D.d_seqs.as<SeqInfo>(std::max<uint32_t>(nseq, 1));
    CoderState* d_state = D.d_state.as<CoderState>(1);
    CallOut* d_calls = D.d_calls.as<CallOut>(std::max<uint32_t>(ncalls, 1));
    CCA_HIP(hipMemcpyAsync(d_items, hi, sizeof(CodeItem) * nitems, hipMemcpyHostToDevice, st));
    CCA_HIP(hipMemcpyAsync(d_first, hf, sizeof(uint32_t) * (ncalls + 1), hipMemcpyHostToDevice, st));

    const HeaderFormat fmt = header_format(I->dialect, I->log2);
    CoderParams p{I->cbs, I->log2, (uint32_t)I->dialect, fmt.count_bits, fmt.symbits_bits, 0};
    uint64_t cap = std::max<uint64_t>(D.scratch_cap, arena_estimate(calls, I->log2, I->cbs));
    for (int attempt = 0;; attempt++) {
        const uint64_t tail = 2ull * (nitems + 1) + 16;  // item offsets (u64)
        uint32_t* scratch = D.d_scratch.as<uint32_t>(cap + tail);
        D.scratch_cap = cap;
        coder_encode(d_items, nitems, d_seqs, nseq, p, symSize, scratch, cap, d_state, d_calls, ncalls, d_first, d_out,
                     out_cap, total_syms, st);
        if (!wait) return 0;
        const size_t rbytes = sizeof(CoderState) + sizeof(CallOut) * ncalls;
        unsigned char* r = (unsigned char*)D.h_result.get(rbytes);
        CCA_HIP(hipMemcpyAsync(r, d_state, sizeof(CoderState), hipMemcpyDeviceToHost, st));
        CCA_HIP(hipMemcpyAsync(r + sizeof(CoderState), d_calls, sizeof(CallOut) * ncalls, hipMemcpyDeviceToHost, st));
        CCA_HIP(hipStreamSynchronize(st));
        const CoderState* hs = (const CoderState*)r;
        if (hs->err == kErrScratch && attempt < 4) {
            cap = std::max<uint64_t>(2 * cap, hs->scratch_used + hs->scratch_used / 2);
            continue;
        }
        if (results) results->assign((const CallOut*)(r + sizeof(CoderState)), (const CallOut*)(r + sizeof(CoderState)) + ncalls);
        return hs->out_words;
    }
}

}  // namespace cca
