// Host side of the device-resident coder: one pinned upload of the item list, the kernel
// pipelines of hip/coder.hip and hip/decoder.hip, one download of the per-call results.
#include "batch.hpp"

#include <algorithm>
#include <cstring>

#include "huffman_table.hpp"

namespace cca {

namespace {

// Arena estimate: histograms, tables and symbol lists of the usual (LDS-class) size, codewords
// at up to one word per symbol, block offsets.  Too small is detected on the device and retried.
uint64_t arena_estimate(const std::vector<ItemSpec>& items, uint32_t log2, uint32_t cbs) {
    const uint64_t alpha = std::min<uint64_t>(1ull << log2, 8192);
    uint64_t words = 4096;
    for (const ItemSpec& it : items) {
        const uint64_t nseq = it.rl ? 2 : 1;
        words += nseq * ((uint64_t)it.n + it.n / cbs + 4 * std::min<uint64_t>(alpha, 2ull * it.n + 64) + 128);
    }
    return words;
}

CoderParams params_of(const cc_instance* I) {
    const HeaderFormat fmt = header_format(I->dialect, I->log2);
    return CoderParams{I->cbs, I->log2, (uint32_t)I->dialect, fmt.count_bits, fmt.symbits_bits, 0};
}

void check_symbols(const cc_instance* I, int symSize) {
    const bool longSymbols = I->log2 > 16;
    CCA_REQUIRE(symSize == (longSymbols ? 4 : 2), CC_ERR_INVALID_ARG,
                "symbol size does not match the instance (log2HuffmanDistinctSymbolCountMax > 16 uses Symbol32)");
    CCA_REQUIRE(I->dialect == CC_DIALECT_GPU || symSize == 2, CC_ERR_INVALID_ARG, "the CPU dialect is Symbol16 only");
}

}  // namespace

uint64_t device_encode_items(cc_instance* I, const std::vector<ItemSpec>& items, uint32_t ncalls, int symSize,
                             uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    DeviceCoder& D = I->coder;
    check_symbols(I, symSize);
    const uint32_t nitems = (uint32_t)items.size();
    const size_t o_first = align_up(sizeof(CodeItem) * std::max<uint32_t>(nitems, 1), 256);
    const size_t stage_bytes = o_first + sizeof(uint32_t) * (ncalls + 1);
    unsigned char* h = (unsigned char*)D.h_stage.get(stage_bytes);
    CodeItem* hi = (CodeItem*)h;
    uint32_t* hf = (uint32_t*)(h + o_first);
    uint32_t seq = 0, call = 0;
    uint64_t total_syms = 0;
    hf[0] = 0;
    for (uint32_t i = 0; i < nitems; i++) {
        const ItemSpec& it = items[i];
        CCA_REQUIRE(it.call >= call && it.call < ncalls, CC_ERR_INVALID_ARG, "items must be grouped by call in order");
        while (call < it.call) hf[++call] = i;
        CCA_REQUIRE(it.ptr != nullptr || it.n == 0, CC_ERR_INVALID_ARG, "null stream");
        hi[i] = CodeItem{it.ptr, it.n, it.rl ? 1u : 0u, seq, it.call};
        seq += it.rl ? 2 : 1;
        total_syms += it.n;
    }
    while (call < ncalls) hf[++call] = nitems;
    const uint32_t nseq = seq;
    CodeItem* d_items = D.d_items.as<CodeItem>(std::max<uint32_t>(nitems, 1));
    uint32_t* d_first = D.d_call_first.as<uint32_t>(ncalls + 1);
    SeqInfo* d_seqs = D.d_seqs.as<SeqInfo>(std::max<uint32_t>(nseq, 1));
    CoderState* d_state = D.d_state.as<CoderState>(1);
    CallOut* d_calls = D.d_calls.as<CallOut>(std::max<uint32_t>(ncalls, 1));
    CCA_HIP(hipMemcpyAsync(d_items, hi, sizeof(CodeItem) * nitems, hipMemcpyHostToDevice, st));
    CCA_HIP(hipMemcpyAsync(d_first, hf, sizeof(uint32_t) * (ncalls + 1), hipMemcpyHostToDevice, st));
    const CoderParams p = params_of(I);
    uint64_t cap = std::max<uint64_t>(D.scratch_cap, arena_estimate(items, I->log2, I->cbs));
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
        if (results)
            results->assign((const CallOut*)(r + sizeof(CoderState)), (const CallOut*)(r + sizeof(CoderState)) + ncalls);
        return hs->out_words;
    }
}

uint64_t device_encode(cc_instance* I, const std::vector<CoderCall>& calls, const std::vector<const void*>& streams,
                       int symSize, uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait) {
    std::vector<ItemSpec> items;
    for (uint32_t c = 0; c < calls.size(); c++) {
        CCA_REQUIRE((uint64_t)calls[c].first_stream + calls[c].nstreams <= streams.size(), CC_ERR_INVALID_ARG,
                    "bad stream range");
        for (uint32_t s = 0; s < calls[c].nstreams; s++)
            items.push_back(ItemSpec{const_cast<void*>(streams[calls[c].first_stream + s]), calls[c].n, calls[c].rl, c});
    }
    return device_encode_items(I, items, (uint32_t)calls.size(), symSize, d_out, out_cap, results, wait);
}

void device_decode_items(cc_instance* I, const std::vector<ItemSpec>& items, const std::vector<CallWords>& calls,
                         int symSize, std::vector<DecCallOut>* results, bool wait) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    DeviceCoder& D = I->coder;
    check_symbols(I, symSize);
    const uint32_t ncalls = (uint32_t)calls.size();
    const uint32_t nitems = (uint32_t)items.size();
    const uint32_t cbs = I->cbs;
    uint32_t nseq = 0, nwaves = 0;
    uint64_t total_syms = 0, arena = 64;
    for (const ItemSpec& it : items) {
        const uint32_t q = it.rl ? 2 : 1;
        const uint32_t nblk = div_up(it.n, cbs);
        nseq += q;
        nwaves += q * div_up(nblk, 64);
        total_syms += it.n;
        arena += q * ((uint64_t)nblk + 16 + (it.rl ? ((uint64_t)it.n * symSize + 3) / 4 + 17 : 0));
    }
    const size_t o_items = align_up(sizeof(DecCall) * std::max<uint32_t>(ncalls, 1), 256);
    const size_t o_wseq = o_items + align_up(sizeof(DecItem) * std::max<uint32_t>(nitems, 1), 256);
    const size_t o_wfirst = o_wseq + align_up(4ull * std::max<uint32_t>(nwaves, 1), 256);
    const size_t total = o_wfirst + 4ull * std::max<uint32_t>(nwaves, 1);
    unsigned char* h = (unsigned char*)D.h_dstage.get(total);
    DecCall* hc = (DecCall*)h;
    DecItem* hi = (DecItem*)(h + o_items);
    uint32_t* hws = (uint32_t*)(h + o_wseq);
    uint32_t* hwf = (uint32_t*)(h + o_wfirst);
    for (uint32_t c = 0; c < ncalls; c++) {
        CCA_REQUIRE(calls[c].words != nullptr || calls[c].nwords == 0, CC_ERR_INVALID_ARG, "null bitstream");
        hc[c] = DecCall{calls[c].words, calls[c].nwords, 0, 0};
    }
    uint32_t seq = 0, wv = 0, call = 0;
    for (uint32_t i = 0; i < nitems; i++) {
        const ItemSpec& it = items[i];
        CCA_REQUIRE(it.call >= call && it.call < ncalls, CC_ERR_INVALID_ARG, "items must be grouped by call in order");
        if (i == 0 || it.call != call) hc[it.call].first_item = i;
        call = it.call;
        hc[it.call].nitems++;
        CCA_REQUIRE(it.ptr != nullptr || it.n == 0, CC_ERR_INVALID_ARG, "null output stream");
        hi[i] = DecItem{it.ptr, it.n, it.rl ? 1u : 0u, seq, it.call};
        const uint32_t nblk = div_up(it.n, cbs);
        for (uint32_t q = 0; q < (it.rl ? 2u : 1u); q++)
            for (uint32_t b = 0; b < nblk; b += 64) {
                hws[wv] = seq + q;
                hwf[wv] = b;
                wv++;
            }
        seq += it.rl ? 2 : 1;
    }
    unsigned char* d = (unsigned char*)D.d_dcalls.get(total);
    CCA_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, st));
    DecSeq* d_seqs = D.d_dseqs.as<DecSeq>(std::max<uint32_t>(nseq, 1));
    CoderState* d_state = D.d_dstate.as<CoderState>(1);
    DecCallOut* d_out = D.d_dout.as<DecCallOut>(std::max<uint32_t>(ncalls, 1));
    uint32_t* scratch = D.d_dscratch.as<uint32_t>(arena);
    coder_decode((const DecCall*)d, ncalls, (const DecItem*)(d + o_items), nitems, d_seqs, nseq, params_of(I), symSize,
                 scratch, arena, d_state, d_out, (const uint32_t*)(d + o_wseq), (const uint32_t*)(d + o_wfirst), nwaves,
                 total_syms, st);
    if (!wait) return;
    const size_t rbytes = sizeof(CoderState) + sizeof(DecCallOut) * ncalls;
    unsigned char* r = (unsigned char*)D.h_dresult.get(rbytes);
    CCA_HIP(hipMemcpyAsync(r, d_state, sizeof(CoderState), hipMemcpyDeviceToHost, st));
    CCA_HIP(hipMemcpyAsync(r + sizeof(CoderState), d_out, sizeof(DecCallOut) * ncalls, hipMemcpyDeviceToHost, st));
    CCA_HIP(hipStreamSynchronize(st));
    if (results)
        results->assign((const DecCallOut*)(r + sizeof(CoderState)),
                        (const DecCallOut*)(r + sizeof(CoderState)) + ncalls);
}

void device_decode(cc_instance* I, const std::vector<DecodeCall>& calls, const std::vector<void*>& outs, int symSize,
                   std::vector<DecCallOut>* results, bool wait) {
    std::vector<ItemSpec> items;
    std::vector<CallWords> cw;
    for (uint32_t c = 0; c < calls.size(); c++) {
        CCA_REQUIRE((uint64_t)calls[c].first_stream + calls[c].nstreams <= outs.size(), CC_ERR_INVALID_ARG,
                    "bad stream range");
        cw.push_back(CallWords{calls[c].words, calls[c].nwords});
        for (uint32_t s = 0; s < calls[c].nstreams; s++)
            items.push_back(ItemSpec{outs[calls[c].first_stream + s], calls[c].n, calls[c].rl, c});
    }
    device_decode_items(I, items, cw, symSize, results, wait);
}

}  // namespace cca
