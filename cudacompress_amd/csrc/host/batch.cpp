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

// Side streams of a coder context (created on first use, on the current device).
const SideStreams* side_streams(DeviceCoder& D, SideStreams& out) {
    if (!D.fork) {
        for (int i = 0; i < kSideStreams; i++) {
            CCA_HIP(hipStreamCreateWithFlags(&D.side[i], hipStreamNonBlocking));
            CCA_HIP(hipEventCreateWithFlags(&D.join[i], hipEventDisableTiming));
        }
        CCA_HIP(hipEventCreateWithFlags(&D.fork, hipEventDisableTiming));
    }
    for (int i = 0; i < kSideStreams; i++) {
        out.s[i] = D.side[i];
        out.join[i] = D.join[i];
    }
    out.fork = D.fork;
    return &out;
}

}  // namespace

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

// ------------------------------------------------------------------------------------- encode
void encode_prepare(cc_instance* I, DeviceCoder& D, const std::vector<ItemSpec>& items, uint32_t ncalls, int symSize,
                    EncodeJob& job, hipStream_t st) {
    check_symbols(I, symSize);
    const uint32_t nitems = (uint32_t)items.size();
    const size_t o_first = align_up(sizeof(CodeItem) * std::max<uint32_t>(nitems, 1), 256);
    const size_t stage_bytes = o_first + sizeof(uint32_t) * (ncalls + 1 + nitems);
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
    // the stats and pack passes take the items largest first (their serial chunk loops start first)
    uint32_t* ho = hf + ncalls + 1;
    for (uint32_t i = 0; i < nitems; i++) ho[i] = i;
    std::stable_sort(ho, ho + nitems, [&](uint32_t a, uint32_t b) { return items[a].n > items[b].n; });
    job.nitems = nitems;
    job.nseq = seq;
    job.ncalls = ncalls;
    job.total_syms = total_syms;
    job.symSize = symSize;
    job.p = params_of(I);
    for (int c = 0; c < 4; c++) job.design_grid[c] = I->tuning.designGrid[c];
    job.lane_k = I->tuning.laneDesignK;
    D.d_items.as<CodeItem>(std::max<uint32_t>(nitems, 1));
    D.d_call_first.as<uint32_t>(ncalls + 1 + std::max<uint32_t>(nitems, 1));
    D.d_seqs.as<SeqInfo>(std::max<uint32_t>(job.nseq, 1));
    D.d_state.as<CoderState>(1);
    D.d_classes.as<uint32_t>(kDesignListBase + (uint64_t)kDesignLists * std::max<uint32_t>(job.nseq, 1));
    D.d_calls.as<CallOut>(std::max<uint32_t>(ncalls, 1));
    D.h_result.get(sizeof(CoderState) + sizeof(CallOut) * std::max<uint32_t>(ncalls, 1));
    CCA_HIP(hipMemcpyAsync(D.d_items.p, hi, sizeof(CodeItem) * nitems, hipMemcpyHostToDevice, st));
    CCA_HIP(hipMemcpyAsync(D.d_call_first.p, hf, sizeof(uint32_t) * (ncalls + 1 + nitems), hipMemcpyHostToDevice, st));
    encode_reserve(D, job, std::max<uint64_t>(D.scratch_cap, arena_estimate(items, I->log2, I->cbs)));
    SideStreams ss;
    side_streams(D, ss);
}

void encode_reserve(DeviceCoder& D, EncodeJob& job, uint64_t words) {
    // item offsets and the layout's block totals (u64) after the arena
    const uint64_t tail = 2ull * (job.nitems + 1) + 2ull * (div_up(job.nitems, 1024u) + 1) + 16;
    D.d_scratch.as<uint32_t>(words + tail);
    D.scratch_cap = words;
    job.scratch_cap = words;
}

void encode_launch(DeviceCoder& D, const EncodeJob& job, uint32_t* d_out, uint64_t out_cap, hipStream_t st,
                   const LayoutLink* link) {
    SideStreams ss;
    coder_encode((const CodeItem*)D.d_items.p, job.nitems, (SeqInfo*)D.d_seqs.p, job.nseq, job.p, job.symSize,
                 (uint32_t*)D.d_scratch.p, job.scratch_cap, (CoderState*)D.d_state.p, (uint32_t*)D.d_classes.p,
                 (CallOut*)D.d_calls.p,
                 job.ncalls, (const uint32_t*)D.d_call_first.p, d_out, out_cap, job.total_syms, st,
                 side_streams(D, ss), link, job.design_grid, job.lane_k,
                 (const uint32_t*)D.d_call_first.p + job.ncalls + 1);
}

CoderState encode_finish(DeviceCoder& D, const EncodeJob& job, std::vector<CallOut>* results, hipStream_t st) {
    unsigned char* r =
        (unsigned char*)D.h_result.get(sizeof(CoderState) + sizeof(CallOut) * std::max<uint32_t>(job.ncalls, 1));
    CCA_HIP(hipMemcpyAsync(r, D.d_state.p, sizeof(CoderState), hipMemcpyDeviceToHost, st));
    if (job.ncalls)
        CCA_HIP(hipMemcpyAsync(r + sizeof(CoderState), D.d_calls.p, sizeof(CallOut) * job.ncalls,
                               hipMemcpyDeviceToHost, st));
    CCA_HIP(hipStreamSynchronize(st));
    if (results)
        results->assign((const CallOut*)(r + sizeof(CoderState)),
                        (const CallOut*)(r + sizeof(CoderState)) + job.ncalls);
    return *(const CoderState*)r;
}

uint64_t device_encode_items(cc_instance* I, const std::vector<ItemSpec>& items, uint32_t ncalls, int symSize,
                             uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    DeviceCoder& D = I->coder;
    EncodeJob job;
    encode_prepare(I, D, items, ncalls, symSize, job, st);
    for (int attempt = 0;; attempt++) {
        encode_launch(D, job, d_out, out_cap, st);
        if (!wait) return 0;
        const CoderState s = encode_finish(D, job, results, st);
        if (s.err == kErrScratch && attempt < 4) {
            encode_reserve(D, job, std::max<uint64_t>(2 * job.scratch_cap, s.scratch_used + s.scratch_used / 2));
            continue;
        }
        return s.out_words;
    }
}

uint64_t device_encode(cc_instance* I, const std::vector<CoderCall>& calls, const std::vector<const void*>& streams,
                       int symSize, uint32_t* d_out, uint64_t out_cap, std::vector<CallOut>* results, bool wait) {
    std::vector<ItemSpec> items;
    for (uint32_t c = 0; c < calls.size(); c++) {
        CCA_REQUIRE((uint64_t)calls[c].first_stream + calls[c].nstreams <= streams.size(), CC_ERR_INVALID_ARG,
                    "bad stream range");
        for (uint32_t s = 0; s < calls[c].nstreams; s++)
            items.push_back(
                ItemSpec{const_cast<void*>(streams[calls[c].first_stream + s]), calls[c].n, calls[c].rl, c});
    }
    return device_encode_items(I, items, (uint32_t)calls.size(), symSize, d_out, out_cap, results, wait);
}

// ------------------------------------------------------------------------------------- decode
namespace {
// Block-decoder entries of sequence `seq` of at most nblk blocks: workgroups of kDecGroupBlocks
// blocks at wl (long sequences), one wave at ws (short ones).
void add_entries(uint32_t seq, uint32_t nblk, uint32_t* hws, uint32_t* hwf, uint32_t& wl, uint32_t& ws) {
    if (dec_long(nblk)) {
        for (uint32_t b = 0; b < nblk; b += kDecGroupBlocks) {
            hws[wl] = seq;
            hwf[wl] = b;
            wl++;
        }
    } else if (nblk) {
        hws[ws] = seq;
        hwf[ws] = 0;
        ws++;
    }
}

// first[i]: the first zero-run expansion chunk of item i (nitems + 1 entries); returns the total.
uint32_t chunk_table(const std::vector<ItemSpec>& items, uint32_t* first) {
    uint64_t k = 0;
    for (size_t i = 0; i < items.size(); i++) {
        first[i] = (uint32_t)k;
        k += rle_chunks(items[i].n, items[i].rl);
    }
    CCA_REQUIRE(k < (1ull << 31), CC_ERR_INVALID_ARG, "too many symbols in one decode job");
    first[items.size()] = (uint32_t)k;
    return (uint32_t)k;
}
}  // namespace

void decode_prepare(cc_instance* I, DeviceCoder& D, const std::vector<ItemSpec>& items,
                    const std::vector<CallWords>& calls, int symSize, DecodeJob& job, hipStream_t st) {
    check_symbols(I, symSize);
    const uint32_t ncalls = (uint32_t)calls.size();
    const uint32_t nitems = (uint32_t)items.size();
    const uint32_t cbs = I->cbs;
    uint32_t nseq = 0, nlong = 0, nshort = 0;
    uint64_t total_syms = 0, arena = 64;
    for (const ItemSpec& it : items) {
        const uint32_t q = it.rl ? 2 : 1;
        const uint32_t nblk = div_up(it.n, cbs);
        nseq += q;
        (dec_long(nblk) ? nlong : nshort) += q * dec_entries(nblk);
        total_syms += it.n;
        // block offsets, lookup + symbol table (nsym <= count <= n), RLE decode buffers
        arena += q * ((uint64_t)nblk + 16 + (1u << kDecLutBits) + it.n + 17 + (it.rl ? ((uint64_t)it.n * symSize + 3) / 4 + 17 : 0));
    }
    const uint32_t nwaves = nlong + nshort;
    job.o_items = align_up(sizeof(DecCall) * std::max<uint32_t>(ncalls, 1), 256);
    job.o_wseq = job.o_items + align_up(sizeof(DecItem) * std::max<uint32_t>(nitems, 1), 256);
    job.o_wfirst = job.o_wseq + align_up(4ull * std::max<uint32_t>(nwaves, 1), 256);
    job.o_ifirst = job.o_wfirst + align_up(4ull * std::max<uint32_t>(nwaves, 1), 256);
    const size_t total = job.o_ifirst + 4ull * (nitems + 1);
    unsigned char* h = (unsigned char*)D.h_dstage.get(total);
    DecCall* hc = (DecCall*)h;
    DecItem* hi = (DecItem*)(h + job.o_items);
    uint32_t* hws = (uint32_t*)(h + job.o_wseq);
    uint32_t* hwf = (uint32_t*)(h + job.o_wfirst);
    job.nchunks = chunk_table(items, (uint32_t*)(h + job.o_ifirst));
    job.o_sums = align_up(total, 256);
    for (uint32_t c = 0; c < ncalls; c++) hc[c] = DecCall{calls[c].words, calls[c].nwords, 0, 0};
    uint32_t seq = 0, wl = 0, ws = nlong, call = 0;
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
            add_entries(seq + q, nblk, hws, hwf, wl, ws);
        seq += it.rl ? 2 : 1;
    }
    job.nitems = nitems;
    job.nseq = nseq;
    job.ncalls = ncalls;
    job.nwaves = nwaves;
    job.nlong = nlong;
    job.total_syms = total_syms;
    job.arena = arena;
    job.symSize = symSize;
    job.p = params_of(I);
    D.d_dcalls.get(job.o_sums + 4ull * std::max<uint32_t>(job.nchunks, 1));
    D.d_dseqs.as<DecSeq>(std::max<uint32_t>(nseq, 1));
    D.d_dstate.as<CoderState>(1);
    D.d_dout.as<DecCallOut>(std::max<uint32_t>(ncalls, 1));
    D.d_dscratch.as<uint32_t>(arena);
    D.h_dresult.get(sizeof(CoderState) + sizeof(DecCallOut) * std::max<uint32_t>(ncalls, 1));
    CCA_HIP(hipMemcpyAsync(D.d_dcalls.p, h, total, hipMemcpyHostToDevice, st));
}

void decode_launch(DeviceCoder& D, const DecodeJob& job, hipStream_t st) {
    const unsigned char* d = (const unsigned char*)D.d_dcalls.p;
    coder_decode((const DecCall*)d, job.ncalls, (const DecItem*)(d + job.o_items), job.nitems, (DecSeq*)D.d_dseqs.p,
                 job.nseq, job.p, job.symSize, (uint32_t*)D.d_dscratch.p, job.arena, (CoderState*)D.d_dstate.p,
                 (DecCallOut*)D.d_dout.p, (const uint32_t*)(d + job.o_wseq), (const uint32_t*)(d + job.o_wfirst),
                 job.nwaves, job.nlong, job.total_syms,
                 RleChunks{(const uint32_t*)(d + job.o_ifirst), job.nchunks,
                           (uint32_t*)((unsigned char*)D.d_dcalls.p + job.o_sums)},
                 st);
}

void decode_finish(DeviceCoder& D, const DecodeJob& job, std::vector<DecCallOut>* results, hipStream_t st) {
    unsigned char* r =
        (unsigned char*)D.h_dresult.get(sizeof(CoderState) + sizeof(DecCallOut) * std::max<uint32_t>(job.ncalls, 1));
    CCA_HIP(hipMemcpyAsync(r, D.d_dstate.p, sizeof(CoderState), hipMemcpyDeviceToHost, st));
    if (job.ncalls)
        CCA_HIP(hipMemcpyAsync(r + sizeof(CoderState), D.d_dout.p, sizeof(DecCallOut) * job.ncalls,
                               hipMemcpyDeviceToHost, st));
    CCA_HIP(hipStreamSynchronize(st));
    if (results)
        results->assign((const DecCallOut*)(r + sizeof(CoderState)),
                        (const DecCallOut*)(r + sizeof(CoderState)) + job.ncalls);
}

void decode_index_build(cc_instance* I, DecodeIndex& X, const std::vector<ItemSpec>& items,
                        const std::vector<CallWords>& calls, int symSize, hipStream_t st) {
    check_symbols(I, symSize);
    const uint32_t ncalls = (uint32_t)calls.size();
    const uint32_t nitems = (uint32_t)items.size();
    const uint32_t cbs = I->cbs;
    X.call_first_seq.assign(ncalls, 0);
    X.call_nseq.assign(ncalls, 0);
    uint32_t nseq = 0;
    uint64_t arena = 64;
    for (const ItemSpec& it : items) {
        const uint32_t q = it.rl ? 2 : 1;
        nseq += q;
        // block offsets, lookup + symbol table (nsym <= count <= n); no RLE decode buffers
        arena += q * ((uint64_t)div_up(it.n, cbs) + 16 + (1u << kDecLutBits) + it.n + 17);
    }
    const size_t o_items = align_up(sizeof(DecCall) * std::max<uint32_t>(ncalls, 1), 256);
    const size_t total = o_items + sizeof(DecItem) * std::max<uint32_t>(nitems, 1);
    unsigned char* h = (unsigned char*)X.h_stage.get(total);
    DecCall* hc = (DecCall*)h;
    DecItem* hi = (DecItem*)(h + o_items);
    for (uint32_t c = 0; c < ncalls; c++) hc[c] = DecCall{calls[c].words, calls[c].nwords, 0, 0};
    uint32_t seq = 0, call = 0;
    for (uint32_t i = 0; i < nitems; i++) {
        const ItemSpec& it = items[i];
        CCA_REQUIRE(it.call >= call && it.call < ncalls, CC_ERR_INVALID_ARG, "items must be grouped by call in order");
        if (i == 0 || it.call != call) {
            hc[it.call].first_item = i;
            X.call_first_seq[it.call] = seq;
        }
        call = it.call;
        hc[it.call].nitems++;
        hi[i] = DecItem{nullptr, it.n, it.rl ? 1u : 0u, seq, it.call};
        seq += it.rl ? 2 : 1;
        X.call_nseq[it.call] += it.rl ? 2 : 1;
    }
    X.nseq = nseq;
    X.arena = arena;
    unsigned char* d = (unsigned char*)X.d_calls.get(total);
    X.d_seqs.as<DecSeq>(std::max<uint32_t>(nseq, 1));
    X.d_state.as<CoderState>(1);
    X.d_out.as<DecCallOut>(std::max<uint32_t>(ncalls, 1));
    X.d_scratch.as<uint32_t>(arena);
    CCA_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, st));
    coder_decode_index((const DecCall*)d, ncalls, (const DecItem*)(d + o_items), (DecSeq*)X.d_seqs.p, nseq,
                       params_of(I), symSize, (uint32_t*)X.d_scratch.p, arena, (CoderState*)X.d_state.p,
                       (DecCallOut*)X.d_out.p, st);
    X.results.resize(ncalls);
    if (ncalls)
        CCA_HIP(hipMemcpyAsync(X.results.data(), X.d_out.p, sizeof(DecCallOut) * ncalls, hipMemcpyDeviceToHost, st));
    CCA_HIP(hipStreamSynchronize(st));
}

void decode_prepare_bound(cc_instance* I, DeviceCoder& D, const DecodeIndex& X, const std::vector<ItemSpec>& items,
                          const std::vector<uint32_t>& index_calls, int symSize, BoundJob& job, hipStream_t st) {
    check_symbols(I, symSize);
    const uint32_t ncalls = (uint32_t)index_calls.size();
    const uint32_t nitems = (uint32_t)items.size();
    const uint32_t cbs = I->cbs;
    uint32_t nseq = 0, nlong = 0, nshort = 0;
    uint64_t total_syms = 0, rl_bytes = 0;
    for (const ItemSpec& it : items) {
        const uint32_t q = it.rl ? 2 : 1;
        const uint32_t nblk = div_up(it.n, cbs);
        nseq += q;
        (dec_long(nblk) ? nlong : nshort) += q * dec_entries(nblk);
        total_syms += it.n;
    }
    const uint32_t nwaves = nlong + nshort;
    job.o_items = align_up(sizeof(DecCall) * std::max<uint32_t>(ncalls, 1), 256);
    job.o_bind = job.o_items + align_up(sizeof(DecItem) * std::max<uint32_t>(nitems, 1), 256);
    job.o_cbrick = job.o_bind + align_up(sizeof(DecBind) * std::max<uint32_t>(nseq, 1), 256);
    job.o_wseq = job.o_cbrick + align_up(4ull * std::max<uint32_t>(ncalls, 1), 256);
    job.o_wfirst = job.o_wseq + align_up(4ull * std::max<uint32_t>(nwaves, 1), 256);
    job.o_ifirst = job.o_wfirst + align_up(4ull * std::max<uint32_t>(nwaves, 1), 256);
    const size_t total = job.o_ifirst + 4ull * (nitems + 1);
    unsigned char* h = (unsigned char*)D.h_dstage.get(total);
    job.nchunks = chunk_table(items, (uint32_t*)(h + job.o_ifirst));
    job.o_sums = align_up(total, 256);
    DecCall* hc = (DecCall*)h;
    DecItem* hi = (DecItem*)(h + job.o_items);
    DecBind* hb = (DecBind*)(h + job.o_bind);
    uint32_t* hcb = (uint32_t*)(h + job.o_cbrick);
    uint32_t* hws = (uint32_t*)(h + job.o_wseq);
    uint32_t* hwf = (uint32_t*)(h + job.o_wfirst);
    for (uint32_t c = 0; c < ncalls; c++) {
        CCA_REQUIRE(index_calls[c] < X.call_first_seq.size(), CC_ERR_INVALID_ARG, "index call out of range");
        hc[c] = DecCall{nullptr, 0, 0, 0};
        hcb[c] = index_calls[c];
    }
    uint32_t seq = 0, wl = 0, ws = nlong, call = 0, local = 0;
    for (uint32_t i = 0; i < nitems; i++) {
        const ItemSpec& it = items[i];
        CCA_REQUIRE(it.call >= call && it.call < ncalls, CC_ERR_INVALID_ARG, "items must be grouped by call in order");
        if (i == 0 || it.call != call) {
            hc[it.call].first_item = i;
            local = 0;
        }
        call = it.call;
        hc[it.call].nitems++;
        CCA_REQUIRE(it.ptr != nullptr || it.n == 0, CC_ERR_INVALID_ARG, "null output stream");
        hi[i] = DecItem{it.ptr, it.n, it.rl ? 1u : 0u, seq, it.call};
        const uint32_t nblk = div_up(it.n, cbs);
        for (uint32_t q = 0; q < (it.rl ? 2u : 1u); q++) {
            CCA_REQUIRE(local < X.call_nseq[index_calls[it.call]], CC_ERR_INVALID_ARG,
                        "request items do not match the index");
            uint64_t off = ~0ull;
            if (it.rl) {
                off = rl_bytes;
                rl_bytes += align_up((uint64_t)it.n * symSize + 4, 64);
            }
            hb[seq + q] = DecBind{X.call_first_seq[index_calls[it.call]] + local, i, off};
            local++;
            add_entries(seq + q, nblk, hws, hwf, wl, ws);
        }
        seq += it.rl ? 2 : 1;
    }
    job.nitems = nitems;
    job.nseq = nseq;
    job.ncalls = ncalls;
    job.nwaves = nwaves;
    job.nlong = nlong;
    job.total_syms = total_syms;
    job.rl_bytes = rl_bytes;
    job.symSize = symSize;
    job.p = params_of(I);
    D.d_dcalls.get(job.o_sums + 4ull * std::max<uint32_t>(job.nchunks, 1));
    D.d_dseqs.as<DecSeq>(std::max<uint32_t>(nseq, 1));
    D.d_dstate.as<CoderState>(1);
    D.d_dout.as<DecCallOut>(std::max<uint32_t>(ncalls, 1));
    D.d_dscratch.get(std::max<uint64_t>(rl_bytes, 64));
    D.h_dresult.get(sizeof(CoderState) + sizeof(DecCallOut) * std::max<uint32_t>(ncalls, 1));
    CCA_HIP(hipMemcpyAsync(D.d_dcalls.p, h, total, hipMemcpyHostToDevice, st));
}

void decode_launch_bound(DeviceCoder& D, const DecodeIndex& X, const BoundJob& job, hipStream_t st) {
    const unsigned char* d = (const unsigned char*)D.d_dcalls.p;
    coder_decode_bound((const DecSeq*)X.d_seqs.p, (const uint32_t*)X.d_scratch.p, (const DecCallOut*)X.d_out.p,
                       (const DecBind*)(d + job.o_bind), (const uint32_t*)(d + job.o_cbrick), (const DecCall*)d,
                       job.ncalls, (const DecItem*)(d + job.o_items), job.nitems, (DecSeq*)D.d_dseqs.p, job.nseq,
                       job.p, job.symSize, (unsigned char*)D.d_dscratch.p, (CoderState*)D.d_dstate.p,
                       (DecCallOut*)D.d_dout.p, (const uint32_t*)(d + job.o_wseq),
                       (const uint32_t*)(d + job.o_wfirst), job.nwaves, job.nlong, job.total_syms,
                       RleChunks{(const uint32_t*)(d + job.o_ifirst), job.nchunks,
                                 (uint32_t*)((unsigned char*)D.d_dcalls.p + job.o_sums)},
                       st);
}

void decode_calls_from_encode(DeviceCoder& D, const DecodeJob& job, const CallOut* d_enc, const uint32_t* base,
                              uint64_t cap, hipStream_t st) {
    coder_calls_from_encode((DecCall*)D.d_dcalls.p, d_enc, base, cap, job.ncalls, st);
}

void device_decode_items(cc_instance* I, const std::vector<ItemSpec>& items, const std::vector<CallWords>& calls,
                         int symSize, std::vector<DecCallOut>* results, bool wait) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    DeviceCoder& D = I->coder;
    for (const CallWords& c : calls)
        CCA_REQUIRE(c.words != nullptr || c.nwords == 0, CC_ERR_INVALID_ARG, "null bitstream");
    DecodeJob job;
    decode_prepare(I, D, items, calls, symSize, job, st);
    decode_launch(D, job, st);
    if (wait) decode_finish(D, job, results, st);
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
