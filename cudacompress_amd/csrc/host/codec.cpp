// Host orchestration of the entropy coder: the equivalent of src/cudaCompress/Encode.cpp:208-917.
//
// Encode (per call, all streams batched into each launch):
//   1. zero-run RLE (rl) or per-stream max (huff)        -> sync #1: pair counts, maxima
//   2. histograms of every coded stream                   -> sync #2: histograms
//   3. table design on the host (libstdc++ heap/sort, as the reference), exact bit counts from
//      histogram x length, encode tables uploaded once
//   4. Huffman block lengths, offsets scan, pack; codewords and offsets downloaded -> sync #3
//   5. bitstream layout (Encode.cpp:402-463 rl, :751-782 huff) written on the host.
// The reference does >= streams + 6 host synchronisations per encodeRLHuff; this does 3.
//
// Decode parses every header on the host (tables, bit counts, offsets -- as Encode.cpp:522-621),
// stages tables, codewords and absolute offsets in one pinned buffer, uploads it with a single
// copy and decodes all coded streams in one launch; RLE decode writes every output element.
#include "codec.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>

#include "../kernels.hpp"
#include "bitio.hpp"
#include "huffman_table.hpp"

namespace cca {
namespace {

constexpr uint32_t kPackSyms = 2048;

template <class F>
void parallel_for(size_t n, size_t work_hint, F f) {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    size_t nt = std::min<size_t>({(size_t)hw, n, (size_t)16});
    if (nt <= 1 || work_hint < (1u << 16)) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<size_t> next{0};
    for (size_t t = 0; t < nt; t++)
        th.emplace_back([&]() {
            for (size_t i = next++; i < n; i = next++) f(i);
        });
    for (auto& x : th) x.join();
}

struct Staging {  // bump allocator over a pinned host buffer mirrored by a device buffer
    unsigned char* h;
    unsigned char* d;
    size_t off = 0;
    template <class T>
    size_t push(size_t count) {
        off = align_up(off, 256);
        size_t o = off;
        off += count * sizeof(T);
        return o;
    }
};

}  // namespace

void encode_streams(cc_instance* I, bool rl, int symSize, const void* const* streams, uint32_t ns, uint32_t n,
                    const uint32_t* startBits) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    const bool longSymbols = I->log2 > 16;
    CCA_REQUIRE(symSize == (longSymbols ? 4 : 2), CC_ERR_INVALID_ARG,
                "symbol size does not match the instance (log2HuffmanDistinctSymbolCountMax > 16 uses Symbol32)");
    CCA_REQUIRE(I->dialect == CC_DIALECT_GPU || symSize == 2, CC_ERR_INVALID_ARG, "the CPU dialect is Symbol16 only");
    CCA_REQUIRE(ns == 0 || streams != nullptr, CC_ERR_INVALID_ARG, "null stream array");
    for (uint32_t i = 0; i < ns; i++) CCA_REQUIRE(streams[i] != nullptr || n == 0, CC_ERR_INVALID_ARG, "null stream");
    if (startBits)
        for (uint32_t i = 0; i < ns; i++) CCA_REQUIRE(startBits[i] < 32, CC_ERR_INVALID_ARG, "start bit offset >= 32");
    const uint32_t cbs = I->cbs;
    const HeaderFormat fmt = header_format(I->dialect, I->log2);
    I->out_words.clear();
    I->seg_begin.assign(ns, 0);
    I->seg_bits.assign(ns, 0);
    if (ns == 0) return;
    SectionTimer call(I, 1), stage(I, 2);
    call(rl ? "Encode RLHuff" : "Encode Huff");

    // ---- 1. RLE / maxima --------------------------------------------------------------------
    stage(rl ? "Encode: Run Length" : "Encode: Maxima");
    const uint64_t cap = align_up(std::max<uint32_t>(n, 1), 64);
    const uint32_t tiles_per = rl ? div_up(n, kRleTile) : 0;
    const uint32_t ntiles = tiles_per * ns;
    {
        Staging s{(unsigned char*)I->h_up.get(0), nullptr};
        size_t o_desc = s.push<RleStream>(ns);
        size_t o_map = s.push<uint32_t>(ntiles);
        unsigned char* h = (unsigned char*)I->h_up.get(s.off);
        unsigned char* d = (unsigned char*)I->d_desc.get(s.off);
        RleStream* hd = (RleStream*)(h + o_desc);
        uint32_t* hm = (uint32_t*)(h + o_map);
        unsigned char* dsym = rl ? (unsigned char*)I->d_rle_sym.get(cap * ns * symSize) : nullptr;
        unsigned char* dzc = rl ? (unsigned char*)I->d_rle_zc.get(cap * ns * symSize) : nullptr;
        for (uint32_t i = 0; i < ns; i++) {
            hd[i].in = streams[i];
            hd[i].out_sym = rl ? dsym + cap * i * symSize : nullptr;
            hd[i].out_zc = rl ? dzc + cap * i * symSize : nullptr;
            hd[i].n = n;
            hd[i].tile_base = i * tiles_per;
            for (uint32_t t = 0; t < tiles_per; t++) hm[i * tiles_per + t] = i;
        }
        CCA_HIP(hipMemcpyAsync(d, h, s.off, hipMemcpyHostToDevice, st));
        RleResult* dres = I->d_results.as<RleResult>(ns);
        if (rl) {
            RleTile* dt = I->d_tiles.as<RleTile>(std::max<uint32_t>(ntiles, 1));
            rle_encode((const RleStream*)(d + o_desc), ns, dt, ntiles, (const uint32_t*)(d + o_map), dres, symSize,
                       (uint64_t)n * ns, st);
        } else {
            stream_max((const RleStream*)(d + o_desc), ns, n, dres, symSize, st);
        }
        CCA_HIP(hipMemcpyAsync(I->h_down.get(sizeof(RleResult) * ns), dres, sizeof(RleResult) * ns,
                               hipMemcpyDeviceToHost, st));
        CCA_HIP(hipStreamSynchronize(st));
    }
    std::vector<RleResult> res((RleResult*)I->h_down.p, (RleResult*)I->h_down.p + ns);

    // ---- 2. histograms ----------------------------------------------------------------------
    stage("Encode: Histograms");
    const uint32_t nt = rl ? 2 * ns : ns;
    struct Coded {
        const void* data;
        uint32_t count, smax;
        uint32_t quirk;
        uint64_t hist_off;
    };
    std::vector<Coded> coded(nt);
    uint32_t callmax_sym = 0, callmax_zc = 0;
    uint64_t hist_total = 0;
    for (uint32_t i = 0; i < ns; i++) {
        const RleResult& r = res[i];
        if (rl) {
            const unsigned char* dsym = (const unsigned char*)I->d_rle_sym.p + cap * i * symSize;
            const unsigned char* dzc = (const unsigned char*)I->d_rle_zc.p + cap * i * symSize;
            coded[2 * i] = Coded{dsym, r.m, r.m ? r.sym_max : 0u, 0, 0};
            coded[2 * i + 1] = Coded{dzc, r.m, r.m ? r.zc_max : 0u, 0, 0};
            callmax_sym = std::max(callmax_sym, coded[2 * i].smax);
            callmax_zc = std::max(callmax_zc, coded[2 * i + 1].smax);
        } else {
            coded[i] = Coded{streams[i], n, r.sym_max, 0, 0};
            callmax_sym = std::max(callmax_sym, r.sym_max);
        }
    }
    if (fmt.dialect == CC_DIALECT_GPU) {
        // HuffmanTable.cu:409-419 (checked per design call: symbols, then zero counts)
        CCA_REQUIRE((uint64_t)callmax_sym + 1 <= (1ull << I->log2), CC_ERR_SYMBOL_OVERFLOW,
                    "symbol exceeds 2^log2HuffmanDistinctSymbolCountMax");
        CCA_REQUIRE((uint64_t)callmax_zc + 1 <= (1ull << I->log2), CC_ERR_SYMBOL_OVERFLOW,
                    "zero count exceeds 2^log2HuffmanDistinctSymbolCountMax");
    }
    for (uint32_t j = 0; j < nt; j++) {
        Coded& c = coded[j];
        const bool is_sym = !rl || (j % 2 == 0);
        // Symbol16 streams are padded with 0xFFFF to a multiple of 8 before the histogram
        // (Histogram.cu:90-96, Encode.cpp:260-264); when the design call has 65536 bins the
        // 32-bit-word histogram pass counts one pad in bin 65535 for an odd symbol count.
        c.quirk = (fmt.dialect == CC_DIALECT_GPU && symSize == 2 && is_sym && c.smax == 65535u && (c.count & 1u)) ? 1u : 0u;
        c.hist_off = hist_total;
        hist_total += align_up((uint64_t)c.smax + 1, 64);
    }
    {
        uint32_t* dh = I->d_hist.as<uint32_t>(hist_total);
        CCA_HIP(hipMemsetAsync(dh, 0, hist_total * 4, st));
        Staging s{nullptr, nullptr};
        size_t o_desc = s.push<HistStream>(nt);
        HistStream* hd = (HistStream*)((unsigned char*)I->h_up.get(s.off) + o_desc);
        for (uint32_t j = 0; j < nt; j++)
            hd[j] = HistStream{coded[j].data, coded[j].count, coded[j].smax + 1, dh + coded[j].hist_off, coded[j].quirk};
        unsigned char* d = (unsigned char*)I->d_desc.get(s.off);
        CCA_HIP(hipMemcpyAsync(d, I->h_up.p, s.off, hipMemcpyHostToDevice, st));
        histogram((const HistStream*)(d + o_desc), hd, nt, symSize, st);
        CCA_HIP(hipMemcpyAsync(I->h_hist.get(hist_total * 4), dh, hist_total * 4, hipMemcpyDeviceToHost, st));
        CCA_HIP(hipStreamSynchronize(st));
    }
    const uint32_t* hh = (const uint32_t*)I->h_hist.p;

    // ---- 3. table design ----------------------------------------------------------------------
    stage("Encode: Design Huffman Tables");
    std::vector<EncTable> tables(nt);
    std::vector<cc_status> rcs(nt, CC_OK);
    std::vector<uint64_t> bits(nt, 0);
    uint64_t design_work = 0;
    for (auto& c : coded) design_work += c.smax + 1;
    parallel_for(nt, design_work, [&](size_t j) {
        rcs[j] = design_table(fmt.dialect, hh + coded[j].hist_off, coded[j].smax, tables[j]);
        if (rcs[j] != CC_OK) return;
        uint64_t b = 0;
        const uint32_t* h = hh + coded[j].hist_off;
        for (uint32_t s = 0; s <= coded[j].smax; s++)
            if (h[s]) b += (uint64_t)h[s] * tables[j].len[s];
        if (coded[j].quirk) b -= tables[j].len[65535];  // the counted pad is not encoded
        bits[j] = b;
    });
    for (uint32_t j = 0; j < nt; j++) {
        CCA_REQUIRE(rcs[j] == CC_OK, rcs[j], "Huffman codeword longer than 31 bits");
        if (fmt.dialect == CC_DIALECT_CPU)
            CCA_REQUIRE(tables[j].symbols.size() < (1u << 14), CC_ERR_SYMBOL_OVERFLOW,
                        "more than 16383 distinct symbols cannot be described in the CPU dialect header");
        CCA_REQUIRE(bits[j] < (1ull << 32), CC_ERR_INVALID_ARG, "coded stream exceeds 2^32 bits");
    }

    // ---- 4. Huffman encode ------------------------------------------------------------------
    stage("Encode: Huffman Encode");
    std::vector<uint64_t> tab_off(nt), word_off(nt), offs_off(nt);
    std::vector<uint32_t> nblk(nt), nchk(nt);
    uint64_t tab_total = 0, word_total = 0, offs_total = 0;
    uint32_t nblocks = 0, nchunks = 0;
    for (uint32_t j = 0; j < nt; j++) {
        tab_off[j] = tab_total;
        tab_total += 2 * align_up((uint64_t)coded[j].smax + 1, 64);
        word_off[j] = word_total;
        word_total += align_up((bits[j] + 31) / 32 + 2, 64);
        nblk[j] = div_up(coded[j].count, cbs);
        nchk[j] = div_up(coded[j].count, kPackSyms);
        offs_off[j] = offs_total;
        offs_total += align_up(std::max<uint32_t>(nblk[j], 1), 64);
        nblocks += nblk[j];
        nchunks += nchk[j];
    }
    uint32_t* dtab = I->d_tables.as<uint32_t>(tab_total);
    uint32_t* dwords = I->d_words.as<uint32_t>(word_total);
    uint32_t* doffs = I->d_offsets.as<uint32_t>(offs_total);
    uint32_t* dbb = I->d_block_bits.as<uint32_t>(std::max<uint32_t>(nblocks, 1));
    uint32_t* dsb = I->d_stream_bits.as<uint32_t>(nt);
    {
        Staging s{nullptr, nullptr};
        size_t o_tab = s.push<uint32_t>(tab_total);
        size_t o_desc = s.push<HuffEncStream>(nt);
        size_t o_bmap = s.push<uint32_t>(nblocks);
        size_t o_cmap = s.push<uint32_t>(nchunks);
        unsigned char* h = (unsigned char*)I->h_up.get(s.off);
        unsigned char* d = (unsigned char*)I->d_desc.get(s.off - o_desc);
        uint32_t* htab = (uint32_t*)(h + o_tab);
        HuffEncStream* hd = (HuffEncStream*)(h + o_desc);
        uint32_t* hb = (uint32_t*)(h + o_bmap);
        uint32_t* hc = (uint32_t*)(h + o_cmap);
        uint32_t bb = 0, cb = 0;
        for (uint32_t j = 0; j < nt; j++) {
            const uint32_t nb = coded[j].smax + 1;
            memcpy(htab + tab_off[j], tables[j].code.data(), 4ull * nb);
            memcpy(htab + tab_off[j] + align_up(nb, 64), tables[j].len.data(), 4ull * nb);
            hd[j] = HuffEncStream{coded[j].data, dtab + tab_off[j], dtab + tab_off[j] + align_up(nb, 64),
                                  dwords + word_off[j], doffs + offs_off[j], coded[j].count, bb, cb, 0};
            for (uint32_t b = 0; b < nblk[j]; b++) hb[bb + b] = j;
            for (uint32_t c = 0; c < nchk[j]; c++) hc[cb + c] = j;
            bb += nblk[j];
            cb += nchk[j];
        }
        CCA_HIP(hipMemcpyAsync(dtab, h + o_tab, 4 * tab_total, hipMemcpyHostToDevice, st));
        CCA_HIP(hipMemcpyAsync(d, h + o_desc, s.off - o_desc, hipMemcpyHostToDevice, st));
        const HuffEncStream* dd = (const HuffEncStream*)d;
        uint64_t total_bits = 0;
        for (uint32_t j = 0; j < nt; j++) total_bits += bits[j];
        huffman_encode(dd, hd, nt, cbs, symSize, dbb, nblocks, (const uint32_t*)(d + (o_bmap - o_desc)), dsb,
                       (const uint32_t*)(d + (o_cmap - o_desc)), nchunks, total_bits, st);
        // download codewords and offsets (exact sizes are known from the histograms)
        uint64_t down_words = 0;
        for (uint32_t j = 0; j < nt; j++) down_words += (bits[j] + 31) / 32 + nblk[j];
        uint32_t* hw = I->h_words.as<uint32_t>(std::max<uint64_t>(down_words, 1));
        uint64_t o = 0;
        for (uint32_t j = 0; j < nt; j++) {
            uint64_t nw = (bits[j] + 31) / 32;
            if (nw) CCA_HIP(hipMemcpyAsync(hw + o, dwords + word_off[j], 4 * nw, hipMemcpyDeviceToHost, st));
            o += nw;
            if (nblk[j]) CCA_HIP(hipMemcpyAsync(hw + o, doffs + offs_off[j], 4ull * nblk[j], hipMemcpyDeviceToHost, st));
            o += nblk[j];
        }
        CCA_HIP(hipStreamSynchronize(st));
    }

    // ---- 5. bitstream layout ------------------------------------------------------------------
    stage("Encode: Bitstream Layout");
    const uint32_t* hw = (const uint32_t*)I->h_words.p;
    std::vector<uint16_t> inc;
    BitWriter w(I->out_words);
    uint64_t o = 0;
    auto put_coded = [&](uint32_t j) {
        write_table(fmt, tables[j], w);
        const uint32_t nb32 = (uint32_t)bits[j];
        w.aligned_u32(&nb32, 1);
        const uint64_t nw = (bits[j] + 31) / 32;
        w.aligned_u32(hw + o, nw);
        o += nw;
        // offsets as 16-bit increments (packInc16CPU, PackInc.cu:210-220)
        const uint32_t* off = hw + o;
        inc.resize(nblk[j]);
        uint32_t prev = 0;
        for (uint32_t b = 0; b < nblk[j]; b++) {
            inc[b] = (uint16_t)(off[b] - prev);
            prev = off[b];
        }
        w.aligned_u16(inc.data(), nblk[j]);
        o += nblk[j];
    };
    for (uint32_t i = 0; i < ns; i++) {
        const uint64_t begin = I->out_words.size();
        I->seg_begin[i] = begin;
        w.set_pos(begin * 32 + (startBits ? startBits[i] : 0u));
        if (rl) {
            const uint32_t m = res[i].m;
            w.aligned_u32(&m, 1);
            put_coded(2 * i);
            put_coded(2 * i + 1);
        } else {
            put_coded(i);
        }
        w.align32();
        I->seg_bits[i] = w.pos() - begin * 32;
    }
}

void decode_streams(cc_instance* I, bool rl, int symSize, const uint32_t* const* bsWords, const uint64_t* bsBits,
                    uint64_t* bsPos, uint32_t nbs, const uint32_t* s2b, void* const* outs, uint32_t ns, uint32_t n) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    const bool longSymbols = I->log2 > 16;
    CCA_REQUIRE(symSize == (longSymbols ? 4 : 2), CC_ERR_INVALID_ARG, "symbol size does not match the instance");
    CCA_REQUIRE(I->dialect == CC_DIALECT_GPU || symSize == 2, CC_ERR_INVALID_ARG, "the CPU dialect is Symbol16 only");
    CCA_REQUIRE(nbs >= 1 && bsWords && bsBits && bsPos, CC_ERR_INVALID_ARG, "no bitstream");
    if (ns == 0) return;
    CCA_REQUIRE(outs != nullptr, CC_ERR_INVALID_ARG, "null output array");
    const uint32_t cbs = I->cbs;
    const HeaderFormat fmt = header_format(I->dialect, I->log2);
    SectionTimer call(I, 1), stage(I, 2);
    call(rl ? "Decode RLHuff" : "Decode Huff");
    stage("Decode: Read Tables + Upload");

    struct Coded {
        DecTable t;
        const uint32_t* words;
        uint32_t nwords;
        const uint16_t* incs;
        uint32_t count;
    };
    const uint32_t nc = rl ? 2 * ns : ns;
    std::vector<Coded> coded(nc);
    std::vector<uint32_t> m(ns, n);
    auto read_coded = [&](BitReader& r, Coded& c, uint32_t count) {
        CCA_REQUIRE(read_table(fmt, r, c.t), CC_ERR_CORRUPT_STREAM, "malformed Huffman table header");
        const uint32_t nbits = r.aligned_u32();
        c.words = r.word_ptr();
        c.nwords = (nbits + 31) / 32;
        r.skip_bits(32ull * c.nwords);
        r.align32();
        c.incs = (const uint16_t*)r.word_ptr();
        r.skip_aligned(div_up(count, cbs), 2);
        c.count = count;
        CCA_REQUIRE(r.ok(), CC_ERR_CORRUPT_STREAM, "bitstream ends inside a coded stream");
        CCA_REQUIRE(count == 0 || !c.t.symbols.empty(), CC_ERR_CORRUPT_STREAM, "empty table for a non-empty stream");
    };
    for (uint32_t i = 0; i < ns; i++) {
        const uint32_t b = s2b ? s2b[i] : 0u;
        CCA_REQUIRE(b < nbs && bsWords[b], CC_ERR_INVALID_ARG, "bad bitstream index");
        BitReader r(bsWords[b], bsBits[b], bsPos[b]);
        if (rl) {
            m[i] = r.aligned_u32();
            CCA_REQUIRE(r.ok() && m[i] <= n, CC_ERR_CORRUPT_STREAM, "run-length pair count exceeds the stream length");
            read_coded(r, coded[2 * i], m[i]);
            read_coded(r, coded[2 * i + 1], m[i]);
        } else {
            read_coded(r, coded[i], n);
        }
        bsPos[b] = r.pos();
    }

    // staging layout
    Staging s{nullptr, nullptr};
    std::vector<size_t> o_tab(nc), o_sym(nc), o_off(nc), o_words(nc);
    uint32_t nwaves = 0;
    std::vector<uint32_t> wave_base(nc);
    for (uint32_t j = 0; j < nc; j++) {
        o_tab[j] = s.push<int32_t>(96);
        o_sym[j] = s.push<unsigned char>(coded[j].t.symbols.size() * symSize);
        o_off[j] = s.push<uint32_t>(div_up(coded[j].count, cbs));
        o_words[j] = s.push<uint32_t>(coded[j].nwords + 2);
        wave_base[j] = nwaves;
        nwaves += div_up(div_up(coded[j].count, cbs), 64);
    }
    size_t o_desc = s.push<HuffDecStream>(nc);
    size_t o_wmap = s.push<uint32_t>(nwaves);
    uint32_t ntiles = 0;
    std::vector<uint32_t> tile_base(ns);
    for (uint32_t i = 0; i < ns && rl; i++) {
        tile_base[i] = ntiles;
        ntiles += div_up(m[i], kRleTile);
    }
    size_t o_rdesc = s.push<RleDecStream>(rl ? ns : 0);
    size_t o_tmap = s.push<uint32_t>(ntiles);
    const size_t total = s.off;

    const int slot = I->dec_slot;
    I->dec_slot = (I->dec_slot + 1) % kDecodeRing;
    CCA_HIP(hipEventSynchronize(I->dec_event[slot]));  // previous upload from this slot done
    unsigned char* h = (unsigned char*)I->h_dec_stage[slot].get(total);
    unsigned char* d = (unsigned char*)I->d_dec_stage.get(total);
    const uint64_t cap = align_up(std::max<uint32_t>(n, 1), 64);
    unsigned char* dsym = rl ? (unsigned char*)I->d_dec_sym.get(cap * ns * symSize) : nullptr;
    unsigned char* dzc = rl ? (unsigned char*)I->d_dec_zc.get(cap * ns * symSize) : nullptr;

    HuffDecStream* hd = (HuffDecStream*)(h + o_desc);
    uint32_t* hwm = (uint32_t*)(h + o_wmap);
    for (uint32_t j = 0; j < nc; j++) {
        const Coded& c = coded[j];
        int32_t* tab = (int32_t*)(h + o_tab[j]);
        memcpy(tab, c.t.first, 128);
        memcpy(tab + 32, c.t.mn, 128);
        memcpy(tab + 64, c.t.mx, 128);
        if (symSize == 2) {
            uint16_t* sy = (uint16_t*)(h + o_sym[j]);
            for (size_t k = 0; k < c.t.symbols.size(); k++) sy[k] = (uint16_t)c.t.symbols[k];
        } else {
            memcpy(h + o_sym[j], c.t.symbols.data(), 4 * c.t.symbols.size());
        }
        uint32_t* off = (uint32_t*)(h + o_off[j]);
        const uint32_t noff = div_up(c.count, cbs);
        uint32_t acc = 0;
        for (uint32_t b = 0; b < noff; b++) {  // unpackInc16CPU, PackInc.cu:222-229
            acc += c.incs[b];
            off[b] = acc;
            CCA_REQUIRE(acc <= 32ull * c.nwords, CC_ERR_CORRUPT_STREAM, "coding block offset beyond the codewords");
        }
        uint32_t* wd = (uint32_t*)(h + o_words[j]);
        memcpy(wd, c.words, 4ull * c.nwords);
        wd[c.nwords] = 0;
        wd[c.nwords + 1] = 0;
        void* out;
        if (rl) out = (j % 2 == 0 ? dsym : dzc) + cap * (j / 2) * symSize;
        else out = outs[j];
        CCA_REQUIRE(out != nullptr || c.count == 0, CC_ERR_INVALID_ARG, "null output stream");
        hd[j] = HuffDecStream{(const uint32_t*)(d + o_words[j]), (const uint32_t*)(d + o_off[j]),
                              (const int32_t*)(d + o_tab[j]), d + o_sym[j], out, c.count,
                              (uint32_t)c.t.symbols.size(), wave_base[j], c.nwords};
        for (uint32_t w = 0; w < div_up(div_up(c.count, cbs), 64); w++) hwm[wave_base[j] + w] = j;
    }
    if (rl) {
        RleDecStream* rd = (RleDecStream*)(h + o_rdesc);
        uint32_t* tm = (uint32_t*)(h + o_tmap);
        for (uint32_t i = 0; i < ns; i++) {
            rd[i] = RleDecStream{dsym + cap * i * symSize, dzc + cap * i * symSize, outs[i], m[i], n, tile_base[i], 0};
            for (uint32_t t = 0; t < div_up(m[i], kRleTile); t++) tm[tile_base[i] + t] = i;
        }
        for (uint32_t i = 0; i < ns; i++) CCA_REQUIRE(outs[i] != nullptr || n == 0, CC_ERR_INVALID_ARG, "null output stream");
    }
    CCA_HIP(hipMemcpyAsync(d, h, total, hipMemcpyHostToDevice, st));
    CCA_HIP(hipEventRecord(I->dec_event[slot], st));
    uint64_t total_syms = 0, total_bits = 0, total_pairs = 0;
    for (uint32_t j = 0; j < nc; j++) {
        total_syms += coded[j].count;
        total_bits += 32ull * coded[j].nwords;
    }
    for (uint32_t i = 0; i < ns && rl; i++) total_pairs += m[i];
    stage("Decode: Huffman Decode");
    huffman_decode((const HuffDecStream*)(d + o_desc), nc, (const uint32_t*)(d + o_wmap), nwaves, cbs, symSize,
                   total_syms, total_bits, st);
    if (rl) {
        stage("Decode: Run Length Decode");
        uint32_t* sums = I->d_dec_sums.as<uint32_t>(std::max<uint32_t>(ntiles, 1));
        rle_decode((const RleDecStream*)(d + o_rdesc), ns, sums, ntiles, (const uint32_t*)(d + o_tmap), n, symSize,
                   total_pairs, (uint64_t)n * ns, st);
    }
}

}  // namespace cca
