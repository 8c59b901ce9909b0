// C-ABI over the reference's own CPU entropy coder (src/CPU of m0bl0/cudaCompress), compiled
// from the reference sources in place by oracle/Makefile into oracle/_ref/.
// TEST INFRASTRUCTURE ONLY: used to pin the oracle and to write golden vectors.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <cudaCompress/BitStream.h>
#include "CPU/EncodeCPU.h"
#include "CPU/HuffmanTableCPU.h"
#include "CPU/RunLengthCPU.h"

using namespace cudaCompress;

extern "C" {

// encodeRLHuffCPU / encodeHuffCPU, src/CPU/EncodeCPU.cpp:98-197, :275-320.  All streams go into one
// BitStream back to back (every entry of ppBitStreams points at the same stream).
int ref_encode(int rl, int cbs, const uint16_t* const* streams, uint32_t nstreams, uint32_t count,
               uint32_t** out_words, uint64_t* out_len, uint64_t* out_bits) {
    std::vector<uint> storage;
    BitStream bs(&storage);
    std::vector<std::vector<Symbol16>> data(nstreams);
    std::vector<std::vector<Symbol16>*> pdata(nstreams);
    std::vector<BitStream*> pbs(nstreams, &bs);
    for (uint32_t s = 0; s < nstreams; s++) {
        data[s].assign(streams[s], streams[s] + count);
        pdata[s] = &data[s];
    }
    bool ok = rl ? encodeRLHuffCPU(pbs.data(), pdata.data(), nstreams, (uint)cbs, true)
                 : encodeHuffCPU(pbs.data(), pdata.data(), nstreams, (uint)cbs);
    if (!ok) return 1;
    size_t nw = bs.getRawSizeUInts();
    uint32_t* o = (uint32_t*)malloc(sizeof(uint32_t) * (nw ? nw : 1));
    if (nw) memcpy(o, bs.getRaw(), nw * 4);
    *out_words = o;
    *out_len = nw;
    *out_bits = bs.getBitPosition();
    return 0;
}

// decodeRLHuffCPU / decodeHuffCPU, src/CPU/EncodeCPU.cpp:200-256, :322-354.
int ref_decode(int rl, int cbs, const uint32_t* words, uint64_t nbits, uint16_t* const* out, uint32_t nstreams,
               uint32_t count, uint64_t* bitpos_out) {
    BitStreamReadOnly bs(words, (uint)nbits);
    std::vector<std::vector<Symbol16>> data(nstreams);
    std::vector<std::vector<Symbol16>*> pdata(nstreams);
    std::vector<BitStreamReadOnly*> pbs(nstreams, &bs);
    for (uint32_t s = 0; s < nstreams; s++) pdata[s] = &data[s];
    bool ok = rl ? decodeRLHuffCPU(pbs.data(), pdata.data(), count, nstreams, (uint)cbs, true)
                 : decodeHuffCPU(pbs.data(), pdata.data(), count, nstreams, (uint)cbs);
    if (!ok) return 1;
    for (uint32_t s = 0; s < nstreams; s++) {
        if (data[s].size() != count) return 2;
        memcpy(out[s], data[s].data(), count * sizeof(uint16_t));
    }
    *bitpos_out = bs.getBitPosition();
    return 0;
}

// HuffmanTableCPU::design, src/CPU/HuffmanTableCPU.cpp:22-105.
int ref_design(const uint16_t* syms, uint32_t n, uint32_t* num_lengths, uint32_t* count_per_length,
               uint32_t* symbols, uint32_t* nsyms) {
    std::vector<Symbol16> v(syms, syms + n);
    HuffmanTableCPU t;
    t.design(v);
    *num_lengths = (uint32_t)t.m_codewordCountPerLength.size();
    for (size_t i = 0; i < t.m_codewordCountPerLength.size() && i < 32; i++) count_per_length[i] = t.m_codewordCountPerLength[i];
    for (size_t i = 0; i < t.m_symbols.size(); i++) symbols[i] = t.m_symbols[i];
    *nsyms = (uint32_t)t.m_symbols.size();
    return 0;
}

// runLengthZeroesEncodeCPU, src/CPU/RunLengthCPU.cpp:8-32.
int ref_rle(const uint16_t* syms, uint32_t n, uint16_t* out_syms, uint16_t* out_zc, uint32_t* out_m) {
    std::vector<Symbol16> v(syms, syms + n), s;
    std::vector<ushort> z;
    runLengthZeroesEncodeCPU(s, z, v, ZERO_COUNT_MAX);
    for (size_t i = 0; i < s.size(); i++) { out_syms[i] = s[i]; out_zc[i] = z[i]; }
    *out_m = (uint32_t)s.size();
    return 0;
}

void ref_free(void* p) { free(p); }

}  // extern "C"
