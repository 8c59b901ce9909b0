// Device-resident entropy decoder for gfx950: parse -> tables -> blocks -> rle.
// Reads the layouts of src/cudaCompress/Encode.cpp:474-646 (decodeRLHuff) and :793-917
// (decodeHuff) straight from a bitstream in device memory: table headers as
// src/cudaCompress/HuffmanTable.cu:117-151 reads them (including the 4-bit symbolBits field whose
// stored 0 means 16 when more than one symbol follows), decode tables as :205-266, block offsets
// as unpackInc16 (src/cudaCompress/PackInc.cu:222-229), canonical decoding per coding block as
// src/cudaCompress/HuffmanKernels.cui:197-371, and the zero-run expansion of
// src/cudaCompress/RunLengthKernels.cui:57-76.  Every read is bounded by the call's word count,
// so a corrupt stream yields an error status, never an out-of-bounds access.
#include <type_traits>

#include "../coder.hpp"
#include "../profile.hpp"
#include "huff_design.hpp"
#include "wave.hpp"

namespace cca {
namespace {

constexpr uint32_t kLut = kDecLutBits;

// n <= 32 bits at bit position pos of w[0, nwords); ok cleared when out of range
__device__ __forceinline__ uint32_t rd_bits(const uint32_t* w, uint64_t nwords, uint64_t pos, uint32_t n, bool& ok) {
    if (n == 0) return 0;
    if (((pos + n - 1) >> 5) >= nwords) {
        ok = false;
        return 0;
    }
    const uint64_t wi = pos >> 5;
    const uint32_t o = (uint32_t)(pos & 31);
    uint64_t win = (uint64_t)w[wi] << 32;
    if (o + n > 32) win |= w[wi + 1];
    return (uint32_t)((win << o) >> (64 - n));
}

__device__ uint64_t arena_take(CoderState* st, uint64_t words, uint64_t cap, bool& ok) {
    words = (words + 15) & ~uint64_t(15);
    const uint64_t off = atomicAdd(&st->scratch_used, (unsigned long long)words);
    if (off + words > cap) {
        ok = false;
        atomicCAS(&st->err, 0u, (uint32_t)kDecScratch);
        return 0;
    }
    return off;
}

__global__ void k_dec_init(CoderState* st) {
    if (threadIdx.x == 0) {
        st->scratch_used = 0;
        st->out_words = 0;
        st->err = 0;
        st->pad = 0;
    }
}

// One wave per call.  The sections of a call are walked in order (their sizes are in-band); each
// section's head is read as one 64-word window by the wave's lanes (lane i: word i of the window),
// and the fields come out of the window with lane reads, so a section costs one coalesced load (two
// when its bit-count word lies past the window) instead of a chain of dependent word reads.  The
// call's arena ranges are taken after the walk with one atomic per 64 sequences: lane k holds the
// k-th sequence's range sizes, a wave scan places them (block offsets, lookup and symbol table,
// and the decoded values of an RLE sequence).  index_mode: a resident decode index (tables and
// block offsets only): RLE sequences get no decode buffer, a bound decode supplies one per request
// (k_dec_bind).
struct ParseWin {
    uint64_t base;  // stream word of lane 0
    uint32_t v;     // this lane's word (0 past the stream)
};
__device__ __forceinline__ void parse_load(ParseWin& W, const DecCall& C, uint64_t base) {
    const uint64_t i = base + (uint64_t)threadIdx.x;
    W.base = base;
    W.v = i < C.nwords ? C.words[i] : 0u;
}
// stream word i (i < C.nwords): from the window, else from memory
__device__ __forceinline__ uint32_t parse_word(const ParseWin& W, const DecCall& C, uint64_t i) {
    const uint64_t d = i - W.base;
    if (i >= W.base && d < 64) return (uint32_t)__builtin_amdgcn_readlane((int)W.v, (int)d);
    return C.words[i];
}
// rd_bits over the window
__device__ __forceinline__ uint32_t parse_bits(const ParseWin& W, const DecCall& C, uint64_t pos, uint32_t n,
                                               bool& ok) {
    if (n == 0) return 0;
    if (((pos + n - 1) >> 5) >= C.nwords) {
        ok = false;
        return 0;
    }
    const uint64_t wi = pos >> 5;
    const uint32_t o = (uint32_t)(pos & 31);
    uint64_t win = (uint64_t)parse_word(W, C, wi) << 32;
    if (o + n > 32) win |= parse_word(W, C, wi + 1);
    return (uint32_t)((win << o) >> (64 - n));
}
__device__ __forceinline__ uint64_t arena_words16(uint64_t words) { return (words + 15) & ~uint64_t(15); }

__global__ __launch_bounds__(64) void k_dec_parse(const DecCall* calls, uint32_t ncalls, const DecItem* items,
                                                  DecSeq* seqs, CoderParams p, int symSize, uint32_t* scratch,
                                                  uint64_t cap, CoderState* st, DecCallOut* outc, int index_mode) {
    const uint32_t c = blockIdx.x;
    if (c >= ncalls) return;
    const int lane = threadIdx.x;
    const DecCall C = calls[c];
    const uint32_t cb = p.count_bits, sbf = p.symbits_bits;
    uint64_t w = 0;
    uint32_t err = 0;
    ParseWin W;
    // the batch of sequences waiting for their arena ranges: lane k holds the k-th one's sizes
    uint32_t kb = 0, my_seq = 0;
    uint64_t my_boff = 0, my_lut = 0, my_out = 0;
    auto place = [&]() {
        const uint64_t mine = (uint32_t)lane < kb ? my_boff + my_lut + my_out : 0ull;
        const uint64_t incl = wave_incl_scan_u64(mine);
        const uint64_t total = dpp::last_lane64(incl);
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd(&st->scratch_used, (unsigned long long)total);
        base = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(base >> 32), 0) << 32) |
               (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, 0);
        if (base + total > cap) {
            err = kDecScratch;
            if (lane == 0) atomicCAS(&st->err, 0u, (uint32_t)kDecScratch);
        } else if ((uint32_t)lane < kb) {
            const uint64_t o = base + incl - mine;
            DecSeq& s = seqs[my_seq];
            s.boff_off = o;
            s.lut_off = o + my_boff;
            if (my_out) s.out = scratch + o + my_boff + my_lut;
        }
        kb = 0;
    };
    for (uint32_t i = C.first_item; i < C.first_item + C.nitems && !err; i++) {
        const DecItem it = items[i];
        uint32_t count = it.n;
        if (it.rl) {
            if (w >= C.nwords) { err = kDecCorrupt; break; }
            parse_load(W, C, w);
            count = parse_word(W, C, w);
            w++;
            if (count > it.n) { err = kDecCorrupt; break; }
        }
        for (uint32_t q = 0; q < (it.rl ? 2u : 1u); q++) {
            bool ok = true;
            const uint64_t hb = w * 32;
            if (q == 1 || !it.rl) parse_load(W, C, w);  // the first sequence of an RLE item: loaded at its count word
            const uint32_t nlen = parse_bits(W, C, hb, 5, ok);
            const uint32_t nsym = parse_bits(W, C, hb + 5 + (uint64_t)cb * nlen, cb, ok);
            uint32_t sb = parse_bits(W, C, hb + 5 + (uint64_t)cb * (nlen + 1), sbf, ok);
            if (p.dialect == CC_DIALECT_GPU && sbf == 4 && sb == 0 && nsym > 1) sb = 16;
            const uint64_t sym_bit = hb + 5 + (uint64_t)cb * (nlen + 1) + sbf;
            const uint64_t tb = 5 + (uint64_t)cb * (nlen + 1) + sbf + (uint64_t)sb * nsym;
            w = (hb + tb + 31) / 32;
            if (!ok || sb > 32 || w >= C.nwords || (count > 0 && nsym == 0) || nsym > count) { err = kDecCorrupt; break; }
            const uint32_t bits = parse_word(W, C, w);
            w++;
            const uint64_t words_off = w;
            const uint64_t nwc = ((uint64_t)bits + 31) / 32;
            w += nwc;
            const uint64_t incs_off = w;
            const uint64_t nblk = ((uint64_t)count + p.cbs - 1) / p.cbs;
            w += (nblk + 1) / 2;
            if (w > C.nwords) { err = kDecCorrupt; break; }
            const uint32_t sq = it.seq + q;
            if (lane == 0) {
                DecSeq& s = seqs[sq];
                s.base = C.words;
                s.nwords = C.nwords;
                s.hdr_bit = hb;
                s.sym_bit = sym_bit;
                s.words_off = words_off;
                s.incs_off = incs_off;
                if (!it.rl || index_mode) s.out = it.out;  // else the arena buffer, placed below
                s.nlen = nlen;
                s.nsym = nsym;
                s.sb = sb;
                s.count = count;
                s.nwords_cw = (uint32_t)nwc;
                s.err = 0;
                s.item = i;
            }
            if ((uint32_t)lane == kb) {
                my_seq = sq;
                my_boff = arena_words16(nblk ? nblk : 1);
                my_lut = arena_words16((1u << kLut) + (uint64_t)nsym + 1);
                my_out = it.rl && !index_mode ? arena_words16(((uint64_t)count * symSize + 3) / 4 + 1) : 0ull;
            }
            if (++kb == 64) {
                place();
                if (err) break;
            }
        }
    }
    if (!err && kb) place();
    if (err) {  // mark every sequence of the call failed
        for (uint32_t k = C.first_item + lane; k < C.first_item + C.nitems; k += 64) {
            const DecItem it = items[k];
            seqs[it.seq].err = err;
            seqs[it.seq].item = k;
            if (it.rl) {
                seqs[it.seq + 1].err = err;
                seqs[it.seq + 1].item = k;
            }
        }
        if (lane == 0) atomicCAS(&st->err, 0u, err);
    }
    if (lane == 0) {
        outc[c].words = w;
        outc[c].err = err;
        outc[c].pad = 0;
    }
}

// One workgroup per coded sequence: first/min/max codeword per length (closed form of the
// canonical assignment) and absolute coding-block offsets.
__global__ __launch_bounds__(256) void k_dec_tables(DecSeq* seqs, CoderParams p, uint32_t* scratch,
                                                    CoderState* st) {
    constexpr uint32_t kSymPer = (1u << kLut) / 256;  // symbols per thread read up front
    __shared__ uint32_t cpl[32];
    __shared__ uint32_t sh[4];
    __shared__ uint32_t bad;
    __shared__ int tab_sh[96];
    __shared__ uint32_t sym_lo[1 << kLut];
    DecSeq& s = seqs[blockIdx.x];
    if (s.err) return;
    // every independent header read is issued before the first barrier: the code-length counts,
    // the first 2^kLut symbols and the first 1024 block-offset increments
    bool ok = true;
    const uint32_t cplv = threadIdx.x < s.nlen
                              ? rd_bits(s.base, s.nwords, s.hdr_bit + 5 + (uint64_t)p.count_bits * threadIdx.x,
                                        p.count_bits, ok)
                              : 0u;
    uint32_t symv[kSymPer];
#pragma unroll
    for (uint32_t j = 0; j < kSymPer; j++) {
        const uint32_t k = threadIdx.x + 256 * j;
        symv[j] = k < s.nsym ? rd_bits(s.base, s.nwords, s.sym_bit + (uint64_t)s.sb * k, s.sb, ok) : 0u;
    }
    const uint32_t nblk = (s.count + p.cbs - 1) / p.cbs;
    const uint32_t* inc = s.base + s.incs_off;
    uint32_t incv[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t b = threadIdx.x * 4 + k;
        incv[k] = b < nblk ? (inc[b >> 1] >> (16 * (b & 1))) & 0xFFFFu : 0u;
    }
    if (threadIdx.x == 0) bad = 0;
    if (threadIdx.x < 32) cpl[threadIdx.x] = cplv;
    __syncthreads();
    if (!ok) atomicOr(&bad, 1u);
    if (threadIdx.x == 0) {
        uint32_t fc[32], fi[32];
        design::first_codes(cpl, (int)s.nlen, fc, fi);
        uint64_t total = 0;
        for (uint32_t l = 0; l < 32; l++) {
            const bool used = l < s.nlen && cpl[l] > 0;
            tab_sh[l] = used ? (int32_t)fi[l] : -1;
            tab_sh[32 + l] = used ? (int32_t)fc[l] : -1;
            tab_sh[64 + l] = used ? (int32_t)(fc[l] + cpl[l] - 1) : -1;
            if (l < s.nlen) total += cpl[l];
        }
        if (s.nlen > 0 && total != s.nsym) bad = 1;
    }
    // symbol table (codeword-index order); canonical codes of length <= kLut belong to the first
    // 2^kLut symbols, kept in LDS for the lookup table
    uint32_t* gsyms = scratch + s.lut_off + (1u << kLut);
    uint32_t* glut = scratch + s.lut_off;
#pragma unroll
    for (uint32_t j = 0; j < kSymPer; j++) {
        const uint32_t k = threadIdx.x + 256 * j;
        if (k < s.nsym) {
            gsyms[k] = symv[j];
            sym_lo[k] = symv[j];
        }
    }
    for (uint32_t k = threadIdx.x + (1u << kLut); k < s.nsym; k += 256) {
        bool ok2 = true;
        gsyms[k] = rd_bits(s.base, s.nwords, s.sym_bit + (uint64_t)s.sb * k, s.sb, ok2);
        if (!ok2) atomicOr(&bad, 1u);
    }
    __syncthreads();
    if (threadIdx.x < 96) s.tab[threadIdx.x] = tab_sh[threadIdx.x];
    // the table is indexed by the next lb = min(longest code, kLut) bits: 2^lb entries
    const uint32_t lb = s.nlen < kLut ? s.nlen : kLut;
    for (uint32_t e = threadIdx.x; e < (1u << lb); e += 256) {
        uint32_t entry = 0;
        for (int l = 1; l <= (int)lb; l++) {
            const int cw = (int)(e >> (lb - l));
            if (cw <= tab_sh[64 + l - 1]) {
                const int idx = tab_sh[l - 1] + cw - tab_sh[32 + l - 1];
                const uint32_t sym = (idx >= 0 && (uint32_t)idx < s.nsym && idx < (1 << kLut)) ? sym_lo[idx] : 0u;
                entry = ((uint32_t)l << 24) | (sym & 0xFFFFFFu);
                break;
            }
        }
        glut[e] = entry;
    }
    // absolute block offsets: running sum of the u16 increments
    uint32_t* boff = scratch + s.boff_off;
    const uint64_t limit = 32ull * s.nwords_cw;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nblk; base += 256 * 4) {
        uint32_t v[4], sum = 0;
        for (int k = 0; k < 4; k++) {
            const uint32_t b = base + threadIdx.x * 4 + k;
            v[k] = base == 0 ? incv[k] : (b < nblk ? (inc[b >> 1] >> (16 * (b & 1))) & 0xFFFFu : 0u);
            sum += v[k];
        }
        uint32_t tot;
        uint32_t ex = block_excl_scan<256>(sum, sh, tot);
        uint32_t acc = carry + ex;
        for (int k = 0; k < 4; k++) {
            const uint32_t b = base + threadIdx.x * 4 + k;
            acc += v[k];
            if (b < nblk) {
                boff[b] = acc;
                if (acc > limit) atomicOr(&bad, 1u);
            }
        }
        carry += tot;
    }
    __syncthreads();
    if (threadIdx.x == 0 && bad) {
        s.err = kDecCorrupt;
        atomicCAS(&st->err, 0u, (uint32_t)kDecCorrupt);
    }
}

// Block decoder: one thread per coding block, W waves per workgroup (W = 4: 256 blocks of one
// long sequence sharing one LDS copy of its lookup table, which then leaves L2 once per 256
// blocks; W = 1: a short sequence's blocks, so that a sequence of a few blocks does not hold a
// four-wave workgroup's LDS).  A lane decodes its block R symbols per round (8 in both kinds:
// less LDS per wave, so more waves hide the latency of the serial lookups; a 64-symbol block is
// held in registers across its rounds and stored at once) and stores them itself (16 bytes per 8
// 16-bit symbols); the waves of a workgroup meet only at the table load.
// The kernel is bound by instruction issue and latency, so a code costs a handful of vector
// instructions: the bit window is two words {hi, lo} read with one funnel shift (`alignbit`, the
// next code starting q bits above lo's lowest bit); codewords reach a lane through an LDS copy of
// the R + 4 words from the 16-byte boundary at or below its next word, loaded at the start of
// every round with 16-byte loads and one wait (a round takes at most one word per code), so
// refilling the bit window is a branch-free select; the round is unrolled, its symbols collect in
// registers at fixed positions; and the zero-count sums run only in the instance for the zero
// counts of RLE items.
// Symbols per round R: 16 for the long sequences' workgroups, 8 for the short ones' waves (which
// then hold less LDS: the copy is R + 4 words, >= R + 3 since it starts up to 3 words early).
template <uint32_t R>
constexpr uint32_t win_words() { return R + 4; }
// row of a lane's run-path symbols in LDS: an odd number of words, so lanes writing the same
// position use distinct banks
template <class S, uint32_t R>
constexpr uint32_t run_row_words() { return sizeof(S) == 2 ? R / 2 + 1 : R + 1; }

template <class S, int W, uint32_t R>
// (at most 72 vector registers: seven waves per SIMD, as the long workgroups' LDS allows)
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(7))) void k_dec_blocks(const DecSeq* seqs, const uint32_t* wave_seq,
                                                       const uint32_t* wave_first, CoderParams p,
                                                       const uint32_t* scratch, const DecItem* items, RleChunks rc) {
    constexpr uint32_t kDecRound = R;
    constexpr uint32_t kWinWords = win_words<R>();
    static_assert(R % 8 == 0 && kWinWords % 4 == 0, "16-byte stores and copies");
    constexpr uint32_t RW = run_row_words<S, R>();
    constexpr uint32_t NW = kDecRound * sizeof(S) / 4;  // words of a round's symbols
    __shared__ uint32_t lut[1 << kLut];
    __shared__ int tab[96];
    __shared__ S run_all[W][64 * RW * 4 / sizeof(S)];
    __shared__ __align__(16) uint32_t win_all[W][64 * kWinWords];
    // the workgroups of one sequence (consecutive in the list) on one XCD: its tables leave HBM once
    const uint32_t wv = xcd_order(blockIdx.x, gridDim.x);
    const uint32_t sq = wave_seq[wv];
    const DecSeq& s = seqs[sq];
    if (s.err) return;
    const uint32_t cbs = p.cbs;
    const uint32_t count = s.count;
    const uint32_t nblk = (count + cbs - 1) / cbs;
    const uint32_t fb0 = wave_first[wv];
    if (fb0 >= nblk) return;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    S* out = static_cast<S*>(s.out);
    const bool multi = s.nlen > 0;
    const uint32_t nsym = s.nsym;
    const uint32_t* gsyms = scratch + s.lut_off + (1u << kLut);
    auto symbol_at = [&](int idx) -> uint32_t { return (idx >= 0 && (uint32_t)idx < nsym) ? gsyms[idx] : 0u; };
    // lookup table: 2^lb entries, indexed by the next lb bits (lb = min(longest code, kLut))
    const uint32_t lb = s.nlen < kLut ? s.nlen : kLut;
    if (multi) {
        // all of the thread's table loads in flight at once, then the LDS stores
        const uint32_t* glut = scratch + s.lut_off;
        constexpr int kPer = (1 << kLut) / (64 * W);
        const uint32_t nl = 1u << lb;
        uint32_t lv[kPer];
#pragma unroll
        for (int k = 0; k < kPer; k++) lv[k] = (uint32_t)(t + 64 * W * k) < nl ? glut[t + 64 * W * k] : 0u;
        int tv[(96 + 64 * W - 1) / (64 * W)];
#pragma unroll
        for (int k = 0; k < (96 + 64 * W - 1) / (64 * W); k++) tv[k] = t + 64 * W * k < 96 ? s.tab[t + 64 * W * k] : 0;
#pragma unroll
        for (int k = 0; k < kPer; k++)
            if ((uint32_t)(t + 64 * W * k) < nl) lut[t + 64 * W * k] = lv[k];
#pragma unroll
        for (int k = 0; k < (96 + 64 * W - 1) / (64 * W); k++)
            if (t + 64 * W * k < 96) tab[t + 64 * W * k] = tv[k];
    }
    __syncthreads();
    const uint32_t fb = fb0 + 64u * (uint32_t)wave;
    if (fb >= nblk) return;  // no block for this wave
    S* run_row = &run_all[wave][lane * RW * 4 / sizeof(S)];
    // the run row and the codeword copy are also accessed as words / 16-byte vectors
    typedef uint32_t __attribute__((may_alias)) u32a;
    typedef uint32_t u32x4a __attribute__((ext_vector_type(4), may_alias));
    uint32_t* win = &win_all[wave][lane * kWinWords];
    const uint32_t single = nsym >= 1 ? symbol_at(0) : 0u;
    const uint32_t b = fb + lane;
    const bool valid = b < nblk;
    // the zero counts of an RLE item also give the zero-run expansion its chunk spans: the sum of
    // (count + 1) over each chunk of kRleChunk pairs (kRleChunk / cbs whole blocks, a power of two
    // dividing the wave's 64 blocks)
    const DecItem zit = rc.nchunks ? items[s.item] : DecItem{};
    const bool zsum = rc.nchunks != 0 && zit.rl && zit.seq + 1 == sq;
    const uint32_t* words = s.base + s.words_off;
    const uint64_t avail = s.nwords - s.words_off;  // readable words from the codeword start
    auto word = [&](uint32_t i) -> uint32_t { return i < avail ? words[i] : 0u; };
    // window {hi, lo} (the next code: alignbit(hi, lo, q), 32 + q bits unread, q in [0, 31] after a
    // refill), word nxt next; the LDS copy holds words [w0, w0 + kWinWords) of the block's stream
    uint32_t hi = 0, lo = 0, nxt = 0;
    int q = 0, w0 = 0;
    if (valid && multi) {
        const uint32_t pos = b == 0 ? 0u : scratch[s.boff_off + b];
        const uint32_t wi = pos >> 5, b0 = pos & 31u;
        if (b0 == 0) {
            lo = word(wi);
            nxt = wi + 1;
        } else {
            hi = word(wi);
            lo = word(wi + 1);
            nxt = wi + 2;
            q = 32 - (int)b0;
        }
    }
    // words from the 16-byte boundary at or below word nxt into the LDS copy.  The call's buffer
    // holds words [-words_off, avail); a start below it stays in the 16-byte block of the buffer's
    // first word.  Where the copy would reach past the buffer's end, words are read one by one.
    const uint32_t* cend = s.base + s.nwords;
    auto load_window = [&]() {
        const uint32_t* gp = reinterpret_cast<const uint32_t*>(reinterpret_cast<uintptr_t>(words + nxt) & ~uintptr_t(15));
        w0 = (int)nxt - (int)(words + nxt - gp);
        const bool fast = !(valid && multi) || gp + kWinWords <= cend;
        if (__all(fast)) {
            if (valid && multi) {
                // a global-space pointer: the loads cannot alias the LDS stores, so all of them
                // are issued before the first store waits
                using G4 = const __attribute__((address_space(1))) u32x4a*;
                const G4 g4 = (G4)(const void*)gp;
                u32x4a v[kWinWords / 4];
#pragma unroll
                for (uint32_t k = 0; k < kWinWords / 4; k++) v[k] = g4[k];
#pragma unroll
                for (uint32_t k = 0; k < kWinWords / 4; k++) reinterpret_cast<u32x4a*>(win)[k] = v[k];
            }
        } else if (valid && multi) {
#pragma unroll
            for (uint32_t k = 0; k < kWinWords; k++) {
                const int i = w0 + (int)k;
                win[k] = (i >= 0 && (uint64_t)i < avail) ? words[i] : 0u;
            }
        }
    };
    auto refill = [&]() {
        const uint32_t w = win[nxt - (uint32_t)w0];
        const bool need = q < 0;
        hi = need ? lo : hi;
        lo = need ? w : lo;
        nxt += need ? 1u : 0u;
        q &= 31;
    };
    const uint32_t sh = 32 - lb;
    const uint32_t mlen = s.nlen;
    // a code longer than the lookup table: the per-length first/min codes; (len << 24) | symbol
    auto slow = [&](uint32_t pk) -> uint32_t {
        uint32_t len = kLut + 1;
        while (len < mlen && (int)(pk >> (32 - len)) > tab[64 + len - 1]) len++;
        const int cw = (int)(pk >> (32 - len));
        return (len << 24) | symbol_at(tab[len - 1] + cw - tab[32 + len - 1]);
    };
    // one codeword of length 1 (then it is "0" and every other codeword starts with a 1): a run of
    // k zero bits is k copies of its symbol, taken at once (the zero-count sequences, where a zero
    // count is nearly every pair, decode a few steps per round instead of kDecRound)
    const bool run0 = multi && s.tab[64] == 0 && s.tab[32] == 0;
    const uint32_t sym0 = run0 ? symbol_at(s.tab[0]) : 0u;
    const uint32_t bend_l = valid ? min((b + 1) * cbs, count) : 0u;  // this lane's block end
    // a lane's round leaves as 16-byte stores when the output is 16-byte aligned (round starts are
    // then aligned: cbs and kDecRound are multiples of 8), symbol by symbol otherwise and at the
    // sequence's end
    const bool wide = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
    auto store_round = [&](uint32_t p0, uint32_t vlen, const uint32_t (&v)[NW]) {
        if (wide && vlen == kDecRound) {
#pragma unroll
            for (uint32_t k = 0; k < NW; k += 4)
                *reinterpret_cast<uint4*>(out + p0 + k * 4 / sizeof(S)) = make_uint4(v[k], v[k + 1], v[k + 2], v[k + 3]);
        } else {
#pragma unroll
            for (uint32_t i = 0; i < kDecRound; i++)
                if (i < vlen) out[p0 + i] = sizeof(S) == 2 ? (S)(v[i >> 1] >> (16 * (i & 1))) : (S)v[i];
        }
    };
    // symbol i of a round into the round's words (fixed positions: the round is unrolled)
    auto put = [](uint32_t (&v)[NW], uint32_t i, uint32_t sym) {
        if (sizeof(S) == 2) v[i >> 1] = (i & 1) ? (v[i >> 1] | (sym << 16)) : (sym & 0xFFFFu);
        else v[i] = sym;
    };
    // one round: the symbols of positions [r, r + kDecRound) of the lane's block into v; every
    // lane steps through the whole round, steps past its block's end are not stored
    auto one_round = [&](uint32_t r, uint32_t (&v)[NW], uint32_t& zs, auto zflag) {
        constexpr bool Z = decltype(zflag)::value;
        const uint32_t p0 = b * cbs + r;
        const uint32_t vlen = bend_l > p0 ? min(bend_l - p0, kDecRound) : 0u;  // symbols of the block here
        if (!multi) {
#pragma unroll
            for (uint32_t i = 0; i < kDecRound; i++) put(v, i, single);
            if constexpr (Z) zs += (single + 1) * vlen;
        } else {
            load_window();
            if (run0) {
                // positions taken at a lane's own pace: its LDS row, prefilled with the run symbol
                const uint32_t fill = sizeof(S) == 2 ? (sym0 & 0xFFFFu) * 0x10001u : sym0;
#pragma unroll
                for (uint32_t k = 0; k < NW; k++) reinterpret_cast<u32a*>(run_row)[k] = fill;
                uint32_t i = 0;
                while (i < kDecRound) {
                    refill();
                    const uint32_t pk = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)q);
                    const bool run = (pk >> 31) == 0u;
                    uint32_t e = lut[pk >> sh];
                    if (!run && e == 0u) e = slow(pk);
                    const uint32_t k = min(pk ? (uint32_t)__clz(pk) : 32u, kDecRound - i);
                    const uint32_t sym = e & 0xFFFFFFu;
                    if (!run) run_row[i] = (S)sym;
                    if constexpr (Z)
                        zs += run ? (sym0 + 1) * (min(i + k, vlen) - min(i, vlen)) : (i < vlen ? sym + 1 : 0u);
                    q -= (int)(run ? k : (e >> 24));
                    i += run ? k : 1u;
                }
#pragma unroll
                for (uint32_t k = 0; k < NW; k++) v[k] = reinterpret_cast<const u32a*>(run_row)[k];
            } else {
#pragma unroll
                for (uint32_t i = 0; i < kDecRound; i++) {
                    refill();
                    const uint32_t pk = __builtin_amdgcn_alignbit(hi, lo, (uint32_t)q);
                    uint32_t e = lut[pk >> sh];
                    if (e == 0u) e = slow(pk);
                    q -= (int)(e >> 24);
                    const uint32_t sym = e & 0xFFFFFFu;
                    put(v, i, sym);
                    if constexpr (Z) zs += i < vlen ? sym + 1 : 0u;
                }
            }
        }
    };
    auto round_len = [&](uint32_t r) {
        const uint32_t p0 = b * cbs + r;
        return bend_l > p0 ? min(bend_l - p0, kDecRound) : 0u;
    };
    // A 64-symbol coding block is held in registers for all its rounds and leaves as back-to-back
    // stores, so that its output lines are whole when they leave L2 (a round at a time, a lane's
    // partial lines were written out 1.6 times over); other sizes store round by round.
    constexpr uint32_t kHeld = 64 / kDecRound;
    auto rounds = [&](auto zflag) -> uint32_t {
        uint32_t zs = 0;
        if (cbs == kHeld * kDecRound) {
            uint32_t v[kHeld][NW];
#pragma unroll
            for (uint32_t h = 0; h < kHeld; h++) one_round(h * kDecRound, v[h], zs, zflag);
#pragma unroll
            for (uint32_t h = 0; h < kHeld; h++) store_round(b * cbs + h * kDecRound, round_len(h * kDecRound), v[h]);
        } else {
            for (uint32_t r = 0; r < cbs; r += kDecRound) {
                uint32_t v[NW];
                one_round(r, v, zs, zflag);
                store_round(b * cbs + r, round_len(r), v);
            }
        }
        return zs;
    };
    if (zsum) {
        uint32_t zs = rounds(std::true_type{});
        const uint32_t bpc = kRleChunk / cbs;
        for (uint32_t o = 1; o < bpc; o <<= 1) zs += (uint32_t)__shfl_xor((int)zs, (int)o);
        const uint32_t c = b / bpc;
        if ((lane & (bpc - 1)) == 0 && c < (zit.n + kRleChunk - 1) / kRleChunk) rc.sums[rc.item_first[s.item] + c] = zs;
    } else {
        (void)rounds(std::false_type{});
    }
}

// Eight consecutive symbols from a 16-byte aligned address.
__device__ __forceinline__ void load8_aligned(const uint16_t* p, uint32_t (&o)[8]) {
    const uint4 w = *reinterpret_cast<const uint4*>(p);
    const uint32_t a[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        o[2 * k] = a[k] & 0xFFFFu;
        o[2 * k + 1] = a[k] >> 16;
    }
}
__device__ __forceinline__ void load8_aligned(const uint32_t* p, uint32_t (&o)[8]) {
    const uint4 w0 = reinterpret_cast<const uint4*>(p)[0], w1 = reinterpret_cast<const uint4*>(p)[1];
    o[0] = w0.x; o[1] = w0.y; o[2] = w0.z; o[3] = w0.w;
    o[4] = w1.x; o[5] = w1.y; o[6] = w1.z; o[7] = w1.w;
}

template <class S>
__device__ __forceinline__ void clear_symbols(S* out, uint32_t n) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(out);
    constexpr uint32_t per = 16 / sizeof(S);
    uint32_t head = (uint32_t)(((16 - (a & 15)) & 15) / sizeof(S));
    if ((a % sizeof(S)) != 0 || head > n) head = n;
    for (uint32_t i = threadIdx.x; i < head; i += blockDim.x) out[i] = (S)0;
    const uint32_t nv = (n - head) / per;
    uint4* v = reinterpret_cast<uint4*>(out + head);
    for (uint32_t j = threadIdx.x; j < nv; j += blockDim.x) v[j] = make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t i = head + nv * per + threadIdx.x; i < n; i += blockDim.x) out[i] = (S)0;
}

// The item of zero-run expansion chunk b: the last i with first[i] <= b (first: nitems + 1 entries).
__device__ __forceinline__ uint32_t chunk_item(const uint32_t* first, uint32_t nitems, uint32_t b) {
    uint32_t lo = 0, hi = nitems;  // first[lo] <= b < first[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid;
    }
    return lo;
}

// One workgroup per chunk of kRleChunk pairs.  Pair i (value v_i after z_i zeros) lands at
// sum_{j<i}(z_j + 1) + z_i: the chunk starts at the sum of the earlier chunks' spans (each written
// by the block decode of the item's zero counts) and scans
// its own pairs (8 consecutive per thread).  It clears its span of output positions with 16-byte
// coalesced stores (to the end of the output for the sequence's last chunk: the trailing zeros),
// then stores its values there, staged in LDS and issued with a 256-pair stride so that one store
// instruction writes 64 consecutive pairs; the span's lines are still in L2 when the values land,
// so each output line leaves L2 once.
template <class S>
__global__ __launch_bounds__(256) void k_dec_rle(const DecItem* items, uint32_t nitems, DecSeq* seqs, RleChunks rc,
                                                 CoderState* st) {
    __shared__ uint32_t sh[4];
    __shared__ uint64_t shs[4];
    __shared__ uint32_t spos[kRleChunk];
    __shared__ uint32_t sval[kRleChunk];
    const uint32_t b = blockIdx.x;
    const uint32_t ii = chunk_item(rc.item_first, nitems, b);
    const DecItem it = items[ii];
    const DecSeq& ss = seqs[it.seq];
    const DecSeq& sz = seqs[it.seq + 1];
    if (ss.err || sz.err) return;
    const uint32_t M = ss.count, n = it.n;
    const uint32_t c = b - rc.item_first[ii];
    const uint32_t clast = M == 0 ? 0u : (M - 1) / kRleChunk;
    if (c > clast) return;
    const uint32_t base = c * kRleChunk;
    const bool last = c == clast;
    // the chunk's first output position: the earlier chunks' spans
    uint64_t before = 0;
    for (uint32_t k = threadIdx.x; k < c; k += 256) before += rc.sums[rc.item_first[ii] + k];
    const uint64_t carry = block_sum_u64<256>(before, shs);
    const S* sym = static_cast<const S*>(ss.out);
    const S* zc = static_cast<const S*>(sz.out);
    S* out = static_cast<S*>(it.out);
    // LDS index swizzle: thread t's 8 consecutive entries and the strided reads are both
    // conflict-free across the 64 banks
    auto sw = [](uint32_t i) { return i ^ ((i >> 6) & 7u); };
    bool bad = false;
    // the decode buffers are 64-byte aligned: a thread's 8 pairs are two 16-byte (Symbol16) or
    // four (Symbol32) loads
    const bool vec = ((reinterpret_cast<uintptr_t>(sym) | reinterpret_cast<uintptr_t>(zc)) & 15u) == 0;
    uint32_t z[8], v[8], sum = 0;
    const uint32_t i0 = base + threadIdx.x * 8;
    if (vec && i0 + 8 <= M) {
        load8_aligned(zc + i0, z);
        load8_aligned(sym + i0, v);
#pragma unroll
        for (int k = 0; k < 8; k++) sum += z[k] + 1;
    } else {
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const uint32_t i = i0 + k;
            z[k] = i < M ? (uint32_t)zc[i] : 0u;
            v[k] = i < M ? (uint32_t)sym[i] : 0u;
            sum += i < M ? z[k] + 1 : 0u;
        }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<256>(sum, sh, tot);
    uint64_t pos = carry + ex;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        pos += z[k];
        const uint32_t e = sw(threadIdx.x * 8 + k);
        const bool inb = pos < n;
        bad |= !inb && base + threadIdx.x * 8 + k < M;
        spos[e] = inb ? (uint32_t)pos : 0xFFFFFFFFu;
        sval[e] = v[k];
        pos++;
    }
    // the chunk's span of output positions, cleared before its values land
    const uint64_t s1 = last ? (uint64_t)n : min(carry + tot, (uint64_t)n);
    if (carry < s1) clear_symbols(out + carry, (uint32_t)(s1 - carry));
    __syncthreads();  // the clearing stores complete before values land on them
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t idx = k * 256 + threadIdx.x;
        const uint32_t q = spos[sw(idx)];
        if (base + idx < M && q != 0xFFFFFFFFu) out[q] = (S)sval[sw(idx)];
    }
    if (bad) {
        seqs[it.seq].err = kDecCorrupt;
        atomicCAS(&st->err, 0u, (uint32_t)kDecCorrupt);
    }
}

// A bound decode's sequences from a resident index: the parsed header, table and block-offset
// positions are copied, the item and the output are the request's (RLE sequences decode into the
// request's buffer at rl_off).  Calls start from the index's parse result of their brick.  One
// thread per word of a record (a whole record per thread would sit in scratch memory).
constexpr uint32_t kSeqWords = sizeof(DecSeq) / 4;
static_assert(sizeof(DecSeq) % 8 == 0, "records of whole 64-bit words");
__global__ void k_dec_bind(const DecSeq* idx, const DecBind* bind, uint32_t nseq, const DecItem* items,
                           DecSeq* seqs, unsigned char* rl_arena, const DecCallOut* idx_out,
                           const uint32_t* call_brick, uint32_t ncalls, DecCallOut* outc, CoderState* st) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        st->scratch_used = 0;
        st->out_words = 0;
        st->err = 0;
        st->pad = 0;
    }
    if (t < ncalls) outc[t] = idx_out[call_brick[t]];
    const uint64_t r = t / kSeqWords;
    const uint32_t w = (uint32_t)(t % kSeqWords);
    if (r < nseq) {
        constexpr uint32_t kItem = offsetof(DecSeq, item) / 4, kOut = offsetof(DecSeq, out) / 4;
        static_assert(offsetof(DecSeq, out) % 8 == 0, "the output pointer is one 64-bit word");
        const DecBind b = bind[r];
        uint32_t v;
        if (w == kItem) {
            v = b.item;
        } else if (w == kOut || w == kOut + 1) {
            const uint64_t o = reinterpret_cast<uint64_t>(b.rl_off != ~0ull ? (void*)(rl_arena + b.rl_off)
                                                                               : items[b.item].out);
            v = w == kOut ? (uint32_t)o : (uint32_t)(o >> 32);
        } else {
            v = reinterpret_cast<const uint32_t*>(idx + b.idx_seq)[w];
        }
        reinterpret_cast<uint32_t*>(seqs + r)[w] = v;
    }
}

// Per-call status: the first error among the call's coded sequences, one wave per call, its
// lanes over the call's items.
__global__ __launch_bounds__(64) void k_dec_finish(const DecCall* calls, uint32_t ncalls, const DecItem* items,
                                                   const DecSeq* seqs, DecCallOut* outc) {
    const uint32_t c = blockIdx.x;
    if (c >= ncalls || outc[c].err) return;
    const DecCall C = calls[c];
    for (uint32_t k0 = 0; k0 < C.nitems; k0 += 64) {
        const uint32_t k = k0 + threadIdx.x;
        uint32_t e = 0;
        if (k < C.nitems) {
            const DecItem it = items[C.first_item + k];
            e = seqs[it.seq].err;
            if (!e && it.rl) e = seqs[it.seq + 1].err;
        }
        const uint64_t m = __ballot(e != 0);
        if (m) {
            const int first = __ffsll((unsigned long long)m) - 1;
            if ((int)threadIdx.x == first) outc[c].err = e;
            return;
        }
    }
}

// Calls of a decode from an encode's device-side results; a failed call, or a range outside the
// output's cap words, decodes from an empty range (its items then report a corrupt stream).
__global__ void k_calls_from_encode(DecCall* calls, const CallOut* enc, const uint32_t* base, uint32_t n,
                                    uint64_t cap) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const CallOut e = enc[c];
    const bool ok = !e.err && e.word_begin <= cap && e.words <= cap - e.word_begin;
    calls[c].words = base + (ok ? e.word_begin : 0u);
    calls[c].nwords = ok ? e.words : 0u;
}

}  // namespace

void coder_calls_from_encode(DecCall* d_calls, const CallOut* d_enc, const uint32_t* base, uint64_t cap,
                             uint32_t ncalls, hipStream_t st) {
    if (ncalls == 0) return;
    CCA_LAUNCH("dec_calls_from_encode", 0, k_calls_from_encode, dim3(div_up(ncalls, 64)), dim3(64), 0, st, d_calls,
               d_enc, base, ncalls, cap);
    CCA_LAUNCH_CHECK();
}

namespace {
// The block decode: the long sequences' four-wave workgroups, then the short ones' single waves.
template <class S>
void launch_blocks(DecSeq* d_seqs, const uint32_t* d_wave_seq, const uint32_t* d_wave_first, uint32_t nwaves,
                   uint32_t nlong, const CoderParams& p, const uint32_t* scratch, const DecItem* d_items,
                   const RleChunks& rc, uint64_t outb, hipStream_t st) {
    if (nlong) {
        CCA_LAUNCH("dec_blocks", outb, (k_dec_blocks<S, 4, 8>), dim3(nlong), dim3(256), 0, st, d_seqs, d_wave_seq,
                   d_wave_first, p, scratch, d_items, rc);
        CCA_LAUNCH_CHECK();
    }
    if (nwaves > nlong) {
        CCA_LAUNCH("dec_blocks_short", 0, (k_dec_blocks<S, 1, 8>), dim3(nwaves - nlong), dim3(64), 0, st, d_seqs,
                   d_wave_seq + nlong, d_wave_first + nlong, p, scratch, d_items, rc);
        CCA_LAUNCH_CHECK();
    }
}

// The zero-run expansion (its chunk spans come from the block decode of the zero counts).
template <class S>
void launch_rle(const DecItem* d_items, uint32_t nitems, DecSeq* d_seqs, const RleChunks& rc, CoderState* d_state,
                uint64_t outb, hipStream_t st) {
    if (rc.nchunks == 0) return;
    CCA_LAUNCH("dec_rle", outb, k_dec_rle<S>, dim3(rc.nchunks), dim3(256), 0, st, d_items, nitems, d_seqs, rc, d_state);
    CCA_LAUNCH_CHECK();
}
}  // namespace

void coder_decode(const DecCall* d_calls, uint32_t ncalls, const DecItem* d_items, uint32_t nitems, DecSeq* d_seqs,
                  uint32_t nseq, const CoderParams& p, int symSize, uint32_t* scratch, uint64_t scratch_cap,
                  CoderState* d_state, DecCallOut* d_out, const uint32_t* d_wave_seq, const uint32_t* d_wave_first,
                  uint32_t nwaves, uint32_t nlong, uint64_t total_symbols, const RleChunks& rc,
                  hipStream_t st) {
    CCA_REQUIRE(symSize == 2 || symSize == 4, CC_ERR_INVALID_ARG, "symbol size must be 2 or 4");
    CCA_LAUNCH("dec_init", 32, k_dec_init, dim3(1), dim3(64), 0, st, d_state);
    CCA_LAUNCH_CHECK();
    if (ncalls == 0) return;
    CCA_LAUNCH("dec_parse", 0, k_dec_parse, dim3(ncalls), dim3(64), 0, st, d_calls, ncalls, d_items, d_seqs,
               p, symSize, scratch, scratch_cap, d_state, d_out, 0);
    CCA_LAUNCH_CHECK();
    if (nseq == 0) return;
    CCA_LAUNCH("dec_tables", 0, k_dec_tables, dim3(nseq), dim3(256), 0, st, d_seqs, p, scratch, d_state);
    CCA_LAUNCH_CHECK();
    const uint64_t outb = total_symbols * symSize;
    if (symSize == 2)
        launch_blocks<uint16_t>(d_seqs, d_wave_seq, d_wave_first, nwaves, nlong, p, scratch, d_items, rc, outb, st);
    else
        launch_blocks<uint32_t>(d_seqs, d_wave_seq, d_wave_first, nwaves, nlong, p, scratch, d_items, rc, outb, st);
    if (symSize == 2) launch_rle<uint16_t>(d_items, nitems, d_seqs, rc, d_state, outb, st);
    else launch_rle<uint32_t>(d_items, nitems, d_seqs, rc, d_state, outb, st);
    if (ncalls) {
        CCA_LAUNCH("dec_finish", 0, k_dec_finish, dim3(ncalls), dim3(64), 0, st, d_calls, ncalls, d_items, d_seqs,
                   d_out);
        CCA_LAUNCH_CHECK();
    }
}

namespace {
template <class S>
void launch_blocks_rle(const DecItem* d_items, uint32_t nitems, DecSeq* d_seqs, const CoderParams& p,
                       const uint32_t* scratch, CoderState* d_state, const uint32_t* d_wave_seq,
                       const uint32_t* d_wave_first, uint32_t nwaves, uint32_t nlong, uint64_t outb,
                       const RleChunks& rc, hipStream_t st) {
    launch_blocks<S>(d_seqs, d_wave_seq, d_wave_first, nwaves, nlong, p, scratch, d_items, rc, outb, st);
    launch_rle<S>(d_items, nitems, d_seqs, rc, d_state, outb, st);
}
}  // namespace

void coder_decode_index(const DecCall* d_calls, uint32_t ncalls, const DecItem* d_items, DecSeq* d_seqs, uint32_t nseq,
                        const CoderParams& p, int symSize, uint32_t* scratch, uint64_t scratch_cap, CoderState* d_state,
                        DecCallOut* d_out, hipStream_t st) {
    CCA_REQUIRE(symSize == 2 || symSize == 4, CC_ERR_INVALID_ARG, "symbol size must be 2 or 4");
    CCA_LAUNCH("dec_init", 32, k_dec_init, dim3(1), dim3(64), 0, st, d_state);
    CCA_LAUNCH_CHECK();
    if (ncalls == 0) return;
    CCA_LAUNCH("dec_parse", 0, k_dec_parse, dim3(ncalls), dim3(64), 0, st, d_calls, ncalls, d_items, d_seqs,
               p, symSize, scratch, scratch_cap, d_state, d_out, 1);
    CCA_LAUNCH_CHECK();
    if (nseq == 0) return;
    CCA_LAUNCH("dec_tables", 0, k_dec_tables, dim3(nseq), dim3(256), 0, st, d_seqs, p, scratch, d_state);
    CCA_LAUNCH_CHECK();
    if (ncalls) {
        CCA_LAUNCH("dec_finish", 0, k_dec_finish, dim3(ncalls), dim3(64), 0, st, d_calls, ncalls, d_items, d_seqs,
                   d_out);
        CCA_LAUNCH_CHECK();
    }
}

void coder_decode_bound(const DecSeq* d_idx_seqs, const uint32_t* idx_scratch, const DecCallOut* d_idx_out,
                        const DecBind* d_bind, const uint32_t* d_call_brick, const DecCall* d_calls, uint32_t ncalls,
                        const DecItem* d_items, uint32_t nitems, DecSeq* d_seqs, uint32_t nseq, const CoderParams& p,
                        int symSize, unsigned char* rl_arena, CoderState* d_state, DecCallOut* d_out,
                        const uint32_t* d_wave_seq, const uint32_t* d_wave_first, uint32_t nwaves,
                        uint32_t nlong, uint64_t total_symbols, const RleChunks& rc, hipStream_t st) {
    CCA_REQUIRE(symSize == 2 || symSize == 4, CC_ERR_INVALID_ARG, "symbol size must be 2 or 4");
    const uint64_t nb = std::max<uint64_t>((uint64_t)nseq * kSeqWords, ncalls);
    if (nb == 0) {
        CCA_LAUNCH("dec_init", 32, k_dec_init, dim3(1), dim3(64), 0, st, d_state);
        CCA_LAUNCH_CHECK();
        return;
    }
    // the job state is reset by the bind (its first thread)
    CCA_LAUNCH("dec_bind", (uint64_t)nseq * 2 * sizeof(DecSeq), k_dec_bind, dim3(div_up(nb, 256)),
               dim3(256), 0, st,
               d_idx_seqs, d_bind, nseq, d_items, d_seqs, rl_arena, d_idx_out, d_call_brick, ncalls, d_out, d_state);
    CCA_LAUNCH_CHECK();
    const uint64_t outb = total_symbols * symSize;
    if (symSize == 2)
        launch_blocks_rle<uint16_t>(d_items, nitems, d_seqs, p, idx_scratch, d_state, d_wave_seq, d_wave_first, nwaves,
                                    nlong, outb, rc, st);
    else
        launch_blocks_rle<uint32_t>(d_items, nitems, d_seqs, p, idx_scratch, d_state, d_wave_seq, d_wave_first, nwaves,
                                    nlong, outb, rc, st);
    if (ncalls) {
        CCA_LAUNCH("dec_finish", 0, k_dec_finish, dim3(ncalls), dim3(64), 0, st, d_calls, ncalls, d_items, d_seqs,
                   d_out);
        CCA_LAUNCH_CHECK();
    }
}

}  // namespace cca
