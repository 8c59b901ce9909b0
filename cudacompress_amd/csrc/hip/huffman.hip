// Histogram and canonical Huffman coding for gfx950.
//
// Histogram (src/cudaCompress/Histogram.cu:110-289): exact per-stream counts.  Each workgroup
// privatises up to 16384 bins in LDS (64 KB), counts a 32768-symbol chunk with LDS atomics and
// flushes non-zero bins with one global atomicAdd each.  Alphabets of up to 2^16 bins (log2 16)
// use 16-bit counters, two per LDS word (128 KB of CDNA4's 160 KB), over chunks of fewer than
// 2^16 symbols so that a counter cannot carry into its neighbour; only larger alphabets (Symbol32)
// send the symbols beyond the LDS bins to global atomics.
//
// Encode (src/cudaCompress/Huffman.cu:73-187, HuffmanKernels.cui:36-190): the bitstream is the
// concatenation of codewords MSB-first in 32-bit words, the same bits a sequential writeBits
// produces.  Three launches:
//   block_bits : one wave per coding block sums its codeword lengths
//   scan       : one workgroup per stream -> bit offset of every coding block (the stored
//                offsets) and the stream's bit count; also clears the words that two pack chunks
//                share, so no memset of the output is needed
//   pack       : 2048 symbols per workgroup; per-thread bit positions from a workgroup scan,
//                codewords OR-ed into an LDS word buffer, interior words stored plainly, the two
//                boundary words atomically OR-ed.
//
// Decode (HuffmanKernels.cui:197-314, BitStreamGPU.cuh): one lane per coding block starting at
// the block's stored bit offset, a 64-bit bit window, a 1024-entry LDS table resolving codes of
// up to 10 bits to (length, symbol) in one lookup and the canonical first/min/max search beyond
// that.  Decoded symbols are staged in LDS 32 per lane and written as coalesced 32-symbol runs.
#include "../kernels.hpp"
#include "wave.hpp"

namespace cca {
namespace {

constexpr int kHistLds = 16384;
constexpr uint32_t kHistChunk = 32768;
constexpr uint32_t kHistWideBins = 65536;            // 16-bit counters, two per LDS word
constexpr uint32_t kHistWideChunk = 60 * 1024;       // < 2^16 symbols per workgroup
constexpr uint32_t kHistWideThreads = 1024;
constexpr int kPackSyms = 2048;  // symbols per pack workgroup (8 per thread)
constexpr int kLutBits = 10;

template <typename S>
__device__ __forceinline__ uint32_t ld(const void* p, uint32_t i) { return (uint32_t)((const S*)p)[i]; }

template <typename S>
__global__ __launch_bounds__(256) void k_hist(const HistStream* __restrict__ streams) {
    const HistStream d = streams[blockIdx.y];
    const uint32_t start = blockIdx.x * kHistChunk;
    if (start >= d.n && !(blockIdx.x == 0 && d.quirk)) return;
    __shared__ uint32_t h[kHistLds];
    const uint32_t nl = min(d.nbins, (uint32_t)kHistLds);
    for (uint32_t i = threadIdx.x; i < nl; i += 256) h[i] = 0;
    __syncthreads();
    const uint32_t end = min(start + kHistChunk, d.n);
    for (uint32_t i = start + threadIdx.x; i < end; i += 256) {
        uint32_t v = ld<S>(d.data, i);
        if (v < nl) atomicAdd(&h[v], 1u);
        else if (v < d.nbins) atomicAdd(&d.hist[v], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nl; i += 256) {
        uint32_t c = h[i];
        if (c) atomicAdd(&d.hist[i], c);
    }
    if (d.quirk && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&d.hist[65535], 1u);
}

// Alphabets of up to 2^16 bins: bin v counts in half (v & 1) of LDS word v >> 1.  A chunk holds
// fewer than 2^16 symbols, so no half overflows into the other.
template <typename S>
__global__ __launch_bounds__(kHistWideThreads) void k_hist_wide(const HistStream* __restrict__ streams) {
    const HistStream d = streams[blockIdx.y];
    const uint32_t start = blockIdx.x * kHistWideChunk;
    if (start >= d.n && !(blockIdx.x == 0 && d.quirk)) return;
    extern __shared__ uint32_t h2[];
    const uint32_t nl = min(d.nbins, kHistWideBins);
    const uint32_t nw = (nl + 1) / 2;
    for (uint32_t i = threadIdx.x; i < nw; i += kHistWideThreads) h2[i] = 0;
    __syncthreads();
    const uint32_t end = min(start + kHistWideChunk, d.n);
    for (uint32_t i = start + threadIdx.x; i < end; i += kHistWideThreads) {
        const uint32_t v = ld<S>(d.data, i);
        if (v < nl) atomicAdd(&h2[v >> 1], (v & 1u) ? 65536u : 1u);
        else if (v < d.nbins) atomicAdd(&d.hist[v], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nw; i += kHistWideThreads) {
        const uint32_t w = h2[i];
        if (w & 0xFFFFu) atomicAdd(&d.hist[2 * i], w & 0xFFFFu);
        if (w >> 16) atomicAdd(&d.hist[2 * i + 1], w >> 16);
    }
    if (d.quirk && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&d.hist[65535], 1u);
}

template <typename S>
__global__ __launch_bounds__(256) void k_block_bits(const HuffEncStream* __restrict__ streams, uint32_t* __restrict__ bits,
                                                    const uint32_t* __restrict__ block_stream, uint32_t nblocks,
                                                    uint32_t cbs) {
    const uint32_t gb = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (gb >= nblocks) return;
    const HuffEncStream d = streams[block_stream[gb]];
    const uint32_t b = gb - d.block_base;
    const uint32_t start = b * cbs, end = min(start + cbs, d.n);
    uint32_t acc = 0;
    for (uint32_t i = start + lane_id(); i < end; i += 64) acc += d.len[ld<S>(d.sym, i)];
    acc = wave_sum(acc);
    if (lane_id() == 0) bits[gb] = acc;
}

__global__ __launch_bounds__(1024) void k_block_scan(const HuffEncStream* __restrict__ streams,
                                                     const uint32_t* __restrict__ bits, uint32_t* __restrict__ stream_bits,
                                                     uint32_t cbs) {
    const HuffEncStream d = streams[blockIdx.x];
    const uint32_t nb = div_up_dev(d.n, cbs);
    const uint32_t per_chunk = kPackSyms / cbs;
    __shared__ uint32_t scratch[16];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += 1024) {
        uint32_t i = c0 + threadIdx.x;
        uint32_t v = i < nb ? bits[d.block_base + i] : 0u;
        uint32_t total;
        uint32_t e = block_excl_scan<1024>(v, scratch, total);
        if (i < nb) {
            d.offsets[i] = carry + e;
            // the first word of every pack chunk is OR-ed into by two chunks: clear it
            if (i % per_chunk == 0) d.words[(carry + e) >> 5] = 0u;
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        d.words[carry >> 5] = 0u;
        stream_bits[blockIdx.x] = carry;
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_pack(const HuffEncStream* __restrict__ streams,
                                              const uint32_t* __restrict__ chunk_stream, uint32_t cbs) {
    const HuffEncStream d = streams[chunk_stream[blockIdx.x]];
    const uint32_t c = blockIdx.x - d.chunk_base;
    const uint32_t s0 = c * kPackSyms;
    const uint32_t i0 = s0 + threadIdx.x * 8;
    uint32_t code[8], len[8], tsum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (i0 + k < d.n) {
            uint32_t s = ld<S>(d.sym, i0 + k);
            code[k] = d.code[s];
            len[k] = d.len[s];
        } else {
            code[k] = 0;
            len[k] = 0;
        }
        tsum += len[k];
    }
    __shared__ uint32_t scratch[4];
    __shared__ uint32_t lw[kPackSyms + 4];
    uint32_t total;
    uint32_t excl = block_excl_scan<256>(tsum, scratch, total);
    const uint32_t startbit = d.offsets[s0 / cbs];
    const uint32_t rel = startbit & 31u;
    const uint32_t nw = total ? (rel + total + 31) / 32 : 0u;
    for (uint32_t w = threadIdx.x; w < nw; w += 256) lw[w] = 0u;
    __syncthreads();
    uint32_t p = rel + excl;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t l = len[k];
        if (l) {
            const uint32_t wi = p >> 5, o = p & 31u;
            if (o + l <= 32) {
                atomicOr(&lw[wi], code[k] << (32 - o - l));
            } else {
                atomicOr(&lw[wi], code[k] >> (o + l - 32));
                atomicOr(&lw[wi + 1], code[k] << (64 - o - l));
            }
            p += l;
        }
    }
    __syncthreads();
    // word 0 may be shared with the previous chunk and the last word with the next one (or is the
    // stream's final word); both were cleared by the scan.  A last word that ends exactly on a
    // word boundary belongs to this chunk alone and is stored like the interior words.
    uint32_t* out = d.words + (startbit >> 5);
    const bool end_shared = ((rel + total) & 31u) != 0;
    for (uint32_t w = threadIdx.x; w < nw; w += 256) {
        if (w == 0 || (w == nw - 1 && end_shared)) atomicOr(&out[w], lw[w]);
        else out[w] = lw[w];
    }
}

template <typename S>
__global__ __launch_bounds__(64) void k_decode(const HuffDecStream* __restrict__ streams,
                                               const uint32_t* __restrict__ wave_stream, uint32_t cbs) {
    const HuffDecStream d = streams[wave_stream[blockIdx.x]];
    const int lane = threadIdx.x;
    const uint32_t nblocks = div_up_dev(d.n, cbs);
    const uint32_t fb = (blockIdx.x - d.wave_base) * 64;
    const uint32_t b = fb + lane;
    __shared__ uint32_t lut[1 << kLutBits];
    __shared__ int tab[96];
    __shared__ uint32_t stage[64][33];
    S* out = (S*)d.out;
    const S* symtab = (const S*)d.symbols;
    const bool multi = d.nsyms > 1;
    if (multi) {
        for (int i = lane; i < 96; i += 64) tab[i] = d.table[i];
        __syncthreads();
        for (uint32_t e = lane; e < (1u << kLutBits); e += 64) {
            uint32_t entry = 0;
            for (int l = 1; l <= kLutBits; l++) {
                int cw = (int)(e >> (kLutBits - l));
                if (cw <= tab[64 + l - 1]) {
                    int idx = tab[l - 1] + cw - tab[32 + l - 1];
                    uint32_t sym = (idx >= 0 && (uint32_t)idx < d.nsyms) ? (uint32_t)symtab[idx] : 0u;
                    entry = ((uint32_t)l << 24) | (sym & 0xFFFFFFu);
                    break;
                }
            }
            lut[e] = entry;
        }
        __syncthreads();
    }
    const uint32_t single = (d.nsyms >= 1) ? (uint32_t)symtab[0] : 0u;
    const bool valid = b < nblocks;
    const uint32_t cnt = valid ? min(cbs, d.n - b * cbs) : 0u;
    // bit window
    uint64_t cache = 0;
    uint32_t bp = 0, nxt = 0;
    if (valid && multi) {
        uint32_t pos = d.offsets[b];
        uint32_t wi = pos >> 5;
        cache = ((uint64_t)d.words[wi] << 32) | (uint64_t)d.words[wi + 1];
        bp = pos & 31u;
        nxt = wi + 2;
    }
    for (uint32_t r = 0; r < cbs; r += 32) {
        const uint32_t rend = min(r + 32, cnt);
        for (uint32_t i = r; i < rend; i++) {
            uint32_t sym;
            if (multi) {
                if (bp >= 32) {
                    // a corrupt stream cannot make a lane read past the two padding words
                    cache = (cache << 32) | (uint64_t)d.words[min(nxt, d.nwords + 1)];
                    nxt++;
                    bp -= 32;
                }
                const uint32_t peek = (uint32_t)(cache >> (32 - bp));
                const uint32_t e = lut[peek >> (32 - kLutBits)];
                uint32_t len;
                if (e) {
                    len = e >> 24;
                    sym = e & 0xFFFFFFu;
                } else {
                    len = kLutBits + 1;
                    while (len < 32 && (int)(peek >> (32 - len)) > tab[64 + len - 1]) len++;
                    int cw = (int)(peek >> (32 - len));
                    int idx = tab[len - 1] + cw - tab[32 + len - 1];
                    sym = (idx >= 0 && (uint32_t)idx < d.nsyms) ? (uint32_t)symtab[idx] : 0u;
                }
                bp += len;
            } else {
                sym = single;
            }
            stage[lane][i - r] = sym;
        }
        __syncthreads();
        for (uint32_t idx = lane; idx < 64 * 32; idx += 64) {
            const uint32_t blk = idx >> 5, k = idx & 31u;
            const uint32_t gb = fb + blk;
            if (gb < nblocks) {
                const uint32_t pos = gb * cbs + r + k;
                const uint32_t bend = min((gb + 1) * cbs, d.n);
                if (pos < bend) out[pos] = (S)stage[blk][k];
            }
        }
        __syncthreads();
    }
}

}  // namespace

void histogram(const HistStream* d_streams, const HistStream* h_streams, uint32_t nstreams, int symSize,
               hipStream_t st) {
    if (nstreams == 0) return;
    CCA_REQUIRE(nstreams <= 65535u, CC_ERR_INVALID_ARG, "too many streams for one launch");
    uint32_t maxbins = 0;
    for (uint32_t i = 0; i < nstreams; i++) maxbins = std::max(maxbins, h_streams[i].nbins);
    const bool wide = maxbins > (uint32_t)kHistLds;
    const uint32_t chunk = wide ? kHistWideChunk : kHistChunk;
    uint32_t maxchunks = 1;
    for (uint32_t i = 0; i < nstreams; i++) maxchunks = std::max(maxchunks, div_up(h_streams[i].n, chunk));
    dim3 grid(maxchunks, nstreams);
    uint64_t bytes = 0;
    for (uint32_t i = 0; i < nstreams; i++) bytes += (uint64_t)h_streams[i].n * symSize + 4ull * h_streams[i].nbins;
    if (wide) {
        const size_t lds = (std::min(maxbins, kHistWideBins) + 1) / 2 * sizeof(uint32_t);
        if (symSize == 2)
            CCA_LAUNCH("histogram", bytes, k_hist_wide<uint16_t>, grid, dim3(kHistWideThreads), lds, st, d_streams);
        else
            CCA_LAUNCH("histogram", bytes, k_hist_wide<uint32_t>, grid, dim3(kHistWideThreads), lds, st, d_streams);
        CCA_LAUNCH_CHECK();
        return;
    }
    if (symSize == 2) CCA_LAUNCH("histogram", bytes, k_hist<uint16_t>, grid, dim3(256), 0, st, d_streams);
    else CCA_LAUNCH("histogram", bytes, k_hist<uint32_t>, grid, dim3(256), 0, st, d_streams);
    CCA_LAUNCH_CHECK();
}

void huffman_encode(const HuffEncStream* d_streams, const HuffEncStream* h_streams, uint32_t nstreams, uint32_t cbs,
                    int symSize, uint32_t* d_block_bits, uint32_t nblocks, const uint32_t* d_block_stream,
                    uint32_t* d_stream_bits, const uint32_t* d_chunk_stream, uint32_t nchunks, uint64_t total_bits,
                    hipStream_t st) {
    if (nstreams == 0) return;
    uint64_t syms = 0;
    for (uint32_t i = 0; i < nstreams; i++) syms += h_streams[i].n;
    const uint64_t inb = syms * symSize;
    CCA_REQUIRE(cbs >= 32 && cbs <= 256 && (cbs & (cbs - 1)) == 0, CC_ERR_INVALID_ARG,
                "codingBlockSize must be a power of two in [32, 256]");
    if (nblocks > 0) {
        if (symSize == 2) CCA_LAUNCH("huff_block_bits", inb + 4ull * nblocks, k_block_bits<uint16_t>, dim3(div_up(nblocks, 4)), dim3(256), 0, st, d_streams, d_block_bits, d_block_stream, nblocks, cbs);
        else CCA_LAUNCH("huff_block_bits", inb + 4ull * nblocks, k_block_bits<uint32_t>, dim3(div_up(nblocks, 4)), dim3(256), 0, st, d_streams, d_block_bits, d_block_stream, nblocks, cbs);
        CCA_LAUNCH_CHECK();
    }
    CCA_LAUNCH("huff_block_scan", 8ull * nblocks, k_block_scan, dim3(nstreams), dim3(1024), 0, st, d_streams, d_block_bits, d_stream_bits, cbs);
    CCA_LAUNCH_CHECK();
    if (nchunks > 0) {
        if (symSize == 2) CCA_LAUNCH("huff_pack", inb + total_bits / 8, k_pack<uint16_t>, dim3(nchunks), dim3(256), 0, st, d_streams, d_chunk_stream, cbs);
        else CCA_LAUNCH("huff_pack", inb + total_bits / 8, k_pack<uint32_t>, dim3(nchunks), dim3(256), 0, st, d_streams, d_chunk_stream, cbs);
        CCA_LAUNCH_CHECK();
    }
}

void huffman_decode(const HuffDecStream* d_streams, uint32_t nstreams, const uint32_t* d_wave_stream, uint32_t nwaves,
                    uint32_t cbs, int symSize, uint64_t total_syms, uint64_t total_bits, hipStream_t st) {
    const uint64_t bytes = total_syms * symSize + total_bits / 8;
    if (nstreams == 0 || nwaves == 0) return;
    CCA_REQUIRE(cbs >= 32 && cbs <= 256 && (cbs & (cbs - 1)) == 0, CC_ERR_INVALID_ARG,
                "codingBlockSize must be a power of two in [32, 256]");
    if (symSize == 2) CCA_LAUNCH("huff_decode", bytes, k_decode<uint16_t>, dim3(nwaves), dim3(64), 0, st, d_streams, d_wave_stream, cbs);
    else CCA_LAUNCH("huff_decode", bytes, k_decode<uint32_t>, dim3(nwaves), dim3(64), 0, st, d_streams, d_wave_stream, cbs);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
