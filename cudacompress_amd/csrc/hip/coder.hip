// Device-resident entropy coder for gfx950: stats -> design -> pack -> layout -> gather.
// The bitstream it writes is word for word the layout of src/cudaCompress/Encode.cpp:402-463
// (encodeRLHuff) and :751-782 (encodeHuff) with the table headers of
// src/cudaCompress/HuffmanTable.cu:470-485; see coder.hpp for the pipeline.
//
// Zero-run RLE is never materialised.  Every input position emits at most one (symbol, zero
// count) pair: a nonzero symbol v emits (v, run % 256) where run is the number of zeros since
// the previous nonzero, and the zero that completes each 256-zero block of a run emits the
// filler (0, 255) -- but only when a nonzero follows later in the stream (trailing zeros emit
// nothing).  This is the pair sequence of src/cudaCompress/RunLengthKernels.cui:5-44 (fillers
// for runs >= 256, src/CPU/RunLengthCPU.cpp:8-32), with each filler attributed to the zero that
// completes it, so a 2048-symbol step of a workgroup never yields more than 2048 pairs.
#include <type_traits>

#include "../coder.hpp"
#include "../profile.hpp"
#include "huff_design.hpp"
#include "wave.hpp"

namespace cca {
namespace {

using design::Ctl;
using design::HeapEntry;
using design::TableOut;
using design::Work;

constexpr uint32_t kDesignMid = 4096;
// LDS packing window of each coded sequence, a ring indexed by output word mod kWindowWords: a chunk
// of kChunk pairs of at most 31 bits spans at most 1985 words plus the partial word before it.
constexpr uint32_t kWindowWords = 2048;
static_assert(kWindowWords >= (kChunk * 31 + 31) / 32 + 2 && (kWindowWords & (kWindowWords - 1)) == 0,
              "window ring");

// ------------------------------------------------------------------------ helpers
// Exclusive scans over a 256-thread workgroup.  sh: 4 entries of LDS.  All threads must call.
__device__ uint64_t wg_excl_sum_u64(uint64_t v, uint64_t* sh, uint64_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t x = wave_incl_scan_u64(v);
    if (lane == 63) sh[w] = x;
    __syncthreads();
    uint64_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t s = sh[i];
        if (i < w) base += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}
__device__ long long wg_excl_max_i64(long long v, long long* sh, long long& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long x = wave_incl_max_i64(v);
    const long long ex = wave_shift_up_i64(x, -1);
    if (lane == 63) sh[w] = x;
    __syncthreads();
    long long base = -1, tot = -1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long long s = sh[i];
        if (i < w) base = base > s ? base : s;
        tot = tot > s ? tot : s;
    }
    __syncthreads();
    total = tot;
    return base > ex ? base : ex;
}
// wg_excl_max_i64 without the closing barrier: sh must not be written again before the next
// barrier (callers alternate two arrays).
__device__ __forceinline__ long long wg_excl_max_i64_open(long long v, long long* sh, long long& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long x = wave_incl_max_i64(v);
    const long long ex = wave_shift_up_i64(x, -1);
    if (lane == 63) sh[w] = x;
    __syncthreads();
    long long base = -1, tot = -1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const long long s = sh[i];
        if (i < w) base = base > s ? base : s;
        tot = tot > s ? tot : s;
    }
    total = tot;
    return base > ex ? base : ex;
}
// Exclusive sum of a and exclusive max of b (identity -1) over a 256-thread workgroup with one
// barrier: sha and shb (4 entries each) must not be written again before the next barrier.
__device__ __forceinline__ void wg_excl_sum_max(uint64_t a, long long b, uint64_t* sha, long long* shb,
                                                uint64_t& ex_a, uint64_t& tot_a, long long& ex_b, long long& tot_b) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t xa = wave_incl_scan_u64(a);
    const long long xb = wave_incl_max_i64(b);
    const long long eb = wave_shift_up_i64(xb, -1);
    if (lane == 63) {
        sha[w] = xa;
        shb[w] = xb;
    }
    __syncthreads();
    uint64_t base_a = 0, ta = 0;
    long long base_b = -1, tb = -1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint64_t sa = sha[i];
        const long long sb = shb[i];
        if (i < w) {
            base_a += sa;
            base_b = base_b > sb ? base_b : sb;
        }
        ta += sa;
        tb = tb > sb ? tb : sb;
    }
    ex_a = base_a + xa - a;
    tot_a = ta;
    ex_b = base_b > eb ? base_b : eb;
    tot_b = tb;
}
__device__ uint32_t wg_sum_u32(uint32_t v, uint32_t* sh) {
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint32_t t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}
__device__ uint32_t wg_max_u32(uint32_t v, uint32_t* sh) {
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    const uint32_t a = max(sh[0], sh[1]), b = max(sh[2], sh[3]);
    __syncthreads();
    return max(a, b);
}

// Scratch arena allocation (one thread).  Returns ~0 on exhaustion and flags the job.
__device__ uint64_t arena_alloc(CoderState* st, uint64_t words, uint64_t cap, SeqInfo* seq) {
    words = (words + 15) & ~uint64_t(15);
    const uint64_t off = atomicAdd(&st->scratch_used, (unsigned long long)words);
    if (off + words > cap) {
        seq->err = kErrScratch;
        atomicCAS(&st->err, 0u, (uint32_t)kErrScratch);
        return ~uint64_t(0);
    }
    return off;
}

__device__ __forceinline__ uint32_t req_bits(uint32_t v) { return v ? 32u - (uint32_t)__clz(v) : 0u; }

template <class S>
__device__ __forceinline__ void load8(const S* in, uint64_t p0, uint32_t n, uint32_t (&v)[kPerThread]) {
    if (p0 + kPerThread <= n && ((reinterpret_cast<uintptr_t>(in + p0) & 15) == 0)) {
        if (sizeof(S) == 2) {
            const uint4 q = *reinterpret_cast<const uint4*>(in + p0);
            const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int j = 0; j < 4; j++) { v[2 * j] = w[j] & 0xFFFFu; v[2 * j + 1] = w[j] >> 16; }
        } else {
            const uint4 a = *reinterpret_cast<const uint4*>(in + p0);
            const uint4 b = *reinterpret_cast<const uint4*>(in + p0 + 4);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
        }
    } else {
#pragma unroll
        for (int j = 0; j < (int)kPerThread; j++) v[j] = (p0 + j < n) ? (uint32_t)in[p0 + j] : 0u;
    }
}

// Pairs of one thread's 8 positions, computed without branches: a pair is emitted for every
// symbol of a plain stream, and in a run-length stream for every nonzero and for every 256th zero
// of a run (a filler, symbol 0 with zero count 255) at positions < limit; the pair's zero count is
// the length of the zero run before it modulo 256.  prev: position of the last nonzero before the
// thread's positions (-1 none); limit: the stream's last nonzero (pass A passes n and removes the
// trailing run's fillers afterwards).  Returns the emit mask (bit j: position p0 + j) and the zero
// counts in zc.
__device__ __forceinline__ uint32_t thread_flags(const uint32_t (&v)[kPerThread], uint64_t p0, uint32_t n, bool rl,
                                                 long long prev, long long limit, uint32_t (&zc)[kPerThread]) {
    const uint32_t nvalid = p0 >= n ? 0u : (n - p0 >= kPerThread ? kPerThread : (uint32_t)(n - p0));
    const long long lim_rel = limit - (long long)p0;
    uint32_t g = (uint32_t)((long long)p0 - prev - 1);  // only its low 8 bits matter
    uint32_t emit = 0;
#pragma unroll
    for (int j = 0; j < (int)kPerThread; j++) {
        const bool nz = v[j] != 0u;
        const uint32_t z = g & 255u;
        const bool e = (uint32_t)j < nvalid && (!rl || nz || (z == 255u && (long long)j < lim_rel));
        g = nz ? 0u : g + 1u;
        zc[j] = rl ? z : 0u;
        emit |= (uint32_t)e << j;
    }
    return emit;
}

// f(sym, zc) for each pair of thread_flags, in order.
template <class F>
__device__ __forceinline__ void thread_pairs(const uint32_t (&v)[kPerThread], uint64_t p0, uint32_t n, bool rl,
                                             long long prev, long long limit, F f) {
    uint32_t zc[kPerThread];
    const uint32_t emit = thread_flags(v, p0, n, rl, prev, limit, zc);
#pragma unroll
    for (int j = 0; j < (int)kPerThread; j++)
        if ((emit >> j) & 1u) f(v[j], zc[j]);
}

__device__ __forceinline__ long long thread_last_nz(const uint32_t (&v)[kPerThread], uint64_t p0, uint32_t n) {
    const uint32_t nvalid = p0 >= n ? 0u : (n - p0 >= kPerThread ? kPerThread : (uint32_t)(n - p0));
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < (int)kPerThread; j++) m |= (uint32_t)(v[j] != 0u && (uint32_t)j < nvalid) << j;
    return m ? (long long)p0 + (31 - __clz(m)) : -1ll;
}

// Chunk-relative forms for the chunk loops of the stats and pack passes.  A chunk of kChunk
// positions starts at `base`; thread t holds positions base + t8 + j (t8 = kPerThread t), nvalid
// of them inside the stream.  A position enters the pair rule only relative to the chunk or
// through its low 8 bits (the zero-run length modulo 256), so per-thread arithmetic is 32-bit and
// the 64-bit stream state (base, the last nonzero before the chunk) stays uniform.
__device__ __forceinline__ uint32_t chunk_nvalid(uint64_t base, uint32_t n, uint32_t t8) {
    const uint64_t rem = (uint64_t)n - base;
    const uint32_t r = rem > (uint64_t)kChunk ? kChunk : (uint32_t)rem;
    return r > t8 ? (r - t8 < kPerThread ? r - t8 : kPerThread) : 0u;
}
// the thread's last nonzero, chunk-relative (-1: none)
__device__ __forceinline__ int thread_last_nz_rel(const uint32_t (&v)[kPerThread], uint32_t t8, uint32_t nvalid) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < (int)kPerThread; j++) m |= (uint32_t)(v[j] != 0u && (uint32_t)j < nvalid) << j;
    return m ? (int)(t8 + 31u - (uint32_t)__clz(m)) : -1;
}
// thread_flags with g0 = (first position - last nonzero before it - 1) mod 2^32 and
// lim = limit - first position (clamped; only lim > j matters)
__device__ __forceinline__ uint32_t thread_flags_rel(const uint32_t (&v)[kPerThread], uint32_t nvalid, bool rl,
                                                     uint32_t g0, int lim, uint32_t (&zc)[kPerThread]) {
    uint32_t g = g0, emit = 0;
#pragma unroll
    for (int j = 0; j < (int)kPerThread; j++) {
        const bool nz = v[j] != 0u;
        const uint32_t z = g & 255u;
        const bool e = (uint32_t)j < nvalid && (!rl || nz || (z == 255u && j < lim));
        g = nz ? 0u : g + 1u;
        zc[j] = rl ? z : 0u;
        emit |= (uint32_t)e << j;
    }
    return emit;
}
// Uniform per-chunk terms: gc = (base - carry - 1) mod 2^32 for the last nonzero `carry` before
// the chunk (-1: none), limc = limit - base clamped to [-1, kChunk + kPerThread].
__device__ __forceinline__ uint32_t chunk_gc(uint64_t base, long long carry) {
    return (uint32_t)(base - (uint64_t)(carry + 1));
}
__device__ __forceinline__ int chunk_lim(uint64_t base, long long limit) {
    const long long d = limit - (long long)base;
    return d < -1 ? -1 : (d > (long long)(kChunk + kPerThread) ? (int)(kChunk + kPerThread) : (int)d);
}
// g0 of a thread from its exclusive chunk-relative last nonzero prel (-1: before the chunk)
__device__ __forceinline__ uint32_t thread_g0(int prel, uint32_t t8, uint32_t gc) {
    return prel >= 0 ? t8 - (uint32_t)prel - 1u : gc + t8;
}

// Exclusive last nonzero of the chunk's threads (chunk-relative positions grow with the thread
// index, -1: none): in a wave a ballot and one lane read (the nearest lower lane holding a
// nonzero), across the waves the entries of the waves before.  No barrier here: the caller writes
// sh[wave] (the wave's last, `wlast`) before its barrier and combines after it (wg_last_combine).
__device__ __forceinline__ int wave_excl_last(int b, int& wlast) {
    const int lane = threadIdx.x & 63;
    const uint64_t m = __ballot(b >= 0);
    const uint64_t below = m & ((1ull << lane) - 1ull);
    const int got = __shfl(b, below ? 63 - __clzll(below) : 0);
    wlast = m ? __builtin_amdgcn_readlane(b, 63 - __clzll(m)) : -1;
    return below ? got : -1;
}
__device__ __forceinline__ void wg_last_combine(int ex, const int* sh, int& excl, int& total) {
    const int w = threadIdx.x >> 6;
    int base = -1, tot = -1;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int s = sh[i];
        if (i < w) base = base > s ? base : s;
        tot = tot > s ? tot : s;
    }
    excl = ex >= 0 ? ex : base;
    total = __builtin_amdgcn_readfirstlane(tot);
}
// a uniform 64-bit value held in vector registers, moved to scalar ones
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(v >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
}

// ------------------------------------------------------------------------ kernels
__global__ void k_coder_init(CoderState* st, uint32_t* classes) {
    if (threadIdx.x < kDesignListBase) classes[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        st->scratch_used = 0;
        st->out_words = 0;
        st->out_base = 0;
        st->err = 0;
        st->pad = 0;
    }
}

// Words of a histogram of nb bins in the arena and, when the table may be designed one per thread
// (k < kLaneMaxK + 1 used symbols), its leaf list after it: (symbol, count) pairs in symbol order
// from the next 4-word boundary.
__device__ __forceinline__ uint64_t leaves_at(uint32_t nb) { return ((uint64_t)nb + 3) & ~uint64_t(3); }
__device__ __forceinline__ uint64_t hist_words(uint32_t nb, uint32_t k, bool leaves) {
    return leaves && k <= design::kLaneMaxK ? leaves_at(nb) + 2ull * k : (uint64_t)nb;
}

// The used bins of h[0, nb) (nb <= kStatBins) as (symbol, count) pairs in symbol order; every
// thread of the workgroup calls.
__device__ void write_leaves(const uint32_t* h, uint32_t nb, uint2* out, uint64_t* sh) {
    constexpr uint32_t per = kStatBins / kCoderThreads;
    const uint32_t lo = threadIdx.x * per;
    uint32_t c = 0;
    for (uint32_t j = 0; j < per; j++) c += (lo + j < nb && h[lo + j]) ? 1u : 0u;
    uint64_t tot;
    uint64_t r = wg_excl_sum_u64(c, sh, tot);
    for (uint32_t j = 0; j < per; j++) {
        const uint32_t b = lo + j;
        if (b < nb && h[b]) out[r++] = make_uint2(b, h[b]);
    }
}

// Pass A: pair counts, histograms (LDS), maxima; tables of the LDS class get their histogram
// copied to the arena (with the leaf list of small tables); streams with symbols >= kStatBins are
// marked wide and get a zeroed global histogram for k_code_stats_wide.
template <class S>
__global__ __launch_bounds__(kCoderThreads) void k_code_stats(const CodeItem* items, SeqInfo* seqs, CoderParams p,
                                                              uint32_t* scratch, uint64_t cap, CoderState* st,
                                                              const uint32_t* order) {
    __shared__ uint32_t hs[kStatBins];
    __shared__ uint32_t hz[256];
    __shared__ int shl[2][4];  // alternate chunks: one barrier per chunk
    __shared__ uint32_t shu[4];
    __shared__ uint64_t sh_off[2];
    __shared__ uint64_t shq[4];
    const uint32_t ii = order ? order[blockIdx.x] : blockIdx.x;
    const CodeItem it = items[ii];
    const S* in = static_cast<const S*>(it.in);
    const bool rl = it.rl != 0;
    const uint32_t n = it.n;
    for (uint32_t i = threadIdx.x; i < kStatBins; i += kCoderThreads) hs[i] = 0;
    for (uint32_t i = threadIdx.x; i < 256; i += kCoderThreads) hz[i] = 0;
    __syncthreads();
    long long carry = -1;
    uint32_t cnt = 0, vmax = 0, wide = 0;
    // histogram updates are coalesced per thread: runs of one symbol and all zero counts of zero
    // are counted in registers (zero runs dominate the zero-count histogram)
    uint32_t run_sym = 0xFFFFFFFFu, run_len = 0, zc0 = 0;
    auto flush_run = [&]() {
        if (run_len) {
            if (run_sym < kStatBins) atomicAdd(&hs[run_sym], run_len);
            else wide = 1;
        }
    };
    const uint32_t t8 = threadIdx.x * kPerThread;
    uint32_t vn[kPerThread];
    load8(in, (uint64_t)t8, n, vn);
    uint32_t par = 0;
    for (uint64_t base = 0; base < n; base += kChunk, par ^= 1) {
        uint32_t v[kPerThread];
#pragma unroll
        for (int j = 0; j < (int)kPerThread; j++) v[j] = vn[j];
        if (base + kChunk < n) load8(in, base + kChunk + t8, n, vn);  // next step's symbols in flight
        const uint32_t nvalid = chunk_nvalid(base, n, t8);
        int wl;
        const int ex = wave_excl_last(thread_last_nz_rel(v, t8, nvalid), wl);
        if ((threadIdx.x & 63) == 0) shl[par][threadIdx.x >> 6] = wl;
        __syncthreads();
        int prel, ctot;
        wg_last_combine(ex, shl[par], prel, ctot);
        uint32_t zc[kPerThread];
        const uint32_t emit = thread_flags_rel(v, nvalid, rl, thread_g0(prel, t8, chunk_gc(base, carry)),
                                               chunk_lim(base, (long long)n) - (int)t8, zc);
#pragma unroll
        for (int j = 0; j < (int)kPerThread; j++) {
            if (!((emit >> j) & 1u)) continue;
            const uint32_t s = v[j], z = zc[j];
            cnt++;
            vmax = max(vmax, s);
            if (s == run_sym) {
                run_len++;
            } else {
                flush_run();
                run_sym = s;
                run_len = 1;
            }
            if (rl) {
                if (z == 0) zc0++;
                else atomicAdd(&hz[z], 1u);
            }
        }
        if (ctot >= 0) carry = (long long)base + ctot;
    }
    flush_run();
    zc0 = wave_sum(zc0);
    if (rl && (threadIdx.x & 63) == 0 && zc0) atomicAdd(&hz[0], zc0);
    uint32_t M = wg_sum_u32(cnt, shu);
    wide = wg_max_u32(wide, shu);
    vmax = wg_max_u32(vmax, shu);
    if (rl) {  // fillers of the trailing run are not emitted
        const uint64_t r = (uint64_t)((long long)n - 1 - carry);
        const uint32_t f = (uint32_t)(r / 256);
        if (threadIdx.x == 0 && f) {
            hs[0] -= f;
            hz[255] -= f;
        }
        M -= f;
        __syncthreads();
    }
    // maxima and distinct counts from the histograms
    uint32_t smax_l = 0, ks = 0, zmax_l = 0, kz = 0;
    if (!wide) {
        for (uint32_t i = threadIdx.x; i < kStatBins; i += kCoderThreads)
            if (hs[i]) { smax_l = i; ks++; }
    }
    for (uint32_t i = threadIdx.x; i < 256; i += kCoderThreads)
        if (hz[i]) { zmax_l = i; kz++; }
    const uint32_t smax = wide ? vmax : wg_max_u32(smax_l, shu);
    ks = wg_sum_u32(ks, shu);
    const uint32_t zmax = wg_max_u32(zmax_l, shu);
    kz = wg_sum_u32(kz, shu);

    SeqInfo* ss = &seqs[it.seq];
    SeqInfo* sz = rl ? &seqs[it.seq + 1] : nullptr;
    const uint64_t lim = 1ull << p.log2;
    if (threadIdx.x == 0) {
        ss->count = M;
        ss->smax = M ? smax : 0u;
        ss->nsym = wide ? 0u : ks;
        ss->nlen = 0;
        ss->err = kErrNone;
        ss->quirk = 0;
        ss->wide = wide;
        ss->item = ii;
        ss->bits = 0;
        ss->last_nz1 = (uint64_t)(carry + 1);  // last nonzero + 1, read by the pack pass
        ss->out_off = 0;
        if (p.dialect == CC_DIALECT_GPU && (uint64_t)ss->smax + 1 > lim) ss->err = kErrOverflow;
        sh_off[0] = ss->err ? ~uint64_t(0) : arena_alloc(st, hist_words(ss->smax + 1, ks, !wide), cap, ss);
        if (sz) {
            sz->count = M;
            sz->smax = M ? zmax : 0u;
            sz->nsym = kz;
            sz->nlen = 0;
            sz->err = kErrNone;
            sz->quirk = 0;
            sz->wide = 0;
            sz->item = ii;
            sz->bits = 0;
            sz->out_off = 0;
            sz->last_nz1 = 0;
            if (p.dialect == CC_DIALECT_GPU && (uint64_t)sz->smax + 1 > lim) sz->err = kErrOverflow;
            sh_off[1] = sz->err ? ~uint64_t(0) : arena_alloc(st, hist_words(sz->smax + 1, kz, true), cap, sz);
        }
    }
    __syncthreads();
    if (sh_off[0] != ~uint64_t(0)) {
        uint32_t* h = scratch + sh_off[0];
        const uint32_t nb = (M ? smax : 0u) + 1;
        for (uint32_t i = threadIdx.x; i < nb; i += kCoderThreads) h[i] = wide ? 0u : hs[i];
        if (threadIdx.x == 0) ss->hist_off = sh_off[0];
    }
    // the uniform branches below (sh_off is shared) keep every thread at the scans' barriers
    if (sh_off[0] != ~uint64_t(0) && !wide && ks <= design::kLaneMaxK) {
        const uint32_t nb = (M ? smax : 0u) + 1;
        write_leaves(hs, nb, reinterpret_cast<uint2*>(scratch + sh_off[0] + leaves_at(nb)), shq);
    }
    if (rl && sh_off[1] != ~uint64_t(0)) {
        uint32_t* h = scratch + sh_off[1];
        const uint32_t nb = (M ? zmax : 0u) + 1;
        for (uint32_t i = threadIdx.x; i < nb; i += kCoderThreads) h[i] = hz[i];
        if (threadIdx.x == 0) sz->hist_off = sh_off[1];
        if (kz <= design::kLaneMaxK)
            write_leaves(hz, nb, reinterpret_cast<uint2*>(scratch + sh_off[1] + leaves_at(nb)), shq);
    }
}

// Pass A for streams with symbols >= kStatBins: histogram with global atomics, then the pad
// quirk (Symbol16, GPU dialect, max symbol 65535, odd count: src/cudaCompress/Histogram.cu:90-105).
template <class S>
__global__ __launch_bounds__(kCoderThreads) void k_code_stats_wide(const CodeItem* items, SeqInfo* seqs,
                                                                   CoderParams p, uint32_t* scratch) {
    __shared__ long long shl[4];
    __shared__ uint32_t shu[4];
    const CodeItem it = items[blockIdx.x];
    SeqInfo* ss = &seqs[it.seq];
    if (!ss->wide || ss->err) return;
    const S* in = static_cast<const S*>(it.in);
    const bool rl = it.rl != 0;
    const uint32_t n = it.n;
    uint32_t* h = scratch + ss->hist_off;
    const long long last_nz = (long long)ss->last_nz1 - 1;
    long long carry = -1;
    for (uint64_t base = 0; base < n; base += kChunk) {
        const uint64_t p0 = base + (uint64_t)threadIdx.x * kPerThread;
        uint32_t v[kPerThread];
        load8(in, p0, n, v);
        long long chunk_max;
        long long prev = wg_excl_max_i64(thread_last_nz(v, p0, n), shl, chunk_max);
        prev = prev > carry ? prev : carry;
        thread_pairs(v, p0, n, rl, prev, rl ? last_nz : (long long)n,
                     [&](uint32_t s, uint32_t) { atomicAdd(&h[s], 1u); });
        carry = chunk_max > carry ? chunk_max : carry;
    }
    __syncthreads();
    const uint32_t nb = ss->smax + 1;
    const uint32_t quirk = (p.dialect == CC_DIALECT_GPU && sizeof(S) == 2 && ss->smax == 65535u && (ss->count & 1u)) ? 1u : 0u;
    uint32_t k = 0;
    for (uint32_t i = threadIdx.x; i < nb; i += kCoderThreads) {
        uint32_t c = atomicAdd(&h[i], (quirk && i == 65535u) ? 1u : 0u);  // device-coherent read
        if (quirk && i == 65535u) c++;
        k += c ? 1u : 0u;
    }
    k = wg_sum_u32(k, shu);
    if (threadIdx.x == 0) {
        ss->nsym = k;
        ss->quirk = quirk;
    }
}

// wave64 executor of the design code (one wave per workgroup)
// Six levels of a walk over a ballot on the scalar unit: j <- 2 j + bit j of mr, from j = 1;
// three scalar instructions a level (shift, bit test into SCC, add with carry).
__device__ __forceinline__ uint32_t walk6(uint64_t mr) {
    uint32_t j = 1, t;
    asm volatile(
        "s_lshl_b32 %1, %0, 1\n\ts_bitcmp1_b64 %2, %0\n\ts_addc_u32 %0, %1, 0\n\t"
        "s_lshl_b32 %1, %0, 1\n\ts_bitcmp1_b64 %2, %0\n\ts_addc_u32 %0, %1, 0\n\t"
        "s_lshl_b32 %1, %0, 1\n\ts_bitcmp1_b64 %2, %0\n\ts_addc_u32 %0, %1, 0\n\t"
        "s_lshl_b32 %1, %0, 1\n\ts_bitcmp1_b64 %2, %0\n\ts_addc_u32 %0, %1, 0\n\t"
        "s_lshl_b32 %1, %0, 1\n\ts_bitcmp1_b64 %2, %0\n\ts_addc_u32 %0, %1, 0\n\t"
        "s_lshl_b32 %1, %0, 1\n\ts_bitcmp1_b64 %2, %0\n\ts_addc_u32 %0, %1, 0"
        : "+s"(j), "=&s"(t)
        : "s"(mr)
        : "scc");
    return j;
}

struct WaveEx {
    __device__ int lane() const { return (int)threadIdx.x; }
    __device__ int lanes() const { return 64; }
    __device__ void barrier() const { __syncthreads(); }

    // make_by_levels: one lane per node of a level, levels bottom-up
    __device__ void heap_make(HeapEntry* h, int len) const {
        if (len < 2) return;
        const int last = (len - 2) / 2;
        int d = 0;
        while ((1 << (d + 1)) - 1 <= last) d++;
        for (; d >= 0; d--) {
            const int lo = (1 << d) - 1, hi = min((1 << (d + 1)) - 2, last);
            for (int x = lo + lane(); x <= hi; x += 64) design::adjust_heap(h, x, len, h[x], design::QueueLess());
            __syncthreads();
        }
    }
    // pop in the form of design::pop_by_walk: per round, lane j < 63 evaluates the child choice
    // of local node j of the six-level subtree below the round's root (one LDS read of both
    // children's counts); the walk over the ballot of choices runs on the scalar unit.  With
    // 1-based heap indices a node's children are 2i and 2i + 1, so the walk's deepest node e
    // carries the whole round's path: its node at round depth d is e >> (6 - d).  Lane k ends up
    // holding path node k (mypath) and path node k+1 (mynext).  Returns the new top.
    __device__ HeapEntry heap_pop(HeapEntry* h, int len) const {
        const HeapEntry v = h[len - 1];
        if (len <= 1) return v;
        const int n = len - 1, ln = lane();
        const int lim = (n - 1) / 2, lastint = (n - 2) / 2;
        // a round ends the pop when its deepest node has no two children, unless that node is the
        // last internal node of an even-sized heap (its lone left child then comes from one more
        // round, whose walk goes left at once)
        const int lim_end = (n & 1) ? lim : lim + 1;
        const int sd = ln < 63 ? 31 - __clz(ln + 1) : 0;  // depth of this lane's local node
        int mypath = 0, mynext = 0, base = 0, m;
        HeapEntry pn;
        uint32_t c1 = 1;  // 1-based index of the round's root
        for (;;) {
            const int g = (int)((c1 - 1) << sd) + ln;
            const bool inner = ln < 63 && g < lim;
            const int gr = inner ? g : 0;  // lanes without a node read the root's children unused
            const bool ge = !(h[2 * gr + 1].count < h[2 * gr + 2].count);
            const uint64_t mr = (__ballot(ge) & __ballot(inner)) << 1;  // bit j: choice at 1-based local node j
            const uint32_t j = walk6(mr);
            const uint32_t e = (c1 << 6) | (j & 63u);
            const int dk = ln - base;
            if (dk >= 1 && dk <= 6) mypath = (int)(e >> (6 - dk)) - 1;
            if (dk >= 0 && dk <= 5) mynext = (int)(e >> (5 - dk)) - 1;
            if ((int)e - 1 >= lim_end) {
                // every lane reads its successor's entry (no entry changes before the path is
                // complete), issued ahead of the path's length
                pn = h[mynext < n ? mynext : 0];
                // the path is real down to round depth d while its node at depth d - 1 has two
                // children (index < lim); indices grow along the path, so these depths are a prefix
                const int steps = (int)__popcll(__ballot(ln >= 1 && ln <= 6 && (int)(e >> ln) - 1 < lim));
                m = base + steps;
                if ((n & 1) == 0 && (int)(e >> (6 - steps)) - 1 == lastint) m++;
                break;
            }
            base += 6;
            c1 = e;
        }
        // lanes 0..m-1 hold their successor; v lands at i* = m - #{k in 1..m : v.count < count(p_k)}
        const int is = m - (int)__popcll(__ballot(ln < m && v.count < pn.count));
        if (ln <= is) h[mypath] = ln < is ? pn : v;
        if (is == 0) return v;
        return HeapEntry{(uint32_t)__builtin_amdgcn_readlane((int)pn.count, 0),
                         (uint32_t)__builtin_amdgcn_readlane((int)pn.node, 0)};
    }
    // push_by_chain: lane t holds ancestor t of the new position; returns the new top
    __device__ HeapEntry heap_push(HeapEntry* h, int pos, HeapEntry v, HeapEntry top) const {
        const int ln = lane();
        const int T = 31 - __clz(pos + 1);
        const int c = ((pos + 1) >> ln) - 1;
        const bool anc = ln >= 1 && ln <= T;
        const HeapEntry e = h[anc ? c : 0];
        const int j = (int)__popcll(__ballot(anc && v.count < e.count));
        if (ln >= 1 && ln <= j) h[((pos + 1) >> (ln - 1)) - 1] = e;
        if (ln == j) h[c] = v;
        return j == T ? v : top;
    }
    template <class P, class Sk>
    __device__ int compact(int lo, int hi, P pred, Sk sink) const {
        int total = 0;
        for (int base = lo; base < hi; base += 64) {
            const int i = base + lane();
            const bool f = i < hi && pred(i);
            const uint64_t m = __ballot(f);
            if (f) sink(i, total + (int)__popcll(m & lanes_below(lane())));
            total += (int)__popcll(m);
        }
        return total;
    }
    template <class P>
    __device__ int count(int lo, int hi, P pred) const {
        int total = 0;
        for (int base = lo; base < hi; base += 64) {
            const int i = base + lane();
            total += (int)__popcll(__ballot(i < hi && pred(i)));
        }
        return total;
    }
    __device__ int max_all(int v) const { return (int)wave_max((uint32_t)v); }
    __device__ uint64_t sum_all(uint64_t v) const { return wave_sum_u64(v); }
    template <class F>
    __device__ void stable_rank_by_depth(const uint32_t* el, int K, uint32_t* cpl, F f) const {
        // lane d-1 counts depth d (depths are 1..31)
        uint32_t mycount = 0;
        for (int base = 0; base < K; base += 64) {
            const int i = base + lane();
            const bool valid = i < K;
            const uint32_t key = valid ? (el[i] >> 24) : 0u;
            uint64_t rem = __ballot(valid);
            while (rem) {
                const int leader = __ffsll((unsigned long long)rem) - 1;
                const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
                const uint64_t m = __ballot(valid && key == d);
                if (lane() == (int)d - 1) mycount += (uint32_t)__popcll(m);
                rem &= ~m;
            }
        }
        __shared__ uint32_t fi[32], fc[32];
        if (lane() < 32) cpl[lane()] = lane() < 31 ? mycount : 0u;
        __syncthreads();
        if (lane() == 0) design::first_codes(cpl, 31, fc, fi);
        __syncthreads();
        uint32_t seen = 0;  // lane d-1: elements of depth d seen so far
        for (int base = 0; base < K; base += 64) {
            const int i = base + lane();
            const bool valid = i < K;
            const uint32_t e = valid ? el[i] : 0u;
            const uint32_t key = e >> 24;
            uint64_t rem = __ballot(valid);
            while (rem) {
                const int leader = __ffsll((unsigned long long)rem) - 1;
                const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)key, leader);
                const uint64_t m = __ballot(valid && key == d);
                const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)seen, (int)d - 1);
                if (valid && key == d) f(i, e, b + (uint32_t)__popcll(m & lanes_below(lane())), fi, fc);
                if (lane() == (int)d - 1) seen += (uint32_t)__popcll(m);
                rem &= ~m;
            }
        }
    }
};

// Design class of a sequence by distinct-symbol count: four LDS classes and one whose workspace
// lives in the arena.  Class lists: counters at classes[0, kDesignLists), list c at
// classes + kDesignListBase + c * nseq.
// Tables of at most lane_k symbols below symbol kStatBins (so that k_code_stats wrote their leaf
// list) whose sequence is shorter than 2^23 go to the per-thread class; the others to the wave
// classes by distinct-symbol count.
__device__ __forceinline__ uint32_t design_class(const SeqInfo& s, uint32_t lane_k) {
    const uint32_t K = s.nsym;
    if (K <= lane_k && s.smax < kStatBins && s.count < design::kLaneMaxCount - 1) return kLaneClass;
    return K <= 256 ? 0u : K <= 1024 ? 1u : K <= 2048 ? 2u : K <= kDesignMid ? 3u : 4u;
}

// One thread per sequence: appends every sequence still to be designed to its class list (one
// atomic per class per wave).
__global__ __launch_bounds__(256) void k_design_classify(const SeqInfo* seqs, uint32_t nseq, uint32_t* classes,
                                                         uint32_t lane_k) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    const bool on = i < nseq && !seqs[i].err;
    const uint32_t c = on ? design_class(seqs[i], lane_k) : 0xFFFFFFFFu;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (uint32_t k = 0; k < kDesignLists; k++) {
        const uint64_t m = __ballot(c == k);
        if (!m) continue;
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(&classes[k], (uint32_t)__popcll(m));
        base = (uint32_t)__builtin_amdgcn_readlane((int)base, leader);
        if (c == k) classes[kDesignListBase + (size_t)k * nseq + base + __popcll(m & lanes_below(lane))] = i;
    }
}

// Designs the table of one sequence with one wave (all lanes call; returns uniformly).  Gws: the
// design workspace comes from the arena instead of LDS (a compile-time choice, so that the LDS
// classes address their workspace with LDS instructions rather than generic ones).
template <class Idx, bool Gws>
__device__ void design_one(SeqInfo* s, const CoderParams& p, uint32_t* scratch, uint64_t cap, CoderState* st,
                           unsigned char* smem) {
    constexpr bool gws = Gws;
    __shared__ Ctl ctl;
    __shared__ uint64_t offs[4];
    __shared__ uint64_t bits_sh;
    __shared__ uint32_t info_sh[3];
    __shared__ uint32_t cpl_sh[32];
    if (s->err) return;
    const uint32_t K = s->nsym;
    const uint32_t nb = s->smax + 1;
    const uint64_t hreg = design::heap_region_bytes<Idx>(K);
    const uint64_t preg = ((uint64_t)sizeof(Idx) * (2ull * K + 2) + 15) & ~uint64_t(15);
    if (threadIdx.x == 0) {
        offs[0] = arena_alloc(st, 2ull * nb, cap, s);                   // code, len
        offs[1] = offs[0] == ~uint64_t(0) ? offs[0] : arena_alloc(st, K ? K : 1, cap, s);  // symbol list
        offs[2] = 0;
        if (gws && offs[1] != ~uint64_t(0))
            offs[2] = arena_alloc(st, (hreg + preg) / 4 + 3 * design::kStackRanges + 4, cap, s);
    }
    __syncthreads();
    if (offs[0] == ~uint64_t(0) || offs[1] == ~uint64_t(0) || offs[2] == ~uint64_t(0)) return;
    unsigned char* ws;
    if constexpr (Gws) ws = reinterpret_cast<unsigned char*>(scratch + offs[2]);
    else ws = smem;
    Work<Idx> w{reinterpret_cast<HeapEntry*>(ws), reinterpret_cast<Idx*>(ws + hreg),
                reinterpret_cast<int*>(ws + hreg + preg)};
    uint32_t info[3];
    TableOut out{scratch + offs[0], scratch + offs[0] + nb, scratch + offs[1], cpl_sh, info_sh};
    WaveEx ex;
    uint64_t bits = 0;
    design::design_table(ex, scratch + s->hist_off, nb, (int)p.dialect, s->quirk, w, &ctl, out, &bits);
    if (threadIdx.x == 0) bits_sh = bits;
    __syncthreads();
    info[0] = info_sh[0];
    info[1] = info_sh[1];
    info[2] = info_sh[2];
    if (threadIdx.x < 32) s->cpl[threadIdx.x] = threadIdx.x < info[0] ? cpl_sh[threadIdx.x] : 0u;
    if (threadIdx.x == 0) {
        s->code_off = offs[0];
        s->syms_off = offs[1];
        s->nlen = info[0];
        s->bits = bits_sh;
        if (info[2]) {
            s->err = info[2];
            atomicCAS(&st->err, 0u, info[2]);
        } else if (p.dialect == CC_DIALECT_CPU && K >= (1u << 14)) {
            s->err = kErrOverflow;  // not describable by the 14-bit CPU-dialect header
            atomicCAS(&st->err, 0u, (uint32_t)kErrOverflow);
        } else if (bits_sh >= (1ull << 32)) {
            s->err = kErrOutput;
            atomicCAS(&st->err, 0u, (uint32_t)kErrOutput);
        } else {
            // the coding-block offsets (the codewords go straight to the output)
            const uint64_t nblk = (s->count + p.cbs - 1) / p.cbs;
            s->boff_off = arena_alloc(st, nblk ? nblk : 1, cap, s);
        }
    }
}

// The tables of one class list, one wave each, the workgroups striding over the list.
template <class Idx, bool Gws>
__global__ __launch_bounds__(64) void k_design(SeqInfo* seqs, CoderParams p, uint32_t* scratch, uint64_t cap,
                                               CoderState* st, const uint32_t* classes, uint32_t nseq, uint32_t cls) {
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t count = classes[cls];
    const uint32_t* list = classes + kDesignListBase + (size_t)cls * nseq;
    // the largest tables set the length of the design phase: their waves take the SIMD's issue
    // slots first when they share it with the many small designs running beside them
    if (cls >= 2 && blockIdx.x < count) __builtin_amdgcn_s_setprio(2);
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        design_one<Idx, Gws>(&seqs[list[i]], p, scratch, cap, st, smem);
        __syncthreads();  // LDS state is reused by the next table
    }
}

// Arena words of one allocation as arena_alloc rounds them.
__device__ __forceinline__ uint64_t arena_words(uint64_t w) { return (w + 15) & ~uint64_t(15); }

// Arena ranges for the threads of a workgroup with one atomic: every thread calls with its own
// word count (0: none) and gets its offset; a range past the capacity is the caller's to reject.
template <int T>
__device__ uint64_t wg_arena_alloc(CoderState* st, uint64_t words, uint64_t* sh) {
    sh[threadIdx.x] = words;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t acc = 0;
        for (int k = 0; k < T; k++) {
            const uint64_t w = sh[k];
            sh[k] = acc;
            acc += w;
        }
        const uint64_t base = acc ? atomicAdd(&st->scratch_used, (unsigned long long)acc) : 0ull;
        for (int k = 0; k < T; k++) sh[k] += base;
    }
    __syncthreads();
    const uint64_t off = sh[threadIdx.x];
    __syncthreads();
    return off;
}

// The per-thread class: one table per thread (design::design_table_lane), kLaneThreads tables per
// workgroup, the heaps and parent links interleaved across the threads in LDS (8 LaneK bytes per
// table: tables of at most LaneK symbols).
constexpr int kLaneThreads = 32;
template <int LaneK>
__global__ __launch_bounds__(kLaneThreads) void k_design_lanes(SeqInfo* seqs, CoderParams p, uint32_t* scratch,
                                                               uint64_t cap, CoderState* st, const uint32_t* classes,
                                                               uint32_t nseq) {
    static_assert(LaneK <= (int)design::kLaneMaxK, "packed node indices");
    constexpr int kStack = design::lane_stack_ints(LaneK);
    __shared__ uint32_t heap_sh[LaneK * kLaneThreads];
    __shared__ uint16_t par_sh[2 * LaneK * kLaneThreads];
    __shared__ int stack_sh[kStack * kLaneThreads];
    __shared__ uint64_t alloc_sh[kLaneThreads];
    const uint32_t count = classes[kLaneClass];
    const uint32_t* list = classes + kDesignListBase + (size_t)kLaneClass * nseq;
    const int t = threadIdx.x;
    for (uint32_t i0 = blockIdx.x * kLaneThreads; i0 < count; i0 += gridDim.x * kLaneThreads) {
        const uint32_t i = i0 + t;
        SeqInfo* s = i < count ? &seqs[list[i]] : nullptr;
        const bool on = s != nullptr && s->err == 0;
        const uint32_t K = on ? s->nsym : 0u, nb = on ? s->smax + 1 : 0u;
        // code and len (2 nb words), then the symbol list (K words), as design_one allocates them
        const uint64_t w0 = on ? arena_words(2ull * nb) + arena_words(K ? K : 1) : 0ull;
        const uint64_t off0 = wg_arena_alloc<kLaneThreads>(st, w0, alloc_sh);
        bool ok = on;
        if (on && off0 + w0 > cap) {
            ok = false;
            s->err = kErrScratch;
            atomicCAS(&st->err, 0u, (uint32_t)kErrScratch);
        }
        uint64_t bits = 0;
        uint32_t info[3] = {0u, 0u, 0u};
        if (ok) {
            const uint64_t code_off = off0, syms_off = off0 + arena_words(2ull * nb);
            // the leaf list k_code_stats wrote after the histogram
            const uint2* leaves = reinterpret_cast<const uint2*>(scratch + s->hist_off + leaves_at(nb));
            TableOut out{scratch + code_off, scratch + code_off + nb, scratch + syms_off, s->cpl, info};
            design::design_table_lane([&](int r) { return leaves[r]; }, (int)K, (int)p.dialect, s->quirk,
                                      design::Strided<uint32_t>{heap_sh + t, kLaneThreads},
                                      design::Strided<uint16_t>{par_sh + t, kLaneThreads},
                                      design::Strided<int>{stack_sh + t, kLaneThreads}, out, &bits);
            s->code_off = code_off;
            s->syms_off = syms_off;
            s->nlen = info[0];
            s->bits = bits;
            if (info[2]) {
                ok = false;
                s->err = info[2];
                atomicCAS(&st->err, 0u, info[2]);
            } else if (bits >= (1ull << 32)) {
                ok = false;
                s->err = kErrOutput;
                atomicCAS(&st->err, 0u, (uint32_t)kErrOutput);
            }
        }
        // the coding-block offsets (the codewords go straight to the output)
        const uint64_t nblk = ok ? (s->count + p.cbs - 1) / p.cbs : 0ull;
        const uint64_t w1 = ok ? arena_words(nblk ? nblk : 1) : 0ull;
        const uint64_t off1 = wg_arena_alloc<kLaneThreads>(st, w1, alloc_sh);
        if (ok) {
            if (off1 + w1 > cap) {
                s->err = kErrScratch;
                atomicCAS(&st->err, 0u, (uint32_t)kErrScratch);
            } else {
                s->boff_off = off1;
            }
        }
    }
}

__device__ __forceinline__ uint64_t table_bits(const SeqInfo& s, const CoderParams& p) {
    const uint32_t sb = req_bits(s.smax);
    return 5ull + (uint64_t)p.count_bits * (s.nlen + 1) + p.symbits_bits + (uint64_t)sb * s.nsym;
}
__device__ __forceinline__ uint64_t seq_words(const SeqInfo& s, const CoderParams& p) {
    const uint64_t nblk = (s.count + p.cbs - 1) / p.cbs;
    return (table_bits(s, p) + 31) / 32 + 1 + (s.bits + 31) / 32 + (nblk + 1) / 2;
}

// 32 header bits starting at bit `pos` of a table header: [nlen:5][cpl x nlen][nsym][sb][symbols].
__device__ uint32_t header_word(const SeqInfo& s, const uint32_t* cpl, const CoderParams& p, const uint32_t* syms,
                                uint64_t pos) {
    const uint32_t cb = p.count_bits, sbf = p.symbits_bits;
    const uint32_t sb = req_bits(s.smax);
    const uint64_t fixed = 5ull + (uint64_t)cb * (s.nlen + 1) + sbf;
    uint32_t out = 0;
    auto put = [&](uint64_t fpos, uint32_t width, uint32_t val) {  // field at fpos..fpos+width
        if (width == 0 || fpos + width <= pos || fpos >= pos + 32) return;
        if (width < 32) val &= (1u << width) - 1u;
        // align the field's MSB to bit position (fpos - pos) counted from the word's MSB
        const long long sh = (long long)(pos + 32) - (long long)(fpos + width);  // left shift amount
        if (sh >= 0) out |= (sh >= 32) ? 0u : (val << sh);
        else out |= (-sh >= 32) ? 0u : (val >> (-sh));
    };
    if (pos < fixed) {
        put(0, 5, s.nlen);
        for (uint32_t l = 0; l < s.nlen; l++) put(5 + (uint64_t)cb * l, cb, cpl[l]);
        put(5 + (uint64_t)cb * s.nlen, cb, s.nsym);
        put(5 + (uint64_t)cb * (s.nlen + 1), sbf, sb);
    }
    if (sb && s.nsym) {
        const uint64_t lo = pos > fixed ? (pos - fixed) / sb : 0;
        for (uint64_t k = lo; k < s.nsym; k++) {
            const uint64_t fp = fixed + k * sb;
            if (fp >= pos + 32) break;
            put(fp, sb, syms[k]);
        }
    }
    return out;
}

// Pass B: pairs recomputed, codewords packed into an LDS window and flushed as whole words
// straight to their place in the output section (laid out by k_layout_* from the designed bit
// counts): each workgroup also writes its item's zero-run count, table headers and bit counts, and
// after the pairs its packInc16 coding-block offsets (src/cudaCompress/PackInc.cu:210-220).  The
// absolute bit offset of every coding block is kept in the arena for the increments.
template <class S>
__global__ __launch_bounds__(kCoderThreads) void k_code_pack(const CodeItem* items, SeqInfo* seqs, CoderParams p,
                                                             uint32_t* scratch, const CoderState* st, uint32_t* out,
                                                             uint64_t out_cap, const uint32_t* order) {
    __shared__ uint32_t cs[kStatBins];
    __shared__ unsigned char lsh[kStatBins];
    __shared__ uint32_t cz[256];
    __shared__ unsigned char lz[256];
    __shared__ uint32_t win[2][kWindowWords];
    __shared__ int shl[4];
    __shared__ uint64_t shq[4];
    if (st->out_base + st->out_words > out_cap) return;  // the host reports the short output
    const CodeItem it = items[order ? order[blockIdx.x] : blockIdx.x];
    const SeqInfo ss = seqs[it.seq];
    const bool rl = it.rl != 0;
    if (ss.err || (rl && seqs[it.seq + 1].err)) return;
    const SeqInfo sz = rl ? seqs[it.seq + 1] : ss;
    const uint32_t M = ss.count;
    // section heads: [M] (RLE items), then per sequence its table header and bit count
    const uint32_t tws = (uint32_t)((table_bits(ss, p) + 31) / 32), twz = (uint32_t)((table_bits(sz, p) + 31) / 32);
    {
        const uint32_t* syms_s = scratch + ss.syms_off;
        uint32_t* hs = out + ss.out_off;
        // the header from the global record: a local record whose address is taken lives in
        // scratch memory, and so would every field the pack reads from it
        const SeqInfo& gs = seqs[it.seq];
        for (uint32_t w = threadIdx.x; w < tws; w += kCoderThreads) hs[w] = header_word(gs, gs.cpl, p, syms_s, 32ull * w);
        if (threadIdx.x == 0) {
            hs[tws] = (uint32_t)ss.bits;
            if (rl) out[ss.out_off - 1] = M;  // aligned u32 M
        }
        if (rl) {
            const uint32_t* syms_z = scratch + sz.syms_off;
            uint32_t* hz = out + sz.out_off;
            const SeqInfo& gz = seqs[it.seq + 1];
            for (uint32_t w = threadIdx.x; w < twz; w += kCoderThreads)
                hz[w] = header_word(gz, gz.cpl, p, syms_z, 32ull * w);
            if (threadIdx.x == 0) hz[twz] = (uint32_t)sz.bits;
        }
    }
    const uint64_t nws = (ss.bits + 31) / 32, nwz = (sz.bits + 31) / 32;  // codeword words of each
    uint32_t* ws = out + ss.out_off + tws + 1;
    uint32_t* wz = out + sz.out_off + twz + 1;
    uint32_t* bs = scratch + ss.boff_off;
    uint32_t* bz = scratch + sz.boff_off;
    const uint32_t cbs = p.cbs;
    if (M > 0) {
        const S* in = static_cast<const S*>(it.in);
        const uint32_t n = it.n;
        const bool sglob = ss.smax >= kStatBins;
        const uint32_t* gcode = scratch + ss.code_off;
        const uint32_t* glen = gcode + ss.smax + 1;
        // tables to LDS, each thread's loads issued four entries at a time
        auto stage_table = [&](const uint32_t* code, const uint32_t* len, uint32_t nb, uint32_t* dc, unsigned char* dl) {
            for (uint32_t i0 = threadIdx.x; i0 < nb; i0 += 4 * kCoderThreads) {
                uint32_t c[4], l[4];
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t i = i0 + k * kCoderThreads;
                    c[k] = i < nb ? code[i] : 0u;
                    l[k] = i < nb ? len[i] : 0u;
                }
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const uint32_t i = i0 + k * kCoderThreads;
                    if (i < nb) {
                        dc[i] = c[k];
                        dl[i] = (unsigned char)l[k];
                    }
                }
            }
        };
        if (!sglob) stage_table(gcode, glen, ss.smax + 1, cs, lsh);
        if (rl) {
            const uint32_t* zc = scratch + sz.code_off;
            stage_table(zc, zc + sz.smax + 1, sz.smax + 1, cz, lz);
        }
        for (uint32_t i = threadIdx.x; i < 2 * kWindowWords; i += kCoderThreads) (&win[0][0])[i] = 0;
        const bool single = ss.nlen == 0;  // one distinct symbol: no codeword bits
        const bool zsingle = rl && sz.nlen == 0;
        const long long limit = rl ? (long long)ss.last_nz1 - 1 : (long long)n;  // last nonzero
        constexpr uint32_t WM = kWindowWords - 1;
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        const uint32_t t8 = threadIdx.x * kPerThread;
        // One barrier per chunk besides the one before the flush: the exclusive last nonzeros of
        // the next chunk (zero-run lengths) ride on this chunk's pair/bit-count scan, so the
        // symbols run two chunks ahead (v: this chunk, vn: the next, vnn: in flight).
        uint32_t v[kPerThread], vn[kPerThread];
        load8(in, (uint64_t)t8, n, v);
        load8(in, (uint64_t)t8 + kChunk, n, vn);
        int prel, ctot;  // this chunk: the exclusive chunk-relative last nonzero, the chunk's last
        {
            int wl;
            const int ex = wave_excl_last(thread_last_nz_rel(v, t8, chunk_nvalid(0, n, t8)), wl);
            if (lane == 0) shl[wv] = wl;
            __syncthreads();
            wg_last_combine(ex, shl, prel, ctot);
            __syncthreads();  // shl is written again by the first chunk's scan
        }
        // the chunk loop specialised on where the code table lives (LDS, or global memory for
        // symbols >= kStatBins)
        auto chunks = [&](auto sg) {
            constexpr bool SG = decltype(sg)::value;
            long long carry = -1;  // last nonzero before this chunk
            uint64_t pairs = 0, cbits_s = 0, cbits_z = 0;
            for (uint64_t base = 0; base < n && pairs < M; base += kChunk) {
                uint32_t vnn[kPerThread];
                if (base + 2 * kChunk < n) load8(in, base + 2 * kChunk + t8, n, vnn);
                const uint32_t nvalid = chunk_nvalid(base, n, t8);
                // this thread's pairs; code lengths looked up once for both the scan and the packing
                uint32_t zc[kPerThread], ls[kPerThread], lzs[kPerThread];
                const uint32_t emit = thread_flags_rel(v, nvalid, rl, thread_g0(prel, t8, chunk_gc(base, carry)),
                                                       chunk_lim(base, limit) - (int)t8, zc);
                uint32_t nbs = 0, nbz = 0;
                // table reads issued together, unconditionally (symbols and zero counts index inside
                // the tables; a length is then zeroed where no pair is emitted)
#pragma unroll
                for (int j = 0; j < (int)kPerThread; j++) {
                    if (SG) ls[j] = (((emit >> j) & 1u) && !single) ? glen[v[j]] : 0u;
                    else ls[j] = lsh[v[j]];
                    lzs[j] = lz[zc[j]];
                }
#pragma unroll
                for (int j = 0; j < (int)kPerThread; j++) {
                    const bool e = (emit >> j) & 1u;
                    ls[j] = (e && !single) ? ls[j] : 0u;
                    lzs[j] = (e && rl && !zsingle) ? lzs[j] : 0u;
                    nbs += ls[j];
                    nbz += lzs[j];
                }
                const uint32_t np = __popc(emit);
                // pair/bit-count scan, and the next chunk's last nonzeros, over one barrier
                const bool more = base + kChunk < n;
                int wl;
                const int exn =
                    wave_excl_last(more ? thread_last_nz_rel(vn, t8, chunk_nvalid(base + kChunk, n, t8)) : -1, wl);
                const uint64_t a = ((uint64_t)np << 40) | ((uint64_t)nbs << 20) | nbz;
                const uint64_t xa = wave_incl_scan_u64(a);
                if (lane == 63) shq[wv] = xa;
                if (lane == 0) shl[wv] = wl;
                __syncthreads();
                uint64_t ba = 0, ta = 0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const uint64_t sa = shq[i];
                    if (i < wv) ba += sa;
                    ta += sa;
                }
                const uint64_t ex = ba + xa - a;
                const uint64_t tot = uniform_u64(ta);
                int prel_n, ctot_n;
                wg_last_combine(exn, shl, prel_n, ctot_n);
                const uint32_t os0 = (uint32_t)cbits_s + (uint32_t)((ex >> 20) & 0xFFFFF);  // mod 2^32
                const uint32_t oz0 = (uint32_t)cbits_z + (uint32_t)(ex & 0xFFFFF);
                if (np) {
                    // the coding-block start among this thread's pairs: a thread emits at most
                    // kPerThread < cbs pairs (cbs a power of two), so at most one, its k-th pair (pair
                    // indices < M only: stats and pack agree on M, the bound keeps the offset writes in
                    // their arrays regardless); its bit offsets add the lengths of the k pairs before it
                    const uint32_t pi0 = (uint32_t)pairs + (uint32_t)(ex >> 40);
                    const uint32_t k = (cbs - (pi0 & (cbs - 1))) & (cbs - 1);
                    if (k < np && pi0 + k < M) {
                        uint32_t os = os0, oz = oz0, e = 0;
#pragma unroll
                        for (int j = 0; j < (int)kPerThread; j++) {
                            const bool before = ((emit >> j) & 1u) && e < k;
                            os += before ? ls[j] : 0u;
                            oz += before ? lzs[j] : 0u;
                            e += (emit >> j) & 1u;
                        }
                        const uint32_t blk = (pi0 + k) >> (31 - __clz(cbs));
                        bs[blk] = os;
                        if (rl) bz[blk] = oz;
                    }
                    // codewords through a 64-bit accumulator, one LDS OR per window word touched
                    // (ring slot: output word mod kWindowWords)
                    if (!single) {
                        uint32_t wi = os0 >> 5, nacc = os0 & 31;
                        uint64_t acc = 0;
                        uint32_t cw[kPerThread];
#pragma unroll
                        for (int j = 0; j < (int)kPerThread; j++) {
                            if (SG) cw[j] = ls[j] ? gcode[v[j]] : 0u;
                            else cw[j] = cs[v[j]];
                        }
#pragma unroll
                        for (int j = 0; j < (int)kPerThread; j++) {
                            const uint32_t l = ls[j];
                            const uint32_t c = l ? cw[j] : 0u;
                            acc = (acc << l) | c;
                            nacc += l;
                            if (nacc >= 32) {
                                nacc -= 32;
                                atomicOr(&win[0][wi & WM], (uint32_t)(acc >> nacc));
                                wi++;
                            }
                        }
                        if (nacc) atomicOr(&win[0][wi & WM], (uint32_t)(acc << (32 - nacc)));
                    }
                    if (rl && !zsingle) {
                        uint32_t wi = oz0 >> 5, nacc = oz0 & 31;
                        uint64_t acc = 0;
                        uint32_t cw[kPerThread];
#pragma unroll
                        for (int j = 0; j < (int)kPerThread; j++) cw[j] = cz[zc[j]];
#pragma unroll
                        for (int j = 0; j < (int)kPerThread; j++) {
                            const uint32_t l = lzs[j];
                            const uint32_t c = l ? cw[j] : 0u;
                            acc = (acc << l) | c;
                            nacc += l;
                            if (nacc >= 32) {
                                nacc -= 32;
                                atomicOr(&win[1][wi & WM], (uint32_t)(acc >> nacc));
                                wi++;
                            }
                        }
                        if (nacc) atomicOr(&win[1][wi & WM], (uint32_t)(acc << (32 - nacc)));
                    }
                }
                __syncthreads();
                pairs += tot >> 40;
                const uint64_t nend_s = cbits_s + ((tot >> 20) & 0xFFFFF), nend_z = cbits_z + (tot & 0xFFFFF);
                // flush the whole words and clear their slots; the partial last word stays in the ring
                // (the next chunk's barrier orders these clears before its ORs)
                const uint64_t wbs = cbits_s >> 5, wbz = cbits_z >> 5;
                const uint32_t full_s = (uint32_t)((nend_s >> 5) - wbs), full_z = (uint32_t)((nend_z >> 5) - wbz);
                for (uint32_t i = threadIdx.x; i < full_s; i += kCoderThreads) {
                    const uint32_t r = (uint32_t)(wbs + i) & WM;
                    if (wbs + i < nws) ws[wbs + i] = win[0][r];
                    win[0][r] = 0;
                }
                if (rl)
                    for (uint32_t i = threadIdx.x; i < full_z; i += kCoderThreads) {
                        const uint32_t r = (uint32_t)(wbz + i) & WM;
                        if (wbz + i < nwz) wz[wbz + i] = win[1][r];
                        win[1][r] = 0;
                    }
                cbits_s = nend_s;
                cbits_z = nend_z;
                if (ctot >= 0) carry = (long long)base + ctot;
                prel = prel_n;
                ctot = ctot_n;
#pragma unroll
                for (int j = 0; j < (int)kPerThread; j++) {
                    v[j] = vn[j];
                    vn[j] = vnn[j];
                }
            }
            if (threadIdx.x == 0) {  // the partial last words (their ORs precede the last barrier)
                if ((cbits_s & 31) && (cbits_s >> 5) < nws) ws[cbits_s >> 5] = win[0][(uint32_t)(cbits_s >> 5) & WM];
                if (rl && (cbits_z & 31) && (cbits_z >> 5) < nwz)
                    wz[cbits_z >> 5] = win[1][(uint32_t)(cbits_z >> 5) & WM];
            }
        };
        if (sglob) chunks(std::true_type{});
        else chunks(std::false_type{});
    }
    // coding-block offsets as packInc16 increments, from the absolute offsets the pairs recorded
    __threadfence_block();
    __syncthreads();
    for (int q = 0; q < (rl ? 2 : 1); q++) {
        const SeqInfo& s = q ? sz : ss;
        const uint32_t* bo = q ? bz : bs;
        uint32_t* o = out + s.out_off + (q ? twz : tws) + 1 + (q ? nwz : nws);
        const uint32_t nblk = (s.count + cbs - 1) / cbs;
        for (uint32_t w = threadIdx.x; w < (nblk + 1) / 2; w += kCoderThreads) {
            const uint32_t b0 = 2 * w, b1 = 2 * w + 1;
            const uint32_t i0 = b0 == 0 ? bo[0] : bo[b0] - bo[b0 - 1];
            const uint32_t i1 = b1 < nblk ? bo[b1] - bo[b1 - 1] : 0u;
            o[w] = (i0 & 0xFFFFu) | (i1 << 16);
        }
    }
}

// Section offsets in call order, in three launches: the words of every item with each
// 1024-item block's total, then every block's offsets (its base is the sum of the totals before
// it), then each call's range and status.
__device__ __forceinline__ uint64_t item_words(const CodeItem& it, const SeqInfo* seqs, const CoderParams& p) {
    const SeqInfo& a = seqs[it.seq];
    return it.rl ? 1 + seq_words(a, p) + seq_words(seqs[it.seq + 1], p) : seq_words(a, p);
}

// Exclusive scan over the 1024 threads of a workgroup (16 waves); *tot: the workgroup's sum.
__device__ __forceinline__ uint64_t wg1024_excl_scan(uint64_t x, uint64_t* sh, uint64_t& tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t inc = wave_incl_scan_u64(x);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    uint64_t b = 0;
    tot = 0;
    for (int k = 0; k < 16; k++) {
        if (k < w) b += sh[k];
        tot += sh[k];
    }
    __syncthreads();
    return b + inc - x;
}

// Pass 1: item words (kept in item_off for pass 2) and block totals.
__global__ __launch_bounds__(1024) void k_layout_sizes(const CodeItem* items, uint32_t nitems, const SeqInfo* seqs,
                                                       CoderParams p, uint64_t* item_off, uint64_t* bsum) {
    __shared__ uint64_t sh[16];
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    uint64_t words = 0;
    if (i < nitems) {
        words = item_words(items[i], seqs, p);
        item_off[i] = words;
    }
    uint64_t tot;
    (void)wg1024_excl_scan(words, sh, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Pass 2: offsets.  The job's output starts after the previous job's (prev) when chained.
__global__ __launch_bounds__(1024) void k_layout_offsets(const CodeItem* items, uint32_t nitems, SeqInfo* seqs,
                                                         CoderParams p, CoderState* st, uint64_t* item_off,
                                                         const uint64_t* bsum, uint64_t out_cap,
                                                         const CoderState* prev) {
    __shared__ uint64_t sh[16];
    const uint64_t base0 = prev ? prev->out_base + prev->out_words : 0;
    // sum of the totals of the blocks before this one
    uint64_t before = 0;
    for (uint32_t k = threadIdx.x; k < blockIdx.x; k += 1024) before += bsum[k];
    uint64_t sum_before;
    (void)wg1024_excl_scan(before, sh, sum_before);
    const uint32_t i = blockIdx.x * 1024 + threadIdx.x;
    const uint64_t words = i < nitems ? item_off[i] : 0;
    uint64_t tot;
    const uint64_t off = base0 + sum_before + wg1024_excl_scan(words, sh, tot);
    if (i < nitems) {
        const CodeItem it = items[i];
        item_off[i] = off;
        SeqInfo& a = seqs[it.seq];
        a.out_off = off + (it.rl ? 1 : 0);
        if (it.rl) seqs[it.seq + 1].out_off = a.out_off + seq_words(a, p);
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        const uint64_t end = base0 + sum_before + tot;
        item_off[nitems] = end;
        st->out_base = base0;
        st->out_words = end - base0;
        if (end > out_cap) atomicCAS(&st->err, 0u, (uint32_t)kErrOutput);
    }
}

// Pass 3: one thread per call.
__global__ __launch_bounds__(256) void k_layout_calls(const CodeItem* items, uint32_t nitems, const SeqInfo* seqs,
                                                      CallOut* calls, uint32_t ncalls, const uint32_t* call_first,
                                                      const uint64_t* item_off, uint64_t out_cap) {
    const uint32_t c = blockIdx.x * 256 + threadIdx.x;
    if (c >= ncalls) return;
    const uint32_t f = call_first[c], l = call_first[c + 1];
    uint32_t err = 0;
    for (uint32_t i = f; i < l && !err; i++) {
        const CodeItem it = items[i];
        err = seqs[it.seq].err;
        if (!err && it.rl) err = seqs[it.seq + 1].err;
    }
    calls[c].word_begin = item_off[f];
    calls[c].words = item_off[l] - item_off[f];
    calls[c].err = err ? err : (item_off[nitems] > out_cap ? (uint32_t)kErrOutput : 0u);
    calls[c].pad = 0;
}

// Concatenation of separately laid out jobs (the groups of a plan): group blockIdx.y's words
// [0, n_g) land at the sum of the earlier groups' word counts; every group's count is its job
// state's out_base + out_words.  Words past dst_cap are not written.
__global__ __launch_bounds__(256) void k_concat_groups(const ConcatGroup* groups, uint32_t* dst, uint64_t dst_cap) {
    const uint32_t g = blockIdx.y;
    uint64_t prefix = 0;
    for (uint32_t h = 0; h < g; h++) prefix += groups[h].st->out_base + groups[h].st->out_words;
    const uint64_t n = groups[g].st->out_base + groups[g].st->out_words;
    const uint64_t end = prefix + n < dst_cap ? n : (dst_cap > prefix ? dst_cap - prefix : 0);
    const uint32_t* src = groups[g].words;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < end; i += (uint64_t)gridDim.x * 256)
        dst[prefix + i] = src[i];
}

}  // namespace

void coder_concat(const ConcatGroup* d_groups, uint32_t ngroups, uint32_t* dst, uint64_t dst_cap, uint64_t bytes,
                  hipStream_t st) {
    if (ngroups == 0) return;
    CCA_LAUNCH("code_concat", bytes, k_concat_groups, dim3(1024, ngroups), dim3(256), 0, st, d_groups, dst, dst_cap);
    CCA_LAUNCH_CHECK();
}

void coder_encode(const CodeItem* d_items, uint32_t nitems, SeqInfo* d_seqs, uint32_t nseq, const CoderParams& p,
                  int symSize, uint32_t* scratch, uint64_t scratch_cap, CoderState* d_state, uint32_t* d_classes,
                  CallOut* d_calls,
                  uint32_t ncalls, const uint32_t* d_call_first_item, uint32_t* d_out, uint64_t out_cap,
                  uint64_t total_symbols, hipStream_t st, const SideStreams* side, const LayoutLink* link,
                  const uint32_t* design_grid, uint32_t lane_k, const uint32_t* d_item_order) {
    CCA_REQUIRE(symSize == 2 || symSize == 4, CC_ERR_INVALID_ARG, "symbol size must be 2 or 4");
    uint64_t* item_off = reinterpret_cast<uint64_t*>(scratch + scratch_cap);  // caller reserves the tail
    CCA_LAUNCH("coder_init", 32, k_coder_init, dim3(1), dim3(64), 0, st, d_state, d_classes);
    CCA_LAUNCH_CHECK();
    if (nitems == 0) return;
    const uint64_t inb = total_symbols * symSize;
    if (symSize == 2) {
        CCA_LAUNCH("code_stats", inb, k_code_stats<uint16_t>, dim3(nitems), dim3(kCoderThreads), 0, st, d_items, d_seqs,
                   p, scratch, scratch_cap, d_state, d_item_order);
        CCA_LAUNCH_CHECK();
        CCA_LAUNCH("code_stats_wide", 0, k_code_stats_wide<uint16_t>, dim3(nitems), dim3(kCoderThreads), 0, st,
                   d_items, d_seqs, p, scratch);
    } else {
        CCA_LAUNCH("code_stats", inb, k_code_stats<uint32_t>, dim3(nitems), dim3(kCoderThreads), 0, st, d_items, d_seqs,
                   p, scratch, scratch_cap, d_state, d_item_order);
        CCA_LAUNCH_CHECK();
        CCA_LAUNCH("code_stats_wide", 0, k_code_stats_wide<uint32_t>, dim3(nitems), dim3(kCoderThreads), 0, st,
                   d_items, d_seqs, p, scratch);
    }
    CCA_LAUNCH_CHECK();
    // Design classes by distinct-symbol count (small LDS footprints keep many tables resident):
    // the sequences are sorted into per-class lists, and each class runs over its list only, the
    // classes concurrently on side streams, the largest tables' classes launched first.
    CCA_LAUNCH("design_classify", 0, k_design_classify, dim3(div_up(nseq, 256u)), dim3(256), 0, st, d_seqs, nseq,
               d_classes, lane_k);
    CCA_LAUNCH_CHECK();
    constexpr int kClasses = 4;
    static_assert(kClasses <= kSideStreams, "one side stream per design class");
    const uint32_t kmax[kClasses] = {256, 1024, 2048, kDesignMid};
    uint32_t grid[kClasses] = {4096, 2048, 1536, 768};  // resident workgroups per class (LDS-bound)
    if (design_grid)
        for (int c = 0; c < kClasses; c++) grid[c] = std::max(1u, design_grid[c]);
    const char* names[kClasses] = {"huff_design_k256", "huff_design_k1024", "huff_design_k2048",
                                   "huff_design_k4096"};
    if (side) {
        CCA_HIP(hipEventRecord(side->fork, st));
        for (int c = 0; c < kClasses; c++) CCA_HIP(hipStreamWaitEvent(side->s[c], side->fork, 0));
    }
    for (int c = kClasses - 1; c >= 0; c--) {
        hipStream_t sc = side ? side->s[c] : st;
        if (c == 0 && lane_k) {
            const dim3 lg(std::min(div_up(nseq, (uint32_t)kLaneThreads), grid[0]));
#define CCA_LANES(K)                                                                                              \
    CCA_LAUNCH("huff_design_lanes", 0, k_design_lanes<K>, lg, dim3(kLaneThreads), 0, sc, d_seqs, p, scratch,    \
               scratch_cap, d_state, d_classes, nseq)
            if (lane_k <= 32) CCA_LANES(32);
            else if (lane_k <= 64) CCA_LANES(64);
            else if (lane_k <= 128) CCA_LANES(128);
            else CCA_LANES(256);
#undef CCA_LANES
            CCA_LAUNCH_CHECK();
        }
        const size_t lds = design::heap_region_bytes<uint16_t>(kmax[c]) + 2 * (2 * kmax[c] + 2) + 16 +
                           12 * design::kStackRanges + 16;
        CCA_LAUNCH(names[c], 0, (k_design<uint16_t, false>), dim3(std::min(nseq, grid[c])), dim3(64), lds, sc, d_seqs,
                   p, scratch, scratch_cap, d_state, d_classes, nseq, (uint32_t)c);
        CCA_LAUNCH_CHECK();
    }
    CCA_LAUNCH("huff_design_global", 0, (k_design<uint32_t, true>), dim3(std::min(nseq, 256u)), dim3(64), 0, st,
               d_seqs, p, scratch, scratch_cap, d_state, d_classes, nseq, 4u);
    CCA_LAUNCH_CHECK();
    if (side) {
        for (int c = 0; c < kClasses; c++) {
            CCA_HIP(hipEventRecord(side->join[c], side->s[c]));
            CCA_HIP(hipStreamWaitEvent(st, side->join[c], 0));
        }
    }
    // every table designed: the sections' sizes are known, so the layout comes before the pack,
    // which writes each section in place
    if (link && link->wait) CCA_HIP(hipStreamWaitEvent(st, link->wait, 0));
    const uint32_t nblocks = div_up(nitems, 1024u);
    uint64_t* bsum = item_off + nitems + 1;  // caller reserves nitems + 1 + nblocks u64
    CCA_LAUNCH("code_layout", 0, k_layout_sizes, dim3(nblocks), dim3(1024), 0, st, d_items, nitems, d_seqs, p,
               item_off, bsum);
    CCA_LAUNCH_CHECK();
    CCA_LAUNCH("code_layout", 0, k_layout_offsets, dim3(nblocks), dim3(1024), 0, st, d_items, nitems, d_seqs, p,
               d_state, item_off, bsum, out_cap, link ? link->prev : nullptr);
    CCA_LAUNCH_CHECK();
    if (ncalls)
        CCA_LAUNCH("code_layout", 0, k_layout_calls, dim3(div_up(ncalls, 256u)), dim3(256), 0, st, d_items, nitems,
                   d_seqs, d_calls, ncalls, d_call_first_item, item_off, out_cap);
    CCA_LAUNCH_CHECK();
    if (link && link->signal) CCA_HIP(hipEventRecord(link->signal, st));
    if (symSize == 2)
        CCA_LAUNCH("code_pack", inb, k_code_pack<uint16_t>, dim3(nitems), dim3(kCoderThreads), 0, st, d_items, d_seqs,
                   p, scratch, d_state, d_out, out_cap, d_item_order);
    else
        CCA_LAUNCH("code_pack", inb, k_code_pack<uint32_t>, dim3(nitems), dim3(kCoderThreads), 0, st, d_items, d_seqs,
                   p, scratch, d_state, d_out, out_cap, d_item_order);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
