// Zero-run length coding for gfx950.
//
// Format (src/cudaCompress/RunLengthKernels.cui:5-44, src/CPU/RunLengthCPU.cpp:8-32): every
// non-zero symbol becomes a pair (symbol, zeros-before % 256), preceded by zeros/256 filler pairs
// (0, 255); trailing zeros produce nothing.  Decode (RunLengthKernels.cui:57-76): the output
// index of pair i is sum_{j<=i} zc[j] + i.
//
// Encode is three launches over 2048-symbol tiles (4 waves x 8 chunks of 64):
//   stats : per tile non-zero count, first/last non-zero position, fillers inside the tile
//   scan  : one workgroup per stream turns tile stats into output offsets (the filler count of
//           a tile's first non-zero depends on the last non-zero of the tiles before it)
//   emit  : re-reads the tile (8 values per lane kept in registers), places pairs with ballot
//           ranks, fillers written cooperatively by the wave.
// A run of >= 256 zeros spans more than 4 chunks, so inside a 64-symbol chunk only the first
// non-zero can carry fillers -- the rest have gaps < 64.
#include "../kernels.hpp"
#include "wave.hpp"

namespace cca {
namespace {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr int kChunks = 8;          // chunks of 64 per wave
constexpr int kWaveSpan = 64 * kChunks;

template <typename S>
__device__ __forceinline__ uint32_t ld(const void* p, uint32_t i) {
    return (uint32_t)((const S*)p)[i];
}
template <typename S>
__device__ __forceinline__ void st(void* p, uint32_t i, uint32_t v) {
    ((S*)p)[i] = (S)v;
}

struct WaveStats {
    uint32_t nnz, first, last, inner_fill, zc_max, sym_max;
};

// Stats of one wave's 512 symbols held in registers v[0..7] (v[c] = symbol at wstart + 64c + lane).
__device__ __forceinline__ WaveStats wave_stats(const uint32_t* v, uint32_t wstart) {
    WaveStats r{0, NONE, NONE, 0, 0, 0};
    const int lane = lane_id();
    uint32_t prev = NONE;
    uint32_t zc_in = 0, smax = 0;
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        uint64_t mask = __ballot(v[c] != 0);
        if (mask == 0) continue;
        int lead = __ffsll((unsigned long long)mask) - 1;
        int hi = 63 - __clzll((unsigned long long)mask);
        uint32_t leadpos = wstart + 64 * c + lead;
        if (prev != NONE) {
            uint32_t gap = leadpos - prev - 1;
            r.inner_fill += gap / 256;
            r.zc_max = max(r.zc_max, gap >= 256 ? 255u : gap);
        } else {
            r.first = leadpos;
        }
        if (v[c] != 0 && lane != lead) {
            uint64_t below = mask & lanes_below(lane);
            int pl = 63 - __clzll((unsigned long long)below);
            zc_in = max(zc_in, (uint32_t)(lane - pl - 1));
        }
        smax = max(smax, v[c]);
        r.nnz += __popcll((unsigned long long)mask);
        prev = wstart + 64 * c + hi;
    }
    r.last = prev;
    r.zc_max = max(r.zc_max, wave_max(zc_in));
    r.sym_max = wave_max(smax);
    return r;
}

template <typename S>
__device__ __forceinline__ void load_wave(const void* in, uint32_t n, uint32_t wstart, uint32_t* v) {
    const int lane = lane_id();
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        uint32_t p = wstart + 64 * c + lane;
        v[c] = p < n ? ld<S>(in, p) : 0u;
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_rle_stats(const RleStream* __restrict__ streams, RleTile* __restrict__ tiles,
                                                   const uint32_t* __restrict__ tile_stream) {
    const uint32_t tile = blockIdx.x;
    const RleStream d = streams[tile_stream[tile]];
    const uint32_t tstart = (tile - d.tile_base) * (uint32_t)kRleTile;
    const int wave = threadIdx.x >> 6;
    const uint32_t wstart = tstart + wave * kWaveSpan;
    uint32_t v[kChunks];
    load_wave<S>(d.in, d.n, wstart, v);
    WaveStats ws = wave_stats(v, wstart);
    __shared__ WaveStats sh[4];
    if (lane_id() == 0) sh[wave] = ws;
    __syncthreads();
    if (threadIdx.x == 0) {
        RleTile t{};
        t.first_pos = NONE;
        t.last_pos = NONE;
        uint32_t zc_max = 0, sym_max = 0;
        for (int w = 0; w < 4; w++) {
            const WaveStats& s = sh[w];
            sym_max = max(sym_max, s.sym_max);
            if (s.nnz == 0) continue;
            zc_max = max(zc_max, s.zc_max);
            t.inner_fill += s.inner_fill;
            if (t.last_pos != NONE) {
                uint32_t gap = s.first - t.last_pos - 1;
                t.inner_fill += gap / 256;
                zc_max = max(zc_max, gap >= 256 ? 255u : gap);
            } else {
                t.first_pos = s.first;
            }
            t.last_pos = s.last;
            t.nnz += s.nnz;
        }
        t.pad[0] = zc_max;
        t.pad[1] = sym_max;
        tiles[tile] = t;
    }
}

// One workgroup per stream: tile offsets, pair count, maxima.  pad[2] of each tile receives the
// filler count of its first non-zero, pad[0] its zero count (tile max zero count is folded into
// the stream result first).
__global__ __launch_bounds__(256) void k_rle_scan(const RleStream* __restrict__ streams, RleTile* __restrict__ tiles,
                                                  RleResult* __restrict__ results) {
    const RleStream d = streams[blockIdx.x];
    const uint32_t ntiles = div_up_dev(d.n, kRleTile);
    __shared__ uint32_t scratch[4];
    __shared__ uint32_t smax_sh[4], zmax_sh[4];
    uint32_t carry_last = 0;  // last non-zero position + 1 before this chunk (0: none)
    uint32_t carry_out = 0;
    uint32_t zmax = 0, smax = 0;
    for (uint32_t c0 = 0; c0 < ntiles; c0 += 256) {
        const uint32_t ti = c0 + threadIdx.x;
        RleTile t{};
        bool ok = ti < ntiles;
        if (ok) t = tiles[d.tile_base + ti];
        // exclusive max-scan of (last + 1) over non-empty tiles
        uint32_t key = (ok && t.nnz > 0) ? t.last_pos + 1 : 0u;
        uint32_t incl = key;
        const int lane = lane_id(), wave = threadIdx.x >> 6;
        for (int off = 1; off < 64; off <<= 1) {
            uint32_t o = __shfl_up(incl, off, 64);
            if (lane >= off) incl = max(incl, o);
        }
        if (lane == 63) scratch[wave] = incl;
        __syncthreads();
        uint32_t before = carry_last;
        for (int w = 0; w < wave; w++) before = max(before, scratch[w]);
        uint32_t excl = max(before, (uint32_t)__shfl_up(incl, 1, 64));
        if (lane == 0) excl = before;
        uint32_t chunk_last = carry_last;
        for (int w = 0; w < 4; w++) chunk_last = max(chunk_last, scratch[w]);
        __syncthreads();
        uint32_t pairs = 0, ffill = 0, fzc = 0;
        if (ok && t.nnz > 0) {
            uint32_t gap = t.first_pos - excl;  // prev position = excl - 1
            ffill = gap / 256;
            fzc = gap % 256;
            pairs = t.nnz + t.inner_fill + ffill;
            zmax = max(zmax, max(t.pad[0], ffill > 0 ? 255u : fzc));
        }
        if (ok) smax = max(smax, t.pad[1]);
        uint32_t total;
        uint32_t off = block_excl_scan<256>(pairs, scratch, total);
        if (ok) {
            RleTile& o = tiles[d.tile_base + ti];
            o.out_base = carry_out + off;
            o.pad[2] = ffill;
            o.pad[0] = fzc;
        }
        carry_out += total;
        carry_last = chunk_last;
    }
    zmax = wave_max(zmax);
    smax = wave_max(smax);
    if (lane_id() == 0) { zmax_sh[threadIdx.x >> 6] = zmax; smax_sh[threadIdx.x >> 6] = smax; }
    __syncthreads();
    if (threadIdx.x == 0) {
        RleResult r;
        r.m = carry_out;
        r.zc_max = max(max(zmax_sh[0], zmax_sh[1]), max(zmax_sh[2], zmax_sh[3]));
        r.sym_max = max(max(smax_sh[0], smax_sh[1]), max(smax_sh[2], smax_sh[3]));
        r.pad = 0;
        results[blockIdx.x] = r;
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_rle_emit(const RleStream* __restrict__ streams, const RleTile* __restrict__ tiles,
                                                  const uint32_t* __restrict__ tile_stream) {
    const uint32_t tile = blockIdx.x;
    const RleStream d = streams[tile_stream[tile]];
    const RleTile t = tiles[tile];
    if (t.nnz == 0) return;  // uniform across the workgroup; no barrier follows for an empty tile
    const uint32_t tstart = (tile - d.tile_base) * (uint32_t)kRleTile;
    const int wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t wstart = tstart + wave * kWaveSpan;
    uint32_t v[kChunks];
    load_wave<S>(d.in, d.n, wstart, v);
    WaveStats ws = wave_stats(v, wstart);
    __shared__ WaveStats sh[4];
    if (lane == 0) sh[wave] = ws;
    __syncthreads();
    // output base and the filler/zero count of this wave's first non-zero
    uint32_t base = t.out_base, prev_last = NONE;
    for (int w = 0; w < wave; w++) {
        const WaveStats& s = sh[w];
        if (s.nnz == 0) continue;
        uint32_t ff = (prev_last == NONE) ? t.pad[2] : (s.first - prev_last - 1) / 256;
        base += s.nnz + s.inner_fill + ff;
        prev_last = s.last;
    }
    if (ws.nnz == 0) return;
    uint32_t cur = base;
    uint32_t prev = prev_last;  // NONE: this wave holds the tile's first non-zero
    S* osym = (S*)d.out_sym;
    S* ozc = (S*)d.out_zc;
#pragma unroll
    for (int c = 0; c < kChunks; c++) {
        uint64_t mask = __ballot(v[c] != 0);
        if (mask == 0) continue;
        int lead = __ffsll((unsigned long long)mask) - 1;
        int hi = 63 - __clzll((unsigned long long)mask);
        uint32_t leadpos = wstart + 64 * c + lead;
        uint32_t fill, zfirst;
        if (prev == NONE) {
            fill = t.pad[2];
            zfirst = t.pad[0];
        } else {
            uint32_t gap = leadpos - prev - 1;
            fill = gap / 256;
            zfirst = gap % 256;
        }
        for (uint32_t f = lane; f < fill; f += 64) {
            osym[cur + f] = (S)0;
            ozc[cur + f] = (S)255;
        }
        cur += fill;
        if (v[c] != 0) {
            uint64_t below = mask & lanes_below(lane);
            uint32_t rank = __popcll((unsigned long long)below);
            uint32_t zc = (lane == lead) ? zfirst : (uint32_t)(lane - (63 - __clzll((unsigned long long)below)) - 1);
            osym[cur + rank] = (S)v[c];
            ozc[cur + rank] = (S)zc;
        }
        cur += __popcll((unsigned long long)mask);
        prev = wstart + 64 * c + hi;
    }
}

// Per-stream max symbol for streams coded without RLE (HuffmanTable.cu:396-401 reduce).
template <typename S>
__global__ __launch_bounds__(256) void k_stream_max(const RleStream* __restrict__ streams, RleResult* __restrict__ results) {
    const RleStream d = streams[blockIdx.y];
    uint32_t m = 0;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < d.n; i += gridDim.x * 256) m = max(m, ld<S>(d.in, i));
    m = wave_max(m);
    if (lane_id() == 0 && m > 0) atomicMax(&results[blockIdx.y].sym_max, m);
}

__global__ void k_init_results(RleResult* r, uint32_t n, const RleStream* streams) {
    uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) { r[i].m = streams[i].n; r[i].sym_max = 0; r[i].zc_max = 0; r[i].pad = 0; }
}

// ---- decode ----------------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(256) void k_rld_sums(const RleDecStream* __restrict__ streams, uint32_t* __restrict__ sums,
                                                  const uint32_t* __restrict__ tile_stream) {
    const uint32_t tile = blockIdx.x;
    const RleDecStream d = streams[tile_stream[tile]];
    const uint32_t start = (tile - d.tile_base) * (uint32_t)kRleTile;
    uint32_t acc = 0;
    for (uint32_t i = start + threadIdx.x; i < min(start + (uint32_t)kRleTile, d.m); i += 256) acc += ld<S>(d.zc, i) + 1u;
    __shared__ uint32_t scratch[4];
    uint32_t total;
    block_excl_scan<256>(acc, scratch, total);
    if (threadIdx.x == 0) sums[tile] = total;
}

__global__ __launch_bounds__(256) void k_rld_scan(const RleDecStream* __restrict__ streams, uint32_t* __restrict__ sums) {
    const RleDecStream d = streams[blockIdx.x];
    const uint32_t nt = div_up_dev(d.m, kRleTile);
    __shared__ uint32_t scratch[4];
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nt; c0 += 256) {
        uint32_t i = c0 + threadIdx.x;
        uint32_t v = i < nt ? sums[d.tile_base + i] : 0u;
        uint32_t total;
        uint32_t e = block_excl_scan<256>(v, scratch, total);
        if (i < nt) sums[d.tile_base + i] = carry + e;
        carry += total;
    }
}

template <typename S>
__global__ __launch_bounds__(256) void k_rld_zero(const RleDecStream* __restrict__ streams) {
    const RleDecStream d = streams[blockIdx.y];
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < d.n; i += gridDim.x * 256) st<S>(d.out, i, 0u);
}

template <typename S>
__global__ __launch_bounds__(256) void k_rld_scatter(const RleDecStream* __restrict__ streams, const uint32_t* __restrict__ sums,
                                                     const uint32_t* __restrict__ tile_stream) {
    const uint32_t tile = blockIdx.x;
    const RleDecStream d = streams[tile_stream[tile]];
    const uint32_t start = (tile - d.tile_base) * (uint32_t)kRleTile;
    const uint32_t end = min(start + (uint32_t)kRleTile, d.m);
    // 8 consecutive pairs per thread
    const uint32_t i0 = start + threadIdx.x * 8;
    uint32_t zc[8], sy[8], loc = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        bool ok = i0 + k < end;
        zc[k] = ok ? ld<S>(d.zc, i0 + k) : 0u;
        sy[k] = ok ? ld<S>(d.sym, i0 + k) : 0u;
        loc += ok ? zc[k] + 1u : 0u;
    }
    __shared__ uint32_t scratch[4];
    uint32_t total;
    uint32_t pos = sums[tile] + block_excl_scan<256>(loc, scratch, total);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        if (i0 + k < end) {
            pos += zc[k];
            if (pos < d.n) st<S>(d.out, pos, sy[k]);
            pos += 1;
        }
    }
}

}  // namespace

void rle_encode(const RleStream* d_streams, uint32_t nstreams, RleTile* d_tiles, uint32_t ntiles,
                const uint32_t* d_tile_stream, RleResult* d_results, int symSize, uint64_t in_elems, hipStream_t st) {
    const uint64_t inb = in_elems * symSize;
    if (nstreams == 0) return;
    if (ntiles > 0) {
        if (symSize == 2) CCA_LAUNCH("rle_stats", inb, k_rle_stats<uint16_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tiles, d_tile_stream);
        else CCA_LAUNCH("rle_stats", inb, k_rle_stats<uint32_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tiles, d_tile_stream);
        CCA_LAUNCH_CHECK();
    }
    CCA_LAUNCH("rle_scan", (uint64_t)ntiles * 2 * sizeof(RleTile), k_rle_scan, dim3(nstreams), dim3(256), 0, st, d_streams, d_tiles, d_results);
    CCA_LAUNCH_CHECK();
    if (ntiles > 0) {
        if (symSize == 2) CCA_LAUNCH("rle_emit", inb, k_rle_emit<uint16_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tiles, d_tile_stream);
        else CCA_LAUNCH("rle_emit", inb, k_rle_emit<uint32_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tiles, d_tile_stream);
        CCA_LAUNCH_CHECK();
    }
}

void stream_max(const RleStream* d_streams, uint32_t nstreams, uint32_t nmax, RleResult* d_results, int symSize,
                hipStream_t st) {
    const uint64_t inb = (uint64_t)nmax * nstreams * symSize;
    if (nstreams == 0) return;
    CCA_LAUNCH("stream_max_init", (uint64_t)nstreams * 16, k_init_results, dim3(div_up(nstreams, 256)), dim3(256), 0, st, d_results, nstreams, d_streams);
    CCA_LAUNCH_CHECK();
    if (nmax == 0) return;
    CCA_REQUIRE(nstreams <= 65535u, CC_ERR_INVALID_ARG, "too many streams for one launch");
    dim3 grid(std::min<uint32_t>(div_up(nmax, 256 * 16), 1024u), nstreams);
    if (symSize == 2) CCA_LAUNCH("stream_max", inb, k_stream_max<uint16_t>, grid, dim3(256), 0, st, d_streams, d_results);
    else CCA_LAUNCH("stream_max", inb, k_stream_max<uint32_t>, grid, dim3(256), 0, st, d_streams, d_results);
    CCA_LAUNCH_CHECK();
}

void rle_decode(const RleDecStream* d_streams, uint32_t nstreams, uint32_t* d_tile_sums, uint32_t ntiles,
                const uint32_t* d_tile_stream, uint32_t nmax_out, int symSize, uint64_t total_pairs, uint64_t total_out,
                hipStream_t st) {
    if (nstreams == 0) return;
    CCA_REQUIRE(nstreams <= 65535u, CC_ERR_INVALID_ARG, "too many streams for one launch");
    if (nmax_out > 0) {
        dim3 grid(std::min<uint32_t>(div_up(nmax_out, 256 * 4), 4096u), nstreams);
        if (symSize == 2) CCA_LAUNCH("rld_zero", total_out * symSize, k_rld_zero<uint16_t>, grid, dim3(256), 0, st, d_streams);
        else CCA_LAUNCH("rld_zero", total_out * symSize, k_rld_zero<uint32_t>, grid, dim3(256), 0, st, d_streams);
        CCA_LAUNCH_CHECK();
    }
    if (ntiles == 0) return;
    if (symSize == 2) CCA_LAUNCH("rld_sums", total_pairs * symSize, k_rld_sums<uint16_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tile_sums, d_tile_stream);
    else CCA_LAUNCH("rld_sums", total_pairs * symSize, k_rld_sums<uint32_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tile_sums, d_tile_stream);
    CCA_LAUNCH_CHECK();
    CCA_LAUNCH("rld_scan", (uint64_t)ntiles * 8, k_rld_scan, dim3(nstreams), dim3(256), 0, st, d_streams, d_tile_sums);
    CCA_LAUNCH_CHECK();
    if (symSize == 2) CCA_LAUNCH("rld_scatter", total_pairs * 3 * symSize, k_rld_scatter<uint16_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tile_sums, d_tile_stream);
    else CCA_LAUNCH("rld_scatter", total_pairs * 3 * symSize, k_rld_scatter<uint32_t>, dim3(ntiles), dim3(256), 0, st, d_streams, d_tile_sums, d_tile_stream);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
