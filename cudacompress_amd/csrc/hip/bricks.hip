// Brick-batched CDF 9/7 passes with fused quantization (see bricks.hpp).
//
// Per coefficient the arithmetic is that of the single-volume passes in dwt.hip (Fwd97::pair /
// Inv97::out: fmaf chains in filter order over whole-sample symmetric extensions), and the
// quantizer is that of quantize.hip (src/cudaCompress/util/Quantize.cuh:76-120), so the brick
// results equal per-brick dwtFloat3DForward + quantizeToSymbols (and unquantize +
// dwtFloat3DInverseFromSymbols) bit for bit.  Octant numbering of the 7 highpass bands follows
// src/examples/CompressVolume.cpp:375-379 (band b = bx + 2 by + 4 bz, b = 1..7).
#include "../bricks.hpp"
#include "../profile.hpp"
#include "dwt_ops.hpp"
#include "quant_ops.hpp"
#include "wave.hpp"

namespace cca {
namespace {

constexpr int SP = kBrickSP;
constexpr int NC = kBrickCols;
constexpr int WF = 2 * SP + 2 * Fwd97::H;   // forward window (values per column)
constexpr int NPI = SP + 2 * Inv97::G;      // inverse window (pairs per column)
constexpr int LW = 2 * NPI + 1;              // LDS row length (>= WF + 1, >= 2 SP)
static_assert(LW >= WF + 1 && LW >= 2 * SP, "LDS row too short");

__device__ __forceinline__ uint32_t find_brick(const uint32_t* tb, uint32_t nb, uint32_t t) {
    uint32_t lo = 0, hi = nb;  // tb[lo] <= t < tb[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (tb[mid] <= t) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Brick data is accessed through global-address-space pointers (the descriptors hold generic
// ones), so loads and stores use global instructions with scalar bases.
template <class T>
__device__ __forceinline__ T gld(const T* p, int i) {
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <class T>
__device__ __forceinline__ void gst(T* p, int i, T v) {
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// Extensions with one reflection, valid when the line is long enough for the window (forward:
// n >= 5; synthesis: m >= 3); shorter lines use the repeated forms of dwt_ops.hpp.
__device__ __forceinline__ int mirror1(int j, int n) {
    j = j < 0 ? -j : j;
    return j >= n ? 2 * (n - 1) - j : j;
}
__device__ __forceinline__ int ext_low1(int g, int m) {
    g = g < 0 ? -g : g;
    return g >= m ? 2 * m - 1 - g : g;
}
__device__ __forceinline__ int ext_high1(int g, int m) {
    g = g < 0 ? -g - 1 : g;
    return g >= m ? 2 * m - 2 - g : g;
}

template <int AX>
struct Geo {
    int n, m, ncols, lx, ly, stride;
    __device__ Geo(int x, int y, int z) : lx(x), ly(y) {
        n = AX == 0 ? x : (AX == 1 ? y : z);
        m = n / 2;
        ncols = AX == 0 ? y * z : (AX == 1 ? x * z : x * y);
        stride = AX == 0 ? 1 : (AX == 1 ? x : x * y);
    }
    // element offset of position 0 of column c in a dense lx*ly*lz brick (sizes < 2^31 elements)
    __device__ int base(int c) const {
        if (AX == 0) return c * lx;
        if (AX == 1) return (c % lx) + (c / lx) * lx * ly;
        return c;
    }
};

// Window of a forward tile: every lane issues all of its loads before writing any to LDS.
// fetch(c, position, offset) -> float.  LONG: one reflection suffices (n >= 5).
template <int AX, bool LONG, class Fetch>
__device__ __forceinline__ void fwd_window(const Geo<AX>& g, int c0, int k0, float (*lds)[LW], Fetch fetch) {
    const int tid = threadIdx.x;
    constexpr int kLoads = (NC * WF + 255) / 256;
    float v[kLoads];
    if (AX == 0) {
#pragma unroll
        for (int t = 0; t < kLoads; t++) {
            const int i = tid + 256 * t;
            const int col = i / WF, r = i % WF;
            const int c = c0 + col, p = 2 * k0 - Fwd97::H + r;
            const bool ok = i < NC * WF && c < g.ncols && p < g.n + Fwd97::H;
            const int pm = LONG ? mirror1(p, g.n) : mirror(p, g.n);
            v[t] = ok ? fetch(c, pm, c * g.lx + pm) : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < kLoads; t++) {
            const int i = tid + 256 * t;
            if (i < NC * WF) lds[i / WF][i % WF] = v[t];
        }
    } else {
        const int col = tid & 63, c = c0 + col;
        const bool cv = c < g.ncols;
        const int cb = cv ? g.base(c) : 0;
#pragma unroll
        for (int t = 0; t < kLoads; t++) {
            const int r = (tid >> 6) + 4 * t, p = 2 * k0 - Fwd97::H + r;
            const bool ok = r < WF && cv && p < g.n + Fwd97::H;
            const int pm = LONG ? mirror1(p, g.n) : mirror(p, g.n);
            v[t] = ok ? fetch(c, pm, cb + pm * g.stride) : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < kLoads; t++) {
            const int r = (tid >> 6) + 4 * t;
            if (r < WF) lds[col][r] = v[t];
        }
    }
}

// Forward tile of NC columns x SP pairs: the window (fwd_window), the analysis in LDS, then
// sink(c, output position, offset, v).
template <int AX, class Fetch, class Sink>
__device__ void fwd_tile(const Geo<AX>& g, uint32_t local, float (*lds)[LW], Fetch fetch, Sink sink) {
    const int ctiles = (g.ncols + NC - 1) / NC;
    const int c0 = (int)(local % ctiles) * NC, k0 = (int)(local / ctiles) * SP;
    const int tid = threadIdx.x;
    if (g.n >= 5) fwd_window<AX, true>(g, c0, k0, lds, fetch);
    else fwd_window<AX, false>(g, c0, k0, lds, fetch);
    __syncthreads();
    const int col = tid & 63;
    float lo[SP / 4], hi[SP / 4];
#pragma unroll
    for (int j = 0; j < SP / 4; j++) {
        const int q = (tid >> 6) + 4 * j;
        lo[j] = hi[j] = 0.0f;
        if (k0 + q < g.m && c0 + col < g.ncols) Fwd97::pair(lds[col], q, lo[j], hi[j]);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SP / 4; j++) {
        const int q = (tid >> 6) + 4 * j;
        lds[col][q] = lo[j];
        lds[col][SP + q] = hi[j];
    }
    __syncthreads();
    if (AX == 0) {
        for (int i = tid; i < NC * 2 * SP; i += 256) {
            const int cc = i / (2 * SP), j = i % (2 * SP);
            const int q = j % SP, k = k0 + q, c = c0 + cc;
            const int po = j >= SP ? g.m + k : k;
            if (c < g.ncols && k < g.m) sink(c, po, c * g.lx + po, lds[cc][j]);
        }
    } else {
        const int c = c0 + col;
        const int cb = c < g.ncols ? g.base(c) : 0;
        if (c < g.ncols)
            for (int j = tid >> 6; j < 2 * SP; j += 4) {
                const int q = j % SP, k = k0 + q;
                const int po = j >= SP ? g.m + k : k;
                if (k < g.m) sink(c, po, cb + po * g.stride, lds[col][j]);
            }
    }
}

// Window of an inverse tile: lane-owned entries are loaded raw (32 bits each), BATCH loads in
// flight at a time, then converted and stored to LDS.  load(c, coefficient index, offset) ->
// uint32, conv(c, index, raw) -> float.  LONG: the one-reflection extensions apply (m >= 3).
template <int AX, bool LONG, int BATCH, class Load, class Conv>
__device__ __forceinline__ void inv_window(const Geo<AX>& g, int c0, int k0, float (*lds)[LW], Load load, Conv conv) {
    const int tid = threadIdx.x;
    constexpr int kLoads = (NC * 2 * NPI + 255) / 256;
    static_assert(kLoads % BATCH == 0, "batches must divide the window");
    auto coef = [&](int v) {  // coefficient index of window entry v (interleaved low/high)
        const int gi = k0 - Inv97::G + (v >> 1);
        if (v & 1) return g.m + (LONG ? ext_high1(gi, g.m) : ext_high(gi, g.m));
        return LONG ? ext_low1(gi, g.m) : ext_low(gi, g.m);
    };
    for (int t0 = 0; t0 < kLoads; t0 += BATCH) {
        uint32_t raw[BATCH];
        if (AX == 0) {
#pragma unroll
            for (int u = 0; u < BATCH; u++) {
                const int i = tid + 256 * (t0 + u);
                const int col = i / (2 * NPI), v = i % (2 * NPI);
                const int c = c0 + col, gi = k0 - Inv97::G + (v >> 1);
                const bool ok = i < NC * 2 * NPI && c < g.ncols && gi < g.m + Inv97::G;
                const int ci = coef(v);
                raw[u] = ok ? load(c, ci, c * g.lx + ci) : 0u;
            }
#pragma unroll
            for (int u = 0; u < BATCH; u++) {
                const int i = tid + 256 * (t0 + u);
                const int col = i / (2 * NPI), v = i % (2 * NPI);
                if (i < NC * 2 * NPI) lds[col][v] = conv(c0 + col, coef(v), raw[u]);
            }
        } else {
            const int col = tid & 63, c = c0 + col;
            const bool cv = c < g.ncols;
            const int cb = cv ? g.base(c) : 0;
#pragma unroll
            for (int u = 0; u < BATCH; u++) {
                const int v = (tid >> 6) + 4 * (t0 + u), gi = k0 - Inv97::G + (v >> 1);
                const bool ok = v < 2 * NPI && cv && gi < g.m + Inv97::G;
                const int ci = coef(v);
                raw[u] = ok ? load(c, ci, cb + ci * g.stride) : 0u;
            }
#pragma unroll
            for (int u = 0; u < BATCH; u++) {
                const int v = (tid >> 6) + 4 * (t0 + u);
                if (v < 2 * NPI) lds[col][v] = conv(c, coef(v), raw[u]);
            }
        }
    }
}

// Inverse tile: the window (inv_window), the synthesis in LDS, sink(c, output position, offset, v).
template <int AX, int BATCH, class Load, class Conv, class Sink>
__device__ void inv_tile(const Geo<AX>& g, uint32_t local, float (*lds)[LW], Load load, Conv conv, Sink sink) {
    const int ctiles = (g.ncols + NC - 1) / NC;
    const int c0 = (int)(local % ctiles) * NC, k0 = (int)(local / ctiles) * SP;
    const int tid = threadIdx.x;
    if (g.m >= 3) inv_window<AX, true, BATCH>(g, c0, k0, lds, load, conv);
    else inv_window<AX, false, BATCH>(g, c0, k0, lds, load, conv);
    __syncthreads();
    const int col = tid & 63;
    // outputs 2 u and 2 u + 1 from one 10-entry window (n is even: both exist or neither)
    float o0[SP / 4], o1[SP / 4];
#pragma unroll
    for (int j = 0; j < SP / 4; j++) {
        const int u = (tid >> 6) + 4 * j;
        o0[j] = o1[j] = 0.0f;
        if (2 * (k0 + u) < g.n && c0 + col < g.ncols) {
            float w[10];
#pragma unroll
            for (int i = 0; i < 10; i++) w[i] = lds[col][2 * u + i];
            o0[j] = Inv97::out(w, 0);
            o1[j] = Inv97::out(w + 1, 1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SP / 4; j++) {
        const int u = (tid >> 6) + 4 * j;
        lds[col][2 * u] = o0[j];
        lds[col][2 * u + 1] = o1[j];
    }
    __syncthreads();
    if (AX == 0) {
        for (int i = tid; i < NC * 2 * SP; i += 256) {
            const int cc = i / (2 * SP), t = i % (2 * SP);
            const int jo = 2 * k0 + t, c = c0 + cc;
            if (c < g.ncols && jo < g.n) sink(c, jo, c * g.lx + jo, lds[cc][t]);
        }
    } else {
        const int c = c0 + col;
        const int cb = c < g.ncols ? g.base(c) : 0;
        if (c < g.ncols)
            for (int t = tid >> 6; t < 2 * SP; t += 4) {
                const int jo = 2 * k0 + t;
                if (jo < g.n) sink(c, jo, cb + jo * g.stride, lds[col][t]);
            }
    }
}

template <int AX>
__global__ __launch_bounds__(256) void k_bfwd(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float lds[NC][LW];
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<AX> g(L.lx, L.ly, L.lz);
    const float* src = AX == 0 ? L.in : L.a;
    float* dst = AX == 0 ? L.a : L.b;
    fwd_tile<AX>(g, blockIdx.x - tb[i], lds, [&](int, int, int off) { return gld(src, off); },
                 [&](int, int, int off, float v) { gst(dst, off, v); });
}

// Z passes: one thread per column (x, y) of the brick level and per segment of kBrickZP output
// pairs along z.  The 10-sample filter window streams through registers (no LDS, no barrier); a
// segment reads its z range once plus the filter halo.  A workgroup is one wave of NC columns,
// numbered band-major (the columns of octant quarter bx + 2 by of the xy plane, then y, then x
// inside it), so the lanes of a wave read and write consecutive elements of one band (and, for a
// level whose quarter holds a multiple of NC columns, all serve one quarter: the lowpass quarter's
// loads and stores differ from the highpass ones).  Tiles (NC columns x kBrickZP pairs) are those
// of brick_tiles along z.  The window and the fmaf order are those of the tiled passes (Fwd97::pair
// over positions 2q - 4 .. 2q + 5, Inv97::out over interleaved pairs u - 2 .. u + 2), so the
// results are identical.
// Workgroups are renumbered XCD by XCD (xcd_order): the waves of the four quarters of the same rows
// share their lines of the dense plane, and neighbours in logical order then share an L2.
constexpr int ZU = 8;  // output pairs per batch of loads (2 ZU loads in flight per thread)

// The column of a z-pass thread.
struct ZColumn {
    int k0, k1;     // output pairs [k0, k1) of the segment
    int q;          // quarter bx + 2 by (-1: past the level's columns)
    uint32_t r;     // (y - by hy) hx + (x - bx hx): offset inside a band plane
    uint32_t xy;    // y lx + x: offset inside a dense plane
    __device__ ZColumn(const Geo<2>& g, uint32_t local, int hx, int hy) {
        const int ctiles = (g.ncols + NC - 1) / NC;
        const int c = (int)(local % (uint32_t)ctiles) * NC + (int)threadIdx.x;
        k0 = (int)(local / (uint32_t)ctiles) * kBrickZP;
        k1 = min(k0 + kBrickZP, g.m);
        const int hxy = hx * hy;
        q = c < g.ncols ? c / hxy : -1;
        const int rr = c - q * hxy, yq = rr / hx, xq = rr - yq * hx;
        r = (uint32_t)rr;
        xy = (uint32_t)(((q >> 1) * hy + yq) * g.lx + (q & 1) * hx + xq);
    }
};

// Extensions of a pair or position index, uniform across the wave: one reflection when the line
// is long enough (LONG), else the repeated forms.
template <bool LONG>
__device__ __forceinline__ int zmirror(int j, int n) { return LONG ? mirror1(j, n) : mirror(j, n); }
template <bool LONG>
__device__ __forceinline__ int zext_low(int g, int m) { return LONG ? ext_low1(g, m) : ext_low(g, m); }
template <bool LONG>
__device__ __forceinline__ int zext_high(int g, int m) { return LONG ? ext_high1(g, m) : ext_high(g, m); }

// element i of the plane at p (p uniform, i per lane)
template <class T>
__device__ __forceinline__ T zld(const T* p, uint32_t i) {
    return ((const __attribute__((address_space(1))) T*)p)[i];
}
template <class T>
__device__ __forceinline__ void zst(T* p, uint32_t i, T v) {
    ((__attribute__((address_space(1))) T*)p)[i] = v;
}

// Forward analysis of one column segment: positions from src (plane stride nc), low outputs to
// lo (floats when LOF, else symbols of lo_sym), high outputs to the symbols of hi (band plane
// stride hxy).
template <bool LOF, bool LONG, class S>
__device__ __forceinline__ void zfwd_column(const ZColumn& zc, const float* src, int n, int nc, float* lo_f,
                                            S* lo_sym, uint32_t ilo, float inv_lo, S* hi, uint32_t ihi,
                                            float inv_hi, int hxy) {
    auto put = [&](int k, float lo, float h) {
        if (LOF) zst(lo_f + k * hxy, ilo, lo);
        else zst(lo_sym + k * hxy, ilo, (S)symbolize_i(quantize_f(lo, inv_lo, 0)));
        zst(hi + k * hxy, ihi, (S)symbolize_i(quantize_f(h, inv_hi, 0)));
    };
    float w[10];  // positions 2k - 4 .. 2k + 5 of the current pair k
#pragma unroll
    for (int j = 0; j < 8; j++) w[j] = zld(src + zmirror<LONG>(2 * zc.k0 - 4 + j, n) * nc, zc.xy);
    int k = zc.k0;
    for (; k + ZU <= zc.k1; k += ZU) {
        float v[2 * ZU];
#pragma unroll
        for (int j = 0; j < 2 * ZU; j++) v[j] = zld(src + zmirror<LONG>(2 * k + 4 + j, n) * nc, zc.xy);
#pragma unroll
        for (int u = 0; u < ZU; u++) {
            w[8] = v[2 * u];
            w[9] = v[2 * u + 1];
            float lo, h;
            Fwd97::pair(w, 0, lo, h);
            put(k + u, lo, h);
#pragma unroll
            for (int t = 0; t < 8; t++) w[t] = w[t + 2];
        }
    }
    for (; k < zc.k1; k++) {
        w[8] = zld(src + zmirror<LONG>(2 * k + 4, n) * nc, zc.xy);
        w[9] = zld(src + zmirror<LONG>(2 * k + 5, n) * nc, zc.xy);
        float lo, h;
        Fwd97::pair(w, 0, lo, h);
        put(k, lo, h);
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = w[t + 2];
    }
}

// Forward z analysis with the quantizer: L.b -> lowpass octant (L.lp, or the symbol stream on the
// coarsest level) and the highpass symbol bands.
template <class S>
__global__ __launch_bounds__(NC) void k_bfwd_zq(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    const uint32_t w = xcd_order(blockIdx.x, gridDim.x);
    const uint32_t i = find_brick(tb, nb, w);
    const BrickLevel L = lv[i];
    const Geo<2> g(L.lx, L.ly, L.lz);
    const int hx = L.lx / 2, hy = L.ly / 2, hxy = hx * hy;
    const uint32_t band = (uint32_t)hxy * (uint32_t)g.m;
    const ZColumn zc(g, w - tb[i], hx, hy);
    if (zc.q < 0) return;
    S* hp = static_cast<S*>(L.hp);
    const uint32_t ihi = (uint32_t)(zc.q + 3) * band + zc.r;
    if (g.n >= 5) {
        if (zc.q != 0) zfwd_column<false, true>(zc, L.b, g.n, g.ncols, nullptr, hp, (zc.q - 1) * band + zc.r, L.inv_hp, hp, ihi, L.inv_hp, hxy);
        else if (L.last) zfwd_column<false, true>(zc, L.b, g.n, g.ncols, nullptr, static_cast<S*>(L.lpsym), zc.r, L.inv_lp, hp, ihi, L.inv_hp, hxy);
        else zfwd_column<true, true, S>(zc, L.b, g.n, g.ncols, L.lp, nullptr, zc.r, 0.0f, hp, ihi, L.inv_hp, hxy);
    } else {
        if (zc.q != 0) zfwd_column<false, false>(zc, L.b, g.n, g.ncols, nullptr, hp, (zc.q - 1) * band + zc.r, L.inv_hp, hp, ihi, L.inv_hp, hxy);
        else if (L.last) zfwd_column<false, false>(zc, L.b, g.n, g.ncols, nullptr, static_cast<S*>(L.lpsym), zc.r, L.inv_lp, hp, ihi, L.inv_hp, hxy);
        else zfwd_column<true, false, S>(zc, L.b, g.n, g.ncols, L.lp, nullptr, zc.r, 0.0f, hp, ihi, L.inv_hp, hxy);
    }
}

// Inverse synthesis of one column segment: low coefficients from lo (floats when LOF, else
// symbols of lo_sym dequantized with step_lo), high ones from the symbols of hi (band plane
// stride hxy), outputs to dst (plane stride nc).
template <bool LOF, bool LONG, class S>
__device__ __forceinline__ void zinv_column(const ZColumn& zc, const float* lo_f, const S* lo_sym, uint32_t ilo,
                                            float step_lo, const S* hi, uint32_t ihi, float step_hi, int hxy,
                                            int m, float* dst, int nc) {
    auto raw_lo = [&](int k) -> uint32_t {
        const int e = zext_low<LONG>(k, m) * hxy;
        return LOF ? __float_as_uint(zld(lo_f + e, ilo)) : (uint32_t)zld(lo_sym + e, ilo);
    };
    auto raw_hi = [&](int k) -> uint32_t { return (uint32_t)zld(hi + zext_high<LONG>(k, m) * hxy, ihi); };
    auto val_lo = [&](uint32_t r) { return LOF ? __uint_as_float(r) : unquantize_i(unsymbolize_u(r), step_lo, 0); };
    auto val_hi = [&](uint32_t r) { return unquantize_i(unsymbolize_u(r), step_hi, 0); };
    float w[10];  // interleaved pairs u - 2 .. u + 2 of the current output pair u
    {
        uint32_t r[8];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            r[2 * j] = raw_lo(zc.k0 - 2 + j);
            r[2 * j + 1] = raw_hi(zc.k0 - 2 + j);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            w[2 * j] = val_lo(r[2 * j]);
            w[2 * j + 1] = val_hi(r[2 * j + 1]);
        }
    }
    auto emit = [&](int u) {
        zst(dst + 2 * u * nc, zc.xy, Inv97::out(w, 0));
        zst(dst + (2 * u + 1) * nc, zc.xy, Inv97::out(w + 1, 1));
#pragma unroll
        for (int t = 0; t < 8; t++) w[t] = w[t + 2];
    };
    int u = zc.k0;
    for (; u + ZU <= zc.k1; u += ZU) {
        uint32_t r[2 * ZU];
#pragma unroll
        for (int j = 0; j < ZU; j++) {
            r[2 * j] = raw_lo(u + 2 + j);
            r[2 * j + 1] = raw_hi(u + 2 + j);
        }
#pragma unroll
        for (int j = 0; j < ZU; j++) {
            w[8] = val_lo(r[2 * j]);
            w[9] = val_hi(r[2 * j + 1]);
            emit(u + j);
        }
    }
    for (; u < zc.k1; u++) {
        w[8] = val_lo(raw_lo(u + 2));
        w[9] = val_hi(raw_hi(u + 2));
        emit(u);
    }
}

// Inverse z synthesis from the symbols (dequantized on the fly) and the lowpass octant (floats of
// the level above, or symbols on the coarsest level): -> L.a.
template <class S>
__global__ __launch_bounds__(NC) void k_binv_zs(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    const uint32_t w = xcd_order(blockIdx.x, gridDim.x);
    const uint32_t i = find_brick(tb, nb, w);
    const BrickLevel L = lv[i];
    const Geo<2> g(L.lx, L.ly, L.lz);
    const int hx = L.lx / 2, hy = L.ly / 2, hxy = hx * hy;
    const uint32_t band = (uint32_t)hxy * (uint32_t)g.m;
    const ZColumn zc(g, w - tb[i], hx, hy);
    if (zc.q < 0) return;
    const S* hp = static_cast<const S*>(L.hp);
    const uint32_t ihi = (uint32_t)(zc.q + 3) * band + zc.r;
    if (g.m >= 3) {
        if (zc.q != 0) zinv_column<false, true>(zc, nullptr, hp, (zc.q - 1) * band + zc.r, L.step_hp, hp, ihi, L.step_hp, hxy, g.m, L.a, g.ncols);
        else if (L.last) zinv_column<false, true>(zc, nullptr, static_cast<const S*>(L.lpsym), zc.r, L.step_lp, hp, ihi, L.step_hp, hxy, g.m, L.a, g.ncols);
        else zinv_column<true, true, S>(zc, L.lp, nullptr, zc.r, 0.0f, hp, ihi, L.step_hp, hxy, g.m, L.a, g.ncols);
    } else {
        if (zc.q != 0) zinv_column<false, false>(zc, nullptr, hp, (zc.q - 1) * band + zc.r, L.step_hp, hp, ihi, L.step_hp, hxy, g.m, L.a, g.ncols);
        else if (L.last) zinv_column<false, false>(zc, nullptr, static_cast<const S*>(L.lpsym), zc.r, L.step_lp, hp, ihi, L.step_hp, hxy, g.m, L.a, g.ncols);
        else zinv_column<true, false, S>(zc, L.lp, nullptr, zc.r, 0.0f, hp, ihi, L.step_hp, hxy, g.m, L.a, g.ncols);
    }
}

template <int AX>
__global__ __launch_bounds__(256) void k_binv(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float lds[NC][LW];
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<AX> g(L.lx, L.ly, L.lz);
    const float* src = AX == 1 ? L.a : L.b;
    float* dst = AX == 1 ? L.b : L.out;
    inv_tile<AX, 18>(g, blockIdx.x - tb[i], lds, [&](int, int, int off) { return __float_as_uint(gld(src, off)); },
                 [&](int, int, uint32_t raw) { return __uint_as_float(raw); },
                 [&](int, int, int off, float v) { gst(dst, off, v); });
}

// ------------------------------------------------ fused x+y passes, streaming along y
// Workgroup: brick bi, slices z0 .. z0 + zk - 1 (thread t: slice t / mx, x pair k = t % mx) and
// output pairs [k0, k1) along y.  A step stages two rows of every slice in LDS (entry 4 + p for
// position p, with the extensions of the x filter written by the threads holding their sources),
// one barrier, then each thread runs the x filter of its pair on both rows.  The y filter windows
// (10 samples of the columns k and mx + k of the x-filtered plane, or of the inverse's input)
// stay in registers, as in the z passes.  Filters and windows are those of the tiled passes
// (Fwd97::pair, Inv97::out over the same extended samples), so the results are identical.
// LDS: two buffers (alternate steps, so one barrier per step suffices) of two rows.
constexpr int XYT = kBrickXYThreads;
constexpr int XRS = kBrickXYRow;
constexpr int XU = 4;  // steps per batch of loads

// 8-byte LDS accesses of the staged rows (a float array): a may_alias vector type, so the accesses
// keep their order against the array's float accesses.
typedef float xyf2 __attribute__((ext_vector_type(2), may_alias));
__device__ __forceinline__ void lds_put2(float* p, float2 v) { *(xyf2*)p = xyf2{v.x, v.y}; }
__device__ __forceinline__ void lds_get2(float* w, const float* p) {
    const xyf2 v = *(const xyf2*)p;
    w[0] = v.x;
    w[1] = v.y;
}

struct XYTile {
    int nx, ny, mx, my;
    int zk;          // slices of the workgroup
    int k0, k1;      // output pairs along y
    int s, k;        // thread: slice, x pair
    bool act;        // s < zk
    size_t plane0;   // element offset of the thread's slice (z0 + s) in the level
    int row0;        // LDS offset of the thread's slice in a staged row
    __device__ XYTile(const BrickLevel& L, uint32_t local) {
        nx = L.lx;
        ny = L.ly;
        mx = nx / 2;
        my = ny / 2;
        const int zg = (L.lz + L.zk - 1) / L.zk;
        const int z0 = (int)(local % (uint32_t)zg) * L.zk;
        k0 = (int)(local / (uint32_t)zg) * kBrickXYSeg;
        k1 = min(k0 + kBrickXYSeg, my);
        zk = min(L.zk, L.lz - z0);
        s = (int)threadIdx.x / mx;
        k = (int)threadIdx.x - s * mx;
        act = s < zk;
        plane0 = (size_t)(z0 + (act ? s : 0)) * (size_t)(nx * ny);
        row0 = s * (nx + 8);
    }
};

// Forward: stage rows a (v0 = positions 2k, 2k + 1) and a + 1 (v1) into buffer b, then the x
// analysis of pair k of both rows: lo[r] (column k), hi[r] (column mx + k).
__device__ __forceinline__ void xy_fwd_ingest(float* st, const XYTile& t, int b, float2 v0, float2 v1, float lo[2],
                                              float hi[2]) {
    float* r0 = st + (2 * b) * XRS + t.row0;
    float* r1 = r0 + XRS;
    if (t.act) {
        lds_put2(r0 + 4 + 2 * t.k, v0);
        lds_put2(r1 + 4 + 2 * t.k, v1);
        if (t.k < 3 || t.k >= t.mx - 3) {
            // whole-sample symmetric extensions: positions -4 .. -1 and nx .. nx + 3
#pragma unroll
            for (int h = 0; h < 8; h++) {
                const int p = h < 4 ? h - 4 : t.nx + h - 4;
                const int x = mirror1(p, t.nx);
                if (x == 2 * t.k) {
                    r0[4 + p] = v0.x;
                    r1[4 + p] = v1.x;
                } else if (x == 2 * t.k + 1) {
                    r0[4 + p] = v0.y;
                    r1[4 + p] = v1.y;
                }
            }
        }
    }
    __syncthreads();
    if (t.act) {
        float w[10];
#pragma unroll
        for (int j = 0; j < 5; j++) lds_get2(w + 2 * j, r0 + 2 * t.k + 2 * j);
        Fwd97::pair(w, 0, lo[0], hi[0]);
#pragma unroll
        for (int j = 0; j < 5; j++) lds_get2(w + 2 * j, r1 + 2 * t.k + 2 * j);
        Fwd97::pair(w, 0, lo[1], hi[1]);
    }
}

// Forward x then y analysis: L.in -> L.b.
__global__ __launch_bounds__(XYT) void k_bfwd_xy(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float st[4 * XRS];
    const uint32_t bi = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[bi];
    const XYTile t(L, blockIdx.x - tb[bi]);
    const int nx = t.nx, ny = t.ny, my = t.my;
    const float* src = L.in + t.plane0 + 2 * t.k;
    float* dst = L.b + t.plane0 + t.k;
    auto row = [&](int p) -> float2 {  // positions 2k, 2k + 1 of row mirror(p)
        const float* q = src + mirror(p, ny) * nx;
        return make_float2(gld(q, 0), gld(q, 1));
    };
    float wl[10], wh[10];  // y windows of the columns k and mx + k: rows 2q - 4 .. 2q + 5
    int b = 0;
    {
        float2 v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) v[j] = row(2 * t.k0 - 4 + j);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float lo[2], hi[2];
            xy_fwd_ingest(st, t, b, v[2 * j], v[2 * j + 1], lo, hi);
            b ^= 1;
            wl[2 * j] = lo[0];
            wl[2 * j + 1] = lo[1];
            wh[2 * j] = hi[0];
            wh[2 * j + 1] = hi[1];
        }
    }
    auto step = [&](int q, float2 v0, float2 v1) {
        float lo[2], hi[2];
        xy_fwd_ingest(st, t, b, v0, v1, lo, hi);
        b ^= 1;
        wl[8] = lo[0];
        wl[9] = lo[1];
        wh[8] = hi[0];
        wh[9] = hi[1];
        if (t.act) {
            float l0, h0, l1, h1;
            Fwd97::pair(wl, 0, l0, h0);
            Fwd97::pair(wh, 0, l1, h1);
            gst(dst, q * nx, l0);
            gst(dst, q * nx + t.mx, l1);
            gst(dst, (my + q) * nx, h0);
            gst(dst, (my + q) * nx + t.mx, h1);
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            wl[i] = wl[i + 2];
            wh[i] = wh[i + 2];
        }
    };
    int q = t.k0;
    for (; q + XU <= t.k1; q += XU) {
        float2 v[2 * XU];
#pragma unroll
        for (int j = 0; j < 2 * XU; j++) v[j] = row(2 * q + 4 + j);
#pragma unroll
        for (int u = 0; u < XU; u++) step(q + u, v[2 * u], v[2 * u + 1]);
    }
    for (; q < t.k1; q++) step(q, row(2 * q + 4), row(2 * q + 5));
}

// Inverse: stage the y-synthesized rows 2u (a0: low and high x coefficient k) and 2u + 1 (a1) into
// buffer b in interleaved order, then the x synthesis of output pair k of both rows to dst rows.
__device__ __forceinline__ void xy_inv_ingest(float* st, const XYTile& t, int b, float2 a0, float2 a1, float* d0,
                                              float* d1) {
    float* r0 = st + (2 * b) * XRS + t.row0;
    float* r1 = r0 + XRS;
    if (t.act) {
        lds_put2(r0 + 4 + 2 * t.k, a0);
        lds_put2(r1 + 4 + 2 * t.k, a1);
        if (t.k < 3 || t.k >= t.mx - 3) {
            // pairs -2, -1, mx, mx + 1: low coefficient ext_low, high coefficient ext_high
#pragma unroll
            for (int h = 0; h < 4; h++) {
                const int g = h < 2 ? h - 2 : t.mx + h - 2;
                if (ext_low1(g, t.mx) == t.k) {
                    r0[4 + 2 * g] = a0.x;
                    r1[4 + 2 * g] = a1.x;
                }
                if (ext_high1(g, t.mx) == t.k) {
                    r0[4 + 2 * g + 1] = a0.y;
                    r1[4 + 2 * g + 1] = a1.y;
                }
            }
        }
    }
    __syncthreads();
    if (t.act) {
        float w[10];
#pragma unroll
        for (int j = 0; j < 5; j++) lds_get2(w + 2 * j, r0 + 2 * t.k + 2 * j);
        gst(d0, 0, Inv97::out(w, 0));
        gst(d0, 1, Inv97::out(w + 1, 1));
#pragma unroll
        for (int j = 0; j < 5; j++) lds_get2(w + 2 * j, r1 + 2 * t.k + 2 * j);
        gst(d1, 0, Inv97::out(w, 0));
        gst(d1, 1, Inv97::out(w + 1, 1));
    }
}

// Inverse y then x synthesis: L.a -> L.out.
__global__ __launch_bounds__(XYT) void k_binv_yx(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float st[4 * XRS];
    const uint32_t bi = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[bi];
    const XYTile t(L, blockIdx.x - tb[bi]);
    const int nx = t.nx, my = t.my;
    const float* src = L.a + t.plane0 + t.k;
    float* dst = L.out + t.plane0 + 2 * t.k;
    // low (row ext_low(g)) and high (row my + ext_high(g)) y coefficient g of the columns k, mx + k
    auto lo_row = [&](int g) -> float2 {
        const float* q = src + ext_low(g, my) * nx;
        return make_float2(gld(q, 0), gld(q, t.mx));
    };
    auto hi_row = [&](int g) -> float2 {
        const float* q = src + (my + ext_high(g, my)) * nx;
        return make_float2(gld(q, 0), gld(q, t.mx));
    };
    float wl[10], wh[10];  // interleaved y pairs u - 2 .. u + 2 of the columns k and mx + k
    {
        float2 v[8];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v[2 * j] = lo_row(t.k0 - 2 + j);
            v[2 * j + 1] = hi_row(t.k0 - 2 + j);
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            wl[j] = v[j].x;
            wh[j] = v[j].y;
        }
    }
    int b = 0;
    auto step = [&](int u, float2 vl, float2 vh) {
        wl[8] = vl.x;
        wl[9] = vh.x;
        wh[8] = vl.y;
        wh[9] = vh.y;
        const float2 a0 = make_float2(Inv97::out(wl, 0), Inv97::out(wh, 0));
        const float2 a1 = make_float2(Inv97::out(wl + 1, 1), Inv97::out(wh + 1, 1));
        xy_inv_ingest(st, t, b, a0, a1, dst + 2 * u * nx, dst + (2 * u + 1) * nx);
        b ^= 1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            wl[i] = wl[i + 2];
            wh[i] = wh[i + 2];
        }
    };
    int u = t.k0;
    for (; u + XU <= t.k1; u += XU) {
        float2 v[2 * XU];
#pragma unroll
        for (int j = 0; j < XU; j++) {
            v[2 * j] = lo_row(u + 2 + j);
            v[2 * j + 1] = hi_row(u + 2 + j);
        }
#pragma unroll
        for (int j = 0; j < XU; j++) step(u + j, v[2 * j], v[2 * j + 1]);
    }
    for (; u < t.k1; u++) step(u, lo_row(u + 2), hi_row(u + 2));
}

}  // namespace

void bricks_forward_level(const BrickLevel* d_levels, const uint32_t* d_tiles, uint32_t nbricks,
                          const uint32_t* h_tiles_total, int symSize, uint64_t level_elems, hipStream_t st) {
    const uint32_t* tx = d_tiles;
    const uint32_t* ty = d_tiles + (nbricks + 1);
    const uint32_t* tz = d_tiles + 2 * (nbricks + 1);
    const uint32_t* txy = d_tiles + 3 * (nbricks + 1);
    if (h_tiles_total[3]) {
        CCA_LAUNCH("bricks_fwd_xy", level_elems * 8, k_bfwd_xy, dim3(h_tiles_total[3]), dim3(XYT), 0,
                   st, d_levels, txy, nbricks);
    } else {
        if (h_tiles_total[0]) CCA_LAUNCH("bricks_fwd_x", level_elems * 8, k_bfwd<0>, dim3(h_tiles_total[0]), dim3(256), 0, st, d_levels, tx, nbricks);
        CCA_LAUNCH_CHECK();
        if (h_tiles_total[1]) CCA_LAUNCH("bricks_fwd_y", level_elems * 8, k_bfwd<1>, dim3(h_tiles_total[1]), dim3(256), 0, st, d_levels, ty, nbricks);
    }
    CCA_LAUNCH_CHECK();
    const uint64_t zb = level_elems * 4 + level_elems / 8 * 4 + level_elems * 7 / 8 * symSize;
    if (h_tiles_total[2]) {
        if (symSize == 2)
            CCA_LAUNCH("bricks_fwd_z_quant", zb, k_bfwd_zq<uint16_t>, dim3(h_tiles_total[2]), dim3(NC), 0, st, d_levels, tz, nbricks);
        else
            CCA_LAUNCH("bricks_fwd_z_quant", zb, k_bfwd_zq<uint32_t>, dim3(h_tiles_total[2]), dim3(NC), 0, st, d_levels, tz, nbricks);
    }
    CCA_LAUNCH_CHECK();
}

void bricks_inverse_level(const BrickLevel* d_levels, const uint32_t* d_tiles, uint32_t nbricks,
                          const uint32_t* h_tiles_total, int symSize, uint64_t level_elems, hipStream_t st) {
    const uint32_t* tx = d_tiles;
    const uint32_t* ty = d_tiles + (nbricks + 1);
    const uint32_t* tz = d_tiles + 2 * (nbricks + 1);
    const uint64_t zb = level_elems * 4 + level_elems / 8 * 4 + level_elems * 7 / 8 * symSize;
    if (h_tiles_total[2]) {
        if (symSize == 2)
            CCA_LAUNCH("bricks_inv_z_dequant", zb, k_binv_zs<uint16_t>, dim3(h_tiles_total[2]), dim3(NC), 0, st, d_levels, tz, nbricks);
        else
            CCA_LAUNCH("bricks_inv_z_dequant", zb, k_binv_zs<uint32_t>, dim3(h_tiles_total[2]), dim3(NC), 0, st, d_levels, tz, nbricks);
    }
    CCA_LAUNCH_CHECK();
    if (h_tiles_total[3]) {
        CCA_LAUNCH("bricks_inv_yx", level_elems * 8, k_binv_yx, dim3(h_tiles_total[3]), dim3(XYT), 0,
                   st, d_levels, d_tiles + 3 * (nbricks + 1), nbricks);
    } else {
        if (h_tiles_total[1]) CCA_LAUNCH("bricks_inv_y", level_elems * 8, k_binv<1>, dim3(h_tiles_total[1]), dim3(256), 0, st, d_levels, ty, nbricks);
        CCA_LAUNCH_CHECK();
        if (h_tiles_total[0]) CCA_LAUNCH("bricks_inv_x", level_elems * 8, k_binv<0>, dim3(h_tiles_total[0]), dim3(256), 0, st, d_levels, tx, nbricks);
    }
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
