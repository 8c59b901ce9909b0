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

// Z passes: one thread per column (x, y) of the brick level and per segment of SP output pairs
// along z.  The 10-sample filter window streams through registers (no LDS, no barrier); a
// segment reads its z range once plus the filter halo.  A workgroup is one wave of NC columns,
// numbered band-major (the columns of octant quarter bx + 2 by of the xy plane, then y, then x
// inside it), so the lanes of a wave read and write consecutive elements of one band (and, for a
// level whose quarter holds a multiple of NC columns, all serve one quarter: the lowpass quarter's
// loads and stores differ from the highpass ones).  Tiles (NC columns x SP pairs) are those of
// brick_tiles along z.  The window and the fmaf order are those of the tiled passes
// (Fwd97::pair over positions 2q - 4 .. 2q + 5, Inv97::out over interleaved pairs u - 2 .. u + 2),
// so the results are identical.
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
        k0 = (int)(local / (uint32_t)ctiles) * SP;
        k1 = min(k0 + SP, g.m);
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
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<2> g(L.lx, L.ly, L.lz);
    const int hx = L.lx / 2, hy = L.ly / 2, hxy = hx * hy;
    const uint32_t band = (uint32_t)hxy * (uint32_t)g.m;
    const ZColumn zc(g, blockIdx.x - tb[i], hx, hy);
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
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<2> g(L.lx, L.ly, L.lz);
    const int hx = L.lx / 2, hy = L.ly / 2, hxy = hx * hy;
    const uint32_t band = (uint32_t)hxy * (uint32_t)g.m;
    const ZColumn zc(g, blockIdx.x - tb[i], hx, hy);
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

// ------------------------------------------------ fused x+y passes over whole z-slices in LDS
// Workgroup w holds zk slices of one brick: R = zk ly rows (x lines) and C = zk lx columns (y
// lines), each line with its extensions in LDS (entry 4 + p for window position p), strides lx + 9
// and ly + 9 (odd: lanes on consecutive lines hit distinct banks).  The filters are Fwd97::pair and
// Inv97::out over the same windows as the tiled passes, so the results are identical.
constexpr int XYT = kBrickXYThreads;
constexpr int XYB = kBrickXYFloats;
constexpr int NIF = (XYB / 2 + XYT - 1) / XYT;  // forward pairs per thread (R mx < XYB / 2)
constexpr int NII = (XYB / 2 + XYT - 1) / XYT;  // inverse output pairs per thread (C my < XYB / 2)

// Exact quotient e / d for 0 <= e < 2^16 and 0 < d < 2^16 (one multiply-high; m wraps to 0 for
// d = 1).
struct Div {
    uint32_t m;
    __device__ explicit Div(int d) : m(0xFFFFFFFFu / (uint32_t)d + 1u) {}
    __device__ int q(int e) const { return m ? (int)__umulhi((uint32_t)e, m) : e; }
};

// e = 0 .. E-1 of a dense run of floats, sixteen loads in flight per thread; put(e, v).
template <class Put>
__device__ __forceinline__ void xy_load(const float* src, int E, Put put) {
    constexpr int U = 16;
    for (int e0 = threadIdx.x; e0 < E; e0 += XYT * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = e0 + XYT * u;
            v[u] = e < E ? gld(src, e) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int e = e0 + XYT * u;
            if (e < E) put(e, v[u]);
        }
    }
}

// Raw buffer accesses (AMD buffer resources: base, size in bytes; an access past the size reads 0
// and a store there is dropped): 32-bit offsets and no bounds branch per element.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
constexpr uint32_t kBufOut = 0x40000000u;  // an offset past every buffer here

// The R rows of nx floats of a slab (row r at element r nx of src), put(r, x, v).  Lanes run along
// x (W, the power of two at or above nx, lanes per row, so that a row needs no division), XYT / W
// rows per step and sixteen loads in flight per thread; raw buffer loads, so rows past R and
// lanes past nx read 0 without a branch.  Rows wider than XYT take the element-linear form.
template <class Put>
__device__ __forceinline__ void xy_load_rows(const float* src, int nx, int R, Put put) {
    const int lw = nx <= 1 ? 0 : 32 - __clz(nx - 1);
    if ((1 << lw) > XYT) {
        const Div dnx(nx);
        xy_load(src, R * nx, [&](int e, float v) {
            const int r = dnx.q(e);
            put(r, e - r * nx, v);
        });
        return;
    }
    const int x = (int)threadIdx.x & ((1 << lw) - 1), r0 = (int)threadIdx.x >> lw, rs = XYT >> lw;
    const __amdgpu_buffer_rsrc_t rsrc = buffer_of(src, (uint32_t)(R * nx) * 4u);
    const uint32_t xo = x < nx ? (uint32_t)x * 4u : kBufOut;
    constexpr int U = 16;
    for (int rb = r0; rb < R; rb += rs * U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; u++)
            v[u] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(
                                                 rsrc, (int)(xo + (uint32_t)((rb + rs * u) * nx) * 4u), 0, 0));
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int r = rb + rs * u;
            if (r < R && x < nx) put(r, x, v[u]);
        }
    }
}

// Analysis extensions of `lines` lines of n samples: positions -4..-1 and n..n+3 mirror the line.
__device__ __forceinline__ void xy_fill_mirror(float* buf, int lines, int stride, int n) {
    for (int h = threadIdx.x; h < lines * 8; h += XYT) {
        const int ln = h >> 3, j = h & 7;
        const int p = j < 4 ? j - 4 : n + j - 4;
        buf[ln * stride + 4 + p] = buf[ln * stride + 4 + mirror(p, n)];
    }
}

// Synthesis extensions of interleaved lines (entry 4 + 2k + h: low (h = 0) or high coefficient k
// of m): pairs -2, -1, m, m + 1 by ext_low / ext_high.
__device__ __forceinline__ void xy_fill_inter(float* buf, int lines, int stride, int m) {
    for (int h = threadIdx.x; h < lines * 8; h += XYT) {
        const int ln = h >> 3, j = h & 7;
        const int v = j < 4 ? j - 4 : 2 * m + j - 4;
        const int k = v >> 1, hi = v & 1;
        const int ks = hi ? ext_high(k, m) : ext_low(k, m);
        buf[ln * stride + 4 + v] = buf[ln * stride + 4 + 2 * ks + hi];
    }
}

// Forward x then y analysis of zk slices: L.in -> L.b.
__global__ __launch_bounds__(XYT) void k_bfwd_xy(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    extern __shared__ float buf[];
    const uint32_t bi = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[bi];
    const int nx = L.lx, ny = L.ly, mx = nx / 2, my = ny / 2;
    const int z0 = (int)(blockIdx.x - tb[bi]) * L.zk;
    const int k = min(L.zk, L.lz - z0);
    const int RS = nx + 9, CS = ny + 9, R = k * ny, C = k * nx;
    const float* src = L.in + (size_t)z0 * nx * ny;
    float* dst = L.b + (size_t)z0 * nx * ny;
    const int tid = threadIdx.x;
    const Div dnx(nx), dny(ny), dR(R), dC(C);
    xy_load_rows(src, nx, R, [&](int r, int x, float v) { buf[r * RS + 4 + x] = v; });
    __syncthreads();
    xy_fill_mirror(buf, R, RS, nx);
    __syncthreads();
    // x analysis: pair q of row r, consecutive lanes on consecutive rows
    float lo[NIF], hi[NIF];
#pragma unroll
    for (int j = 0; j < NIF; j++) {
        const int it = tid + XYT * j;
        lo[j] = hi[j] = 0.0f;
        if (it < R * mx) {
            const int q = dR.q(it), r = it - q * R;
            Fwd97::pair(&buf[r * RS], q, lo[j], hi[j]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NIF; j++) {
        const int it = tid + XYT * j;
        if (it < R * mx) {
            const int q = dR.q(it), r = it - q * R;
            const int s = dny.q(r), y = r - s * ny, c = s * nx + q;
            buf[c * CS + 4 + y] = lo[j];
            buf[(c + mx) * CS + 4 + y] = hi[j];
        }
    }
    __syncthreads();
    xy_fill_mirror(buf, C, CS, ny);
    __syncthreads();
    // y analysis: pair q of column c, consecutive lanes on consecutive x (coalesced stores)
    for (int it = tid; it < C * my; it += XYT) {
        const int q = dC.q(it), c = it - q * C;
        const int s = dnx.q(c), x = c - s * nx;
        float l, h;
        Fwd97::pair(&buf[c * CS], q, l, h);
        gst(dst, (s * ny + q) * nx + x, l);
        gst(dst, (s * ny + my + q) * nx + x, h);
    }
}

// Inverse y then x synthesis of zk slices: L.a -> L.out.
__global__ __launch_bounds__(XYT) void k_binv_yx(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    extern __shared__ float buf[];
    const uint32_t bi = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[bi];
    const int nx = L.lx, ny = L.ly, mx = nx / 2, my = ny / 2;
    const int z0 = (int)(blockIdx.x - tb[bi]) * L.zk;
    const int k = min(L.zk, L.lz - z0);
    const int RS = nx + 9, CS = ny + 9, R = k * ny, C = k * nx;
    const float* src = L.a + (size_t)z0 * nx * ny;
    float* dst = L.out + (size_t)z0 * nx * ny;
    const int tid = threadIdx.x;
    const Div dnx(nx), dny(ny), dC(C);
    // columns, interleaved: row y of a slice holds low (y < my) or high coefficient y - my
    xy_load_rows(src, nx, R, [&](int r, int x, float v) {
        const int s = dny.q(r), y = r - s * ny;
        const int w = y < my ? 2 * y : 2 * (y - my) + 1;
        buf[(s * nx + x) * CS + 4 + w] = v;
    });
    __syncthreads();
    xy_fill_inter(buf, C, CS, my);
    __syncthreads();
    // y synthesis: outputs 2 u and 2 u + 1 of column c from one 10-entry window, consecutive lanes
    // on consecutive columns
    float o0[NII], o1[NII];
#pragma unroll
    for (int j = 0; j < NII; j++) {
        const int it = tid + XYT * j;
        o0[j] = o1[j] = 0.0f;
        if (it < C * my) {
            const int u = dC.q(it), c = it - u * C;
            float w[10];
#pragma unroll
            for (int i = 0; i < 10; i++) w[i] = buf[c * CS + 2 * u + i];
            o0[j] = Inv97::out(w, 0);
            o1[j] = Inv97::out(w + 1, 1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NII; j++) {
        const int it = tid + XYT * j;
        if (it < C * my) {
            const int u = dC.q(it), c = it - u * C;
            const int s = dnx.q(c), x = c - s * nx;
            const int w = x < mx ? 2 * x : 2 * (x - mx) + 1;
            const int r = s * ny + 2 * u;
            buf[r * RS + 4 + w] = o0[j];
            buf[(r + 1) * RS + 4 + w] = o1[j];
        }
    }
    __syncthreads();
    xy_fill_inter(buf, R, RS, mx);
    __syncthreads();
    // x synthesis: outputs 2 u and 2 u + 1 of row r, plane elements 2 it and 2 it + 1
    const Div dmx(mx);
    const __amdgpu_buffer_rsrc_t rdst = buffer_of(dst, (uint32_t)(R * nx) * 4u);
    for (int it = tid; it < R * mx; it += XYT) {
        const int r = dmx.q(it), u = it - r * mx;
        float w[10];
#pragma unroll
        for (int i = 0; i < 10; i++) w[i] = buf[r * RS + 2 * u + i];
        const float a = Inv97::out(w, 0), b = Inv97::out(w + 1, 1);
        // outputs 2 it and 2 it + 1 are adjacent: one 8-byte store
        typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 ab = {__float_as_uint(a), __float_as_uint(b)};
        __builtin_amdgcn_raw_buffer_store_b64(ab, rdst, (int)((uint32_t)it * 8u), 0, 0);
    }
}

}  // namespace

void bricks_forward_level(const BrickLevel* d_levels, const uint32_t* d_tiles, uint32_t nbricks,
                          const uint32_t* h_tiles_total, int symSize, uint64_t level_elems, hipStream_t st) {
    const uint32_t* tx = d_tiles;
    const uint32_t* ty = d_tiles + (nbricks + 1);
    const uint32_t* tz = d_tiles + 2 * (nbricks + 1);
    const uint32_t* txy = d_tiles + 3 * (nbricks + 1);
    if (h_tiles_total[3]) {
        CCA_LAUNCH("bricks_fwd_xy", level_elems * 8, k_bfwd_xy, dim3(h_tiles_total[3]), dim3(XYT), XYB * sizeof(float),
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
        CCA_LAUNCH("bricks_inv_yx", level_elems * 8, k_binv_yx, dim3(h_tiles_total[3]), dim3(XYT), XYB * sizeof(float),
                   st, d_levels, d_tiles + 3 * (nbricks + 1), nbricks);
    } else {
        if (h_tiles_total[1]) CCA_LAUNCH("bricks_inv_y", level_elems * 8, k_binv<1>, dim3(h_tiles_total[1]), dim3(256), 0, st, d_levels, ty, nbricks);
        CCA_LAUNCH_CHECK();
        if (h_tiles_total[0]) CCA_LAUNCH("bricks_inv_x", level_elems * 8, k_binv<0>, dim3(h_tiles_total[0]), dim3(256), 0, st, d_levels, tx, nbricks);
    }
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
