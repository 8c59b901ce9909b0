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

template <int AX>
struct Geo {
    int n, m, ncols, lx, ly;
    __device__ Geo(int x, int y, int z) : lx(x), ly(y) {
        n = AX == 0 ? x : (AX == 1 ? y : z);
        m = n / 2;
        ncols = AX == 0 ? y * z : (AX == 1 ? x * z : x * y);
    }
    // element (column c, position p along the axis) of a dense lx*ly*lz brick
    __device__ long long at(int c, int p) const {
        if (AX == 0) return (long long)c * lx + p;
        if (AX == 1) return (long long)(c % lx) + (long long)p * lx + (long long)(c / lx) * lx * ly;
        return (long long)c + (long long)p * lx * ly;
    }
};

// Forward tile: fetch(c, p) -> float, sink(c, p_out, v).
template <int AX, class Fetch, class Sink>
__device__ void fwd_tile(const Geo<AX>& g, uint32_t local, float (*lds)[LW], Fetch fetch, Sink sink) {
    const int ctiles = (g.ncols + NC - 1) / NC;
    const int c0 = (int)(local % ctiles) * NC, k0 = (int)(local / ctiles) * SP;
    const int tid = threadIdx.x;
    for (int i = tid; i < NC * WF; i += 256) {
        const int col = AX == 0 ? i / WF : i % NC, r = AX == 0 ? i % WF : i / NC;
        const int c = c0 + col, p = 2 * k0 - Fwd97::H + r;
        if (c < g.ncols && p < g.n + Fwd97::H) lds[col][r] = fetch(c, mirror(p, g.n));
    }
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
    for (int i = tid; i < NC * 2 * SP; i += 256) {
        const int cc = AX == 0 ? i / (2 * SP) : i % NC, j = AX == 0 ? i % (2 * SP) : i / NC;
        const int q = j % SP, k = k0 + q, c = c0 + cc;
        if (c < g.ncols && k < g.m) sink(c, j >= SP ? g.m + k : k, lds[cc][j]);
    }
}

// Inverse tile: fetch(c, coefficient index along the axis), sink(c, output position, v).
template <int AX, class Fetch, class Sink>
__device__ void inv_tile(const Geo<AX>& g, uint32_t local, float (*lds)[LW], Fetch fetch, Sink sink) {
    const int ctiles = (g.ncols + NC - 1) / NC;
    const int c0 = (int)(local % ctiles) * NC, k0 = (int)(local / ctiles) * SP;
    const int tid = threadIdx.x;
    for (int i = tid; i < NC * 2 * NPI; i += 256) {
        const int col = AX == 0 ? i / (2 * NPI) : i % NC, v = AX == 0 ? i % (2 * NPI) : i / NC;
        const int q = v >> 1, gi = k0 - Inv97::G + q, c = c0 + col;
        if (c < g.ncols && gi < g.m + Inv97::G)
            lds[col][v] = fetch(c, (v & 1) ? g.m + ext_high(gi, g.m) : ext_low(gi, g.m));
    }
    __syncthreads();
    const int col = tid & 63;
    float o[2 * SP / 4];
#pragma unroll
    for (int j = 0; j < 2 * SP / 4; j++) {
        const int t = (tid >> 6) + 4 * j;
        o[j] = 0.0f;
        if (2 * k0 + t < g.n && c0 + col < g.ncols) o[j] = Inv97::out(&lds[col][t], t & 1);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 2 * SP / 4; j++) lds[col][(tid >> 6) + 4 * j] = o[j];
    __syncthreads();
    for (int i = tid; i < NC * 2 * SP; i += 256) {
        const int cc = AX == 0 ? i / (2 * SP) : i % NC, t = AX == 0 ? i % (2 * SP) : i / NC;
        const int jo = 2 * k0 + t, c = c0 + cc;
        if (c < g.ncols && jo < g.n) sink(c, jo, lds[cc][t]);
    }
}

// band index of a level coefficient and its position inside the dense band
struct Band {
    int b;
    long long li;
};
__device__ __forceinline__ Band band_of(int x, int y, int z, int hx, int hy, int hz) {
    const int bx = x >= hx, by = y >= hy, bz = z >= hz;
    Band r;
    r.b = bx + 2 * by + 4 * bz;
    r.li = ((long long)(z - bz * hz) * hy + (y - by * hy)) * hx + (x - bx * hx);
    return r;
}

template <int AX>
__global__ __launch_bounds__(256) void k_bfwd(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float lds[NC][LW];
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<AX> g(L.lx, L.ly, L.lz);
    const float* src = AX == 0 ? L.in : L.a;
    float* dst = AX == 0 ? L.a : L.b;
    fwd_tile<AX>(g, blockIdx.x - tb[i], lds, [&](int c, int p) { return src[g.at(c, p)]; },
                 [&](int c, int p, float v) { dst[g.at(c, p)] = v; });
}

template <class S>
__global__ __launch_bounds__(256) void k_bfwd_zq(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float lds[NC][LW];
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<2> g(L.lx, L.ly, L.lz);
    const int hx = L.lx / 2, hy = L.ly / 2, hz = L.lz / 2;
    const long long band = (long long)hx * hy * hz;
    S* hp = static_cast<S*>(L.hp);
    S* lps = static_cast<S*>(L.lpsym);
    fwd_tile<2>(g, blockIdx.x - tb[i], lds, [&](int c, int p) { return L.b[g.at(c, p)]; },
                [&](int c, int p, float v) {
                    const Band bd = band_of(c % L.lx, c / L.lx, p, hx, hy, hz);
                    if (bd.b == 0) {
                        if (L.last) lps[bd.li] = (S)symbolize_i(quantize_f(v, L.inv_lp, 0));
                        else L.lp[bd.li] = v;
                    } else {
                        hp[(bd.b - 1) * band + bd.li] = (S)symbolize_i(quantize_f(v, L.inv_hp, 0));
                    }
                });
}

template <class S>
__global__ __launch_bounds__(256) void k_binv_zs(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float lds[NC][LW];
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<2> g(L.lx, L.ly, L.lz);
    const int hx = L.lx / 2, hy = L.ly / 2, hz = L.lz / 2;
    const long long band = (long long)hx * hy * hz;
    const S* hp = static_cast<const S*>(L.hp);
    const S* lps = static_cast<const S*>(L.lpsym);
    inv_tile<2>(g, blockIdx.x - tb[i], lds,
                [&](int c, int p) -> float {
                    const Band bd = band_of(c % L.lx, c / L.lx, p, hx, hy, hz);
                    if (bd.b == 0)
                        return L.last ? unquantize_i(unsymbolize_u((uint32_t)lps[bd.li]), L.step_lp, 0) : L.lp[bd.li];
                    return unquantize_i(unsymbolize_u((uint32_t)hp[(bd.b - 1) * band + bd.li]), L.step_hp, 0);
                },
                [&](int c, int p, float v) { L.a[g.at(c, p)] = v; });
}

template <int AX>
__global__ __launch_bounds__(256) void k_binv(const BrickLevel* lv, const uint32_t* tb, uint32_t nb) {
    __shared__ float lds[NC][LW];
    const uint32_t i = find_brick(tb, nb, blockIdx.x);
    const BrickLevel L = lv[i];
    const Geo<AX> g(L.lx, L.ly, L.lz);
    const float* src = AX == 1 ? L.a : L.b;
    float* dst = AX == 1 ? L.b : L.out;
    inv_tile<AX>(g, blockIdx.x - tb[i], lds, [&](int c, int p) { return src[g.at(c, p)]; },
                 [&](int c, int p, float v) { dst[g.at(c, p)] = v; });
}

}  // namespace

void bricks_forward_level(const BrickLevel* d_levels, const uint32_t* d_tiles, uint32_t nbricks,
                          const uint32_t* h_tiles_total, int symSize, uint64_t level_elems, hipStream_t st) {
    const uint32_t* tx = d_tiles;
    const uint32_t* ty = d_tiles + (nbricks + 1);
    const uint32_t* tz = d_tiles + 2 * (nbricks + 1);
    if (h_tiles_total[0]) CCA_LAUNCH("bricks_fwd_x", level_elems * 8, k_bfwd<0>, dim3(h_tiles_total[0]), dim3(256), 0, st, d_levels, tx, nbricks);
    CCA_LAUNCH_CHECK();
    if (h_tiles_total[1]) CCA_LAUNCH("bricks_fwd_y", level_elems * 8, k_bfwd<1>, dim3(h_tiles_total[1]), dim3(256), 0, st, d_levels, ty, nbricks);
    CCA_LAUNCH_CHECK();
    const uint64_t zb = level_elems * 4 + level_elems / 8 * 4 + level_elems * 7 / 8 * symSize;
    if (h_tiles_total[2]) {
        if (symSize == 2)
            CCA_LAUNCH("bricks_fwd_z_quant", zb, k_bfwd_zq<uint16_t>, dim3(h_tiles_total[2]), dim3(256), 0, st, d_levels, tz, nbricks);
        else
            CCA_LAUNCH("bricks_fwd_z_quant", zb, k_bfwd_zq<uint32_t>, dim3(h_tiles_total[2]), dim3(256), 0, st, d_levels, tz, nbricks);
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
            CCA_LAUNCH("bricks_inv_z_dequant", zb, k_binv_zs<uint16_t>, dim3(h_tiles_total[2]), dim3(256), 0, st, d_levels, tz, nbricks);
        else
            CCA_LAUNCH("bricks_inv_z_dequant", zb, k_binv_zs<uint32_t>, dim3(h_tiles_total[2]), dim3(256), 0, st, d_levels, tz, nbricks);
    }
    CCA_LAUNCH_CHECK();
    if (h_tiles_total[1]) CCA_LAUNCH("bricks_inv_y", level_elems * 8, k_binv<1>, dim3(h_tiles_total[1]), dim3(256), 0, st, d_levels, ty, nbricks);
    CCA_LAUNCH_CHECK();
    if (h_tiles_total[0]) CCA_LAUNCH("bricks_inv_x", level_elems * 8, k_binv<0>, dim3(h_tiles_total[0]), dim3(256), 0, st, d_levels, tx, nbricks);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
