// Separable wavelet transforms for gfx950: CDF 9/7 (fp32) and reversible CDF 5/3 (int16).
//
// Math: src/cudaCompress/util/DWTFloatKernels.cui:14-81 (9/7 analysis), :433-503 (synthesis),
// src/cudaCompress/util/DWTFloatFilters.cuh:42-49 (coefficients), src/cudaCompress/util/
// DWTIntKernels.cui:15-100, :597-687 (5/3 lifting), src/cudaCompress/util/DWTCommon.cuh:54-142
// (extensions, div2/div4).  Pass order and buffer use follow src/cudaCompress/util/DWTFloat.cu:
// 85-463 and DWTInt.cu:79-317 so callers that alias source, buffers and destination the way the
// reference's callers do (src/examples/CompressVolume.cpp:341-347, :503-507) get the same result.
//
// Layout of the work: one pass per axis.  The X pass gives each wave64 one row and walks it in
// chunks of 64 output pairs staged in LDS with a 4-element halo; Y/Z passes give each workgroup
// a 64-column x tile (one coalesced 256 B row per LDS line) and a chunk of 32 output pairs along
// the transformed axis.  9/7 sums are explicit fmaf chains in filter order (no reassociation),
// which is what the reference's `sum += c*v` loop compiles to.
#include "../kernels.hpp"
#include "dwt_ops.hpp"

#include <type_traits>

namespace cca {
namespace {

constexpr int XP = 64;  // pairs per X-pass chunk (one per lane)
constexpr int SP = 32;  // pairs per Y/Z-pass chunk

// Byte images enter the first analysis pass as value - 128 and leave the last synthesis pass
// rounded half up and clamped to [0, 255] (toFloat/fromFloat, src/cudaCompress/util/
// DWTCommon.cuh:16-34).
template <class T>
__device__ __forceinline__ T load_in(T v) { return v; }
__device__ __forceinline__ float load_in(uint8_t v) { return float(v) - 128.0f; }
template <class TOut, class T>
__device__ __forceinline__ TOut store_out(T v) {
    if constexpr (std::is_same<TOut, uint8_t>::value)
        return (uint8_t)fminf(fmaxf(v + 128.5f, 0.0f), 255.0f);
    else
        return v;
}

// ---- X pass (contiguous axis): one wave per row ------------------------------------------------
// LO: only the lowpass half is written (the lowpass-only transforms).
template <class Op, int CH, class TIn = typename Op::T, bool LO = false>
__global__ __launch_bounds__(256) void fwd_x(typename Op::T* __restrict__ dst, const TIn* __restrict__ src,
                                             int n, int rows, int sy, int dRow, int dSlice, int sRow, int sSlice) {
    using T = typename Op::T;
    constexpr int W = 2 * XP + 2 * Op::H;
    __shared__ T lds[4][W + 2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    const bool active = row < rows;
    const int y = active ? row % sy : 0, z = active ? row / sy : 0;
    const TIn* s = src + (long long)z * sSlice + (long long)y * sRow;
    T* d = dst + (long long)z * dSlice + (long long)y * dRow;
    const int m = n / 2;
    for (int k0 = 0; k0 < m; k0 += XP) {
        if (active) {
            for (int i = lane; i < W; i += 64) {
                int p = 2 * k0 - Op::H + i;
                if (p < n + Op::H) lds[wave][i] = load_in(s[(long long)mirror(p, n) * CH]);
            }
        }
        __syncthreads();
        const int k = k0 + lane;
        if (active && k < m) {
            T lo, hi;
            Op::pair(lds[wave], lane, lo, hi);
            d[k] = lo;
            if (!LO) d[m + k] = hi;
        }
        __syncthreads();
    }
}

template <class Op, int CH, class TOut = typename Op::T>
__global__ __launch_bounds__(256) void inv_x(TOut* __restrict__ dst, const typename Op::T* __restrict__ src,
                                             int n, int rows, int sy, int dRow, int dSlice, int sRow, int sSlice) {
    using T = typename Op::T;
    constexpr int W = 2 * (XP + 2 * Op::G);
    __shared__ T lds[4][W + 2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    const bool active = row < rows;
    const int y = active ? row % sy : 0, z = active ? row / sy : 0;
    const T* s = src + (long long)z * sSlice + (long long)y * sRow;
    TOut* d = dst + (long long)z * dSlice + (long long)y * dRow;
    const int m = n / 2;
    for (int k0 = 0; k0 < m; k0 += XP) {
        if (active) {
            for (int q = lane; q < XP + 2 * Op::G; q += 64) {
                int g = k0 - Op::G + q;
                if (g < m + Op::G) {
                    lds[wave][2 * q] = s[ext_low(g, m)];
                    lds[wave][2 * q + 1] = s[m + ext_high(g, m)];
                }
            }
        }
        __syncthreads();
        if (active) {
            if constexpr (Op::G == 2) {
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    int t = lane + 64 * h;
                    int j = 2 * k0 + t;
                    if (j < n) d[(long long)j * CH] = store_out<TOut>(Op::out(lds[wave] + t, t & 1));
                }
            } else {
                int k = k0 + lane;
                if (k < m) {
                    T e, o;
                    Op::pair_out(lds[wave] + 2 * lane, e, o);
                    d[(long long)(2 * k) * CH] = store_out<TOut>(e);
                    d[(long long)(2 * k + 1) * CH] = store_out<TOut>(o);
                }
            }
        }
        __syncthreads();
    }
}

// ---- Y/Z pass (strided axis): 64 x-columns x SP pairs per workgroup ----------------------------
struct PlainFetch {
    template <class T>
    __device__ T get(const T* src, int x, int other, int pos, long long oStride, long long aStride) const {
        return src[x + other * oStride + pos * aStride];
    }
};

template <class Op, bool LO = false>
__global__ __launch_bounds__(256) void fwd_s(typename Op::T* __restrict__ dst, const typename Op::T* __restrict__ src,
                                             int sx, int n, long long aStride, long long oStride) {
    using T = typename Op::T;
    constexpr int W = 2 * SP + 2 * Op::H;
    __shared__ T lds[W][64 + 1];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + tx;
    const int k0 = blockIdx.y * SP;
    const int other = blockIdx.z;
    const int m = n / 2;
    const bool xin = x < sx;
    const T* s = src + other * oStride + x;
    for (int r = ty; r < W; r += 4) {
        int p = 2 * k0 - Op::H + r;
        if (xin && p < n + Op::H) lds[r][tx] = s[mirror(p, n) * aStride];
    }
    __syncthreads();
    T* d = dst + other * oStride + x;
    for (int q = ty; q < SP; q += 4) {
        int k = k0 + q;
        if (xin && k < m) {
            T w[2 * Op::H + 2];
#pragma unroll
            for (int i = 0; i < 2 * Op::H + 2; i++) w[i] = lds[2 * q + i][tx];
            T lo, hi;
            Op::pair(w, 0, lo, hi);
            d[k * aStride] = lo;
            if (!LO) d[(m + k) * aStride] = hi;
        }
    }
}

template <class Op, class Fetch>
__global__ __launch_bounds__(256) void inv_s(typename Op::T* __restrict__ dst, const typename Op::T* __restrict__ src,
                                             int sx, int n, long long aStride, long long oStride,
                                             long long dStrideA, long long dStrideO, Fetch fetch) {
    using T = typename Op::T;
    constexpr int NP = SP + 2 * Op::G;
    __shared__ T lds[2 * NP][64 + 1];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + tx;
    const int k0 = blockIdx.y * SP;
    const int other = blockIdx.z;
    const int m = n / 2;
    const bool xin = x < sx;
    for (int q = ty; q < NP; q += 4) {
        int g = k0 - Op::G + q;
        if (xin && g < m + Op::G) {
            lds[2 * q][tx] = fetch.get(src, x, other, ext_low(g, m), oStride, aStride);
            lds[2 * q + 1][tx] = fetch.get(src, x, other, m + ext_high(g, m), oStride, aStride);
        }
    }
    __syncthreads();
    T* d = dst + other * dStrideO + x;
    if constexpr (Op::G == 2) {
        for (int t = ty; t < 2 * SP; t += 4) {
            int j = 2 * k0 + t;
            if (xin && j < n) {
                T v[9];
#pragma unroll
                for (int i = 0; i < 9; i++) v[i] = lds[t + i][tx];
                d[j * dStrideA] = Op::out(v, t & 1);
            }
        }
    } else {
        for (int q = ty; q < SP; q += 4) {
            int k = k0 + q;
            if (xin && k < m) {
                T v[6];
#pragma unroll
                for (int i = 0; i < 6; i++) v[i] = lds[2 * q + i][tx];
                T e, o;
                Op::pair_out(v, e, o);
                d[(2 * k) * dStrideA] = e;
                d[(2 * k + 1) * dStrideA] = o;
            }
        }
    }
}

// Fused dequantize + first synthesis pass of dwtFloat{2D,3D}InverseFromSymbols
// (src/cudaCompress/util/DWTFloatFromSymbolsKernels.cui:19-464): octant b = bx + 2by (+ 4bz)
// comes from the lowpass array when b == 0, else from dense symbol band hp[b-1], dequantized
// with the deadzone rule (src/cudaCompress/util/Quantize.cuh:103-110).
template <typename S>
struct SymbolFetch {
    const float* lp;
    const S* const* hp;
    long long lpRow, lpSlice;
    int hx, hy, hz;  // half sizes
    int is3d;
    float step;
    __device__ float deq(uint32_t sym) const {
        uint32_t neg = sym & 1u;
        int q = (int)((1u - 2u * neg) * ((sym + neg) / 2u));
        float sg = q < 0 ? -1.0f : (q > 0 ? 1.0f : 0.0f);
        return (float(q) + sg * 0.5f) * step;
    }
    template <class T>
    __device__ float get(const T*, int x, int other, int pos, long long, long long) const {
        int y, z;
        if (is3d) { y = other; z = pos; } else { y = pos; z = 0; }
        int bx = x >= hx, by = y >= hy, bz = is3d ? (z >= hz) : 0;
        int b = bx + 2 * by + 4 * bz;
        if (b == 0) return lp[x + y * lpRow + z * lpSlice];
        int lx = x - bx * hx, ly = y - by * hy, lz = z - bz * hz;
        return deq((uint32_t)hp[b - 1][((long long)lz * hy + ly) * hx + lx]);
    }
};

inline dim3 x_grid(int rows) { return dim3(div_up(rows, 4)); }
inline dim3 s_grid(int sx, int n, int other) { return dim3(div_up(sx, 64), div_up(n / 2, SP), other); }

void check_dims(int sx, int sy, int sz) {
    CCA_REQUIRE(sx >= 2 && sy >= 2 && sz >= 1, CC_ERR_INVALID_ARG, "DWT sizes must be >= 2 (sizeZ >= 1)");
    CCA_REQUIRE(sx % 2 == 0 && sy % 2 == 0 && (sz == 1 || sz % 2 == 0), CC_ERR_INVALID_ARG,
                "DWT sizes must be even");
    CCA_REQUIRE(sz <= 65535, CC_ERR_INVALID_ARG, "sizeZ too large for the launch grid");
}

template <int CH, bool LO = false, class TIn>
void launch_fwd97_x(float* dst, const TIn* src, int sx, int sy, int sz, int dRow, int dSlice, int sRow, int sSlice,
                    hipStream_t st) {
    const uint64_t bytes = (uint64_t)sx * sy * sz * (sizeof(TIn) + (LO ? 2 : 4));
    CCA_LAUNCH("dwt97_fwd_x", bytes, (fwd_x<Fwd97, CH, TIn, LO>), x_grid(sy * sz), dim3(256), 0, st, dst, src, sx,
               sy * sz, sy, dRow, dSlice, sRow, sSlice);
}
template <int CH, class TOut>
void launch_inv97_x(TOut* dst, const float* src, int sx, int sy, int sz, int dRow, int dSlice, int sRow, int sSlice,
                    hipStream_t st) {
    CCA_LAUNCH("dwt97_inv_x", (uint64_t)sx * sy * sz * (4 + sizeof(TOut)), (inv_x<Inv97, CH, TOut>), x_grid(sy * sz),
               dim3(256), 0, st, dst, src, sx, sy * sz, sy, dRow, dSlice, sRow, sSlice);
}
template <bool LO = false, class TIn>
void launch_fwd97_x_ch(int ch, float* dst, const TIn* src, int sx, int sy, int sz, int dRow, int dSlice, int sRow,
                       int sSlice, hipStream_t st) {
    switch (ch) {
        case 1: launch_fwd97_x<1, LO>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
        case 2: launch_fwd97_x<2, LO>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
        case 3: launch_fwd97_x<3, LO>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
        default: launch_fwd97_x<4, LO>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
    }
}
template <class TOut>
void launch_inv97_x_ch(int ch, TOut* dst, const float* src, int sx, int sy, int sz, int dRow, int dSlice, int sRow,
                       int sSlice, hipStream_t st) {
    switch (ch) {
        case 1: launch_inv97_x<1>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
        case 2: launch_inv97_x<2>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
        case 3: launch_inv97_x<3>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
        default: launch_inv97_x<4>(dst, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st); break;
    }
}

// dwtFloatForward<T> (src/cudaCompress/util/DWTFloat.cu:85-201) for float and byte sources.
template <class TIn>
void dwt97_forward_t(float* dst, float* buf2, float* buf1, const TIn* src, int sx, int sy, int sz, int ch, int dRow,
                     int dSlice, int sRow, int sSlice, hipStream_t st);
// dwtFloatInverse<T> (DWTFloat.cu:233-322) for float and byte destinations.
template <class TOut>
void dwt97_inverse_t(TOut* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int ch, int dRow,
                     int dSlice, int sRow, int sSlice, hipStream_t st);
// dwtFloatInverseFromSymbols<T> (DWTFloat.cu:326-463) for float and byte destinations.
template <class TOut>
void dwt97_inverse_from_symbols_t(TOut* dst, float* buf2, float* buf1, const float* lowpass,
                                  const void* const* dd_highpass, int symbolSize, float step, int sx, int sy, int sz,
                                  int ch, int dRow, int dSlice, int lpRow, int lpSlice, hipStream_t st);

}  // namespace

void dwt97_forward(float* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int ch, int dRow,
                   int dSlice, int sRow, int sSlice, hipStream_t st) {
    dwt97_forward_t(dst, buf2, buf1, src, sx, sy, sz, ch, dRow, dSlice, sRow, sSlice, st);
}
void dwt97_forward_from_byte(float* dst, float* buf2, float* buf1, const uint8_t* src, int sx, int sy, int sz, int ch,
                             int dRow, int dSlice, int sRow, int sSlice, hipStream_t st) {
    dwt97_forward_t(dst, buf2, buf1, src, sx, sy, sz, ch, dRow, dSlice, sRow, sSlice, st);
}
void dwt97_inverse(float* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int ch, int dRow,
                   int dSlice, int sRow, int sSlice, hipStream_t st) {
    dwt97_inverse_t(dst, buf2, buf1, src, sx, sy, sz, ch, dRow, dSlice, sRow, sSlice, st);
}
void dwt97_inverse_to_byte(uint8_t* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int ch,
                           int dRow, int dSlice, int sRow, int sSlice, hipStream_t st) {
    dwt97_inverse_t(dst, buf2, buf1, src, sx, sy, sz, ch, dRow, dSlice, sRow, sSlice, st);
}
void dwt97_inverse_from_symbols(float* dst, float* buf2, float* buf1, const float* lowpass,
                                const void* const* dd_highpass, int symbolSize, float step, int sx, int sy, int sz,
                                int ch, int dRow, int dSlice, int lpRow, int lpSlice, hipStream_t st) {
    dwt97_inverse_from_symbols_t(dst, buf2, buf1, lowpass, dd_highpass, symbolSize, step, sx, sy, sz, ch, dRow, dSlice,
                                 lpRow, lpSlice, st);
}
void dwt97_inverse_from_symbols_to_byte(uint8_t* dst, float* buf2, float* buf1, const float* lowpass,
                                        const void* const* dd_highpass, float step, int sx, int sy, int sz, int ch,
                                        int dRow, int dSlice, int lpRow, int lpSlice, hipStream_t st) {
    dwt97_inverse_from_symbols_t(dst, buf2, buf1, lowpass, dd_highpass, 2, step, sx, sy, sz, ch, dRow, dSlice, lpRow,
                                 lpSlice, st);
}

// One level of the 9/7 analysis at double precision (no reference counterpart; the FP64 twin of
// dwtFloat3DForward / dwtFloat2DForward used to state the FP32 transform's error).  dst and src
// share the row and slice pitches and must not overlap.  3D: X src -> dst, Y dst -> buf,
// Z buf -> dst; 2D: X src -> buf, Y buf -> dst.
void dwt97_forward_f64(double* dst, double* buf, const double* src, int sx, int sy, int sz, int row, int slice,
                       hipStream_t st) {
    check_dims(sx, sy, sz);
    CCA_REQUIRE(dst && buf && src, CC_ERR_INVALID_ARG, "null DWT buffer");
    if (row <= 0) row = sx;
    if (slice <= 0) slice = sy * row;
    CCA_REQUIRE(row >= sx && slice >= sy * row, CC_ERR_INVALID_ARG, "pitches smaller than the level");
    const bool do3d = sz > 1;
    const uint64_t bytes = (uint64_t)sx * sy * sz * 16;
    double* xdst = do3d ? dst : buf;
    CCA_LAUNCH("dwt97d_fwd_x", bytes, (fwd_x<Fwd97d, 1>), x_grid(sy * sz), dim3(256), 0, st, xdst, src, sx, sy * sz,
               sy, row, slice, row, slice);
    CCA_LAUNCH_CHECK();
    CCA_LAUNCH("dwt97d_fwd_y", bytes, (fwd_s<Fwd97d>), s_grid(sx, sy, sz), dim3(256), 0, st, do3d ? buf : dst, xdst,
               sx, sy, (long long)row, (long long)slice);
    CCA_LAUNCH_CHECK();
    if (do3d) {
        CCA_LAUNCH("dwt97d_fwd_z", bytes, (fwd_s<Fwd97d>), s_grid(sx, sz, sy), dim3(256), 0, st, dst, buf, sx, sz,
                   (long long)slice, (long long)row);
        CCA_LAUNCH_CHECK();
    }
}

// dwtFloat2DForwardLowpassOnlyFromByte (DWTFloat.cu:18-81, :535-552; kernels in
// src/cudaCompress/util/DWTFloat2DLowpassKernels.cui): the X pass writes only the row lowpass
// halves to buf (row pitch dRow), the Y pass only the column lowpass halves to dst.  The sums are
// the same fmaf chains as the full transform, so dst equals its LL band.
void dwt97_forward_lowpass_only_from_byte(float* dst, float* buf, const uint8_t* src, int sx, int sy, int ch,
                                          int dRow, int sRow, hipStream_t st) {
    check_dims(sx, sy, 1);
    CCA_REQUIRE(ch >= 1 && ch <= 4, CC_ERR_INVALID_ARG, "srcChannelCount must be 1..4");
    CCA_REQUIRE(dst && buf && src, CC_ERR_INVALID_ARG, "null DWT buffer");
    if (dRow <= 0) dRow = sx / 2;
    if (sRow <= 0) sRow = sx * ch;
    CCA_REQUIRE(dRow >= sx / 2, CC_ERR_INVALID_ARG, "dstRowPitch must be >= sizeX / 2");
    launch_fwd97_x_ch<true>(ch, buf, src, sx, sy, 1, dRow, 0, sRow, 0, st);
    CCA_LAUNCH_CHECK();
    CCA_LAUNCH("dwt97_fwd_y_lowpass", (uint64_t)sx / 2 * sy * 6, (fwd_s<Fwd97, true>), s_grid(sx / 2, sy, 1), dim3(256),
               0, st, dst, buf, sx / 2, sy, (long long)dRow, 0LL);
    CCA_LAUNCH_CHECK();
}

// dwtIntForwardLowpassOnly (src/cudaCompress/util/DWTInt.cu:16-76, kernels in
// DWTInt2DLowpassKernels.cui): the LL band of the 2D 5/3 transform.  The reference launches
// only whole 512-column and 64-row tiles; every even size is computed here, identical where the
// reference computes.
void dwt53_forward_lowpass_only(int16_t* dst, int16_t* buf, const int16_t* src, int sx, int sy, int dRow, int sRow,
                                hipStream_t st) {
    check_dims(sx, sy, 1);
    CCA_REQUIRE(dst && buf && src, CC_ERR_INVALID_ARG, "null DWT buffer");
    if (dRow <= 0) dRow = sx / 2;
    if (sRow <= 0) sRow = sx;
    CCA_REQUIRE(dRow >= sx / 2, CC_ERR_INVALID_ARG, "dstRowPitch must be >= sizeX / 2");
    CCA_LAUNCH("dwt53_fwd_x_lowpass", (uint64_t)sx * sy * 3, (fwd_x<Fwd53, 1, int16_t, true>), x_grid(sy), dim3(256), 0,
               st, buf, src, sx, sy, sy, dRow, 0, sRow, 0);
    CCA_LAUNCH_CHECK();
    CCA_LAUNCH("dwt53_fwd_y_lowpass", (uint64_t)sx / 2 * sy * 3, (fwd_s<Fwd53, true>), s_grid(sx / 2, sy, 1), dim3(256), 0,
               st, dst, buf, sx / 2, sy, (long long)dRow, 0LL);
    CCA_LAUNCH_CHECK();
}

namespace {

template <class TIn>
void dwt97_forward_t(float* dst, float* buf2, float* buf1, const TIn* src, int sx, int sy, int sz, int ch, int dRow,
                     int dSlice, int sRow, int sSlice, hipStream_t st) {
    check_dims(sx, sy, sz);
    CCA_REQUIRE(ch >= 1 && ch <= 4, CC_ERR_INVALID_ARG, "srcChannelCount must be 1..4");
    if (dRow <= 0) dRow = sx;
    if (dSlice <= 0) dSlice = sy * dRow;
    if (sRow <= 0) sRow = sx * ch;
    if (sSlice <= 0) sSlice = sy * sRow;
    const bool do3d = sz > 1;
    CCA_REQUIRE(dst && buf1 && src && (!do3d || buf2), CC_ERR_INVALID_ARG, "null DWT buffer");
    // X: src -> buf1, Y: buf1 -> (3D ? buf2 : dst), Z: buf2 -> dst (DWTFloat.cu:113-201)
    launch_fwd97_x_ch(ch, buf1, src, sx, sy, sz, dRow, dSlice, sRow, sSlice, st);
    CCA_LAUNCH_CHECK();
    float* ydst = do3d ? buf2 : dst;
    CCA_LAUNCH("dwt97_fwd_y", (uint64_t)sx * sy * sz * 8, (fwd_s<Fwd97>), s_grid(sx, sy, sz), dim3(256), 0, st, ydst, buf1, sx, sy, (long long)dRow,
                       (long long)dSlice);
    CCA_LAUNCH_CHECK();
    if (do3d) {
        CCA_LAUNCH("dwt97_fwd_z", (uint64_t)sx * sy * sz * 8, (fwd_s<Fwd97>), s_grid(sx, sz, sy), dim3(256), 0, st, dst, buf2, sx, sz,
                           (long long)dSlice, (long long)dRow);
        CCA_LAUNCH_CHECK();
    }
}

template <class TOut>
void dwt97_inverse_t(TOut* dst, float* buf2, float* buf1, const float* src, int sx, int sy, int sz, int ch, int dRow,
                     int dSlice, int sRow, int sSlice, hipStream_t st) {
    check_dims(sx, sy, sz);
    CCA_REQUIRE(ch >= 1 && ch <= 4, CC_ERR_INVALID_ARG, "dstChannelCount must be 1..4");
    if (dRow <= 0) dRow = sx * ch;
    if (dSlice <= 0) dSlice = sy * dRow;
    if (sRow <= 0) sRow = sx;
    if (sSlice <= 0) sSlice = sy * sRow;
    const bool do3d = sz > 1;
    CCA_REQUIRE(dst && buf1 && src && (!do3d || buf2), CC_ERR_INVALID_ARG, "null DWT buffer");
    // Z: src -> buf2, Y: (3D ? buf2 : src) -> buf1, X: buf1 -> dst (DWTFloat.cu:233-322)
    if (do3d) {
        CCA_LAUNCH("dwt97_inv_z", (uint64_t)sx * sy * sz * 8, (inv_s<Inv97, PlainFetch>), s_grid(sx, sz, sy), dim3(256), 0, st, buf2, src, sx, sz,
                           (long long)sSlice, (long long)sRow, (long long)sSlice, (long long)sRow, PlainFetch{});
        CCA_LAUNCH_CHECK();
    }
    const float* ysrc = do3d ? buf2 : src;
    CCA_LAUNCH("dwt97_inv_y", (uint64_t)sx * sy * sz * 8, (inv_s<Inv97, PlainFetch>), s_grid(sx, sy, sz), dim3(256), 0, st, buf1, ysrc, sx, sy,
                       (long long)sRow, (long long)sSlice, (long long)sRow, (long long)sSlice, PlainFetch{});
    CCA_LAUNCH_CHECK();
    launch_inv97_x_ch(ch, dst, buf1, sx, sy, sz, dRow, dSlice, sRow, sSlice, st);
    CCA_LAUNCH_CHECK();
}

template <class TOut>
void dwt97_inverse_from_symbols_t(TOut* dst, float* buf2, float* buf1, const float* lowpass,
                                  const void* const* dd_highpass, int symbolSize, float step, int sx, int sy, int sz,
                                  int ch, int dRow, int dSlice, int lpRow, int lpSlice, hipStream_t st) {
    check_dims(sx, sy, sz);
    CCA_REQUIRE(ch >= 1 && ch <= 4, CC_ERR_INVALID_ARG, "dstChannelCount must be 1..4");
    CCA_REQUIRE(symbolSize == 2 || symbolSize == 4, CC_ERR_INVALID_ARG, "symbolSize must be 2 or 4");
    if (dRow <= 0) dRow = sx * ch;
    if (dSlice <= 0) dSlice = sy * dRow;
    if (lpRow <= 0) lpRow = sx;
    if (lpSlice <= 0) lpSlice = sy * lpRow;
    const bool do3d = sz > 1;
    CCA_REQUIRE(dst && buf1 && lowpass && dd_highpass && (!do3d || buf2), CC_ERR_INVALID_ARG, "null DWT buffer");
    // buffers are dense with pitch sizeX (DWTFloat.cu:355-356)
    const long long bRow = sx, bSlice = (long long)sx * sy;
    auto first_pass = [&](auto fetch) {
        if (do3d) {
            CCA_LAUNCH("dwt97_inv_z_from_symbols", (uint64_t)sx * sy * sz * 4 / 8 + (uint64_t)sx * sy * sz * 7 / 8 * symbolSize + (uint64_t)sx * sy * sz * 4, (inv_s<Inv97, decltype(fetch)>), s_grid(sx, sz, sy), dim3(256), 0, st, buf2,
                               (const float*)nullptr, sx, sz, bSlice, bRow, bSlice, bRow, fetch);
        } else {
            CCA_LAUNCH("dwt97_inv_y_from_symbols", (uint64_t)sx * sy * sz / 4 * 4 + (uint64_t)sx * sy * sz * 3 / 4 * symbolSize + (uint64_t)sx * sy * sz * 4, (inv_s<Inv97, decltype(fetch)>), s_grid(sx, sy, 1), dim3(256), 0, st, buf1,
                               (const float*)nullptr, sx, sy, bRow, bSlice, bRow, bSlice, fetch);
        }
    };
    if (symbolSize == 2) {
        SymbolFetch<uint16_t> f{lowpass, (const uint16_t* const*)dd_highpass, lpRow, lpSlice, sx / 2, sy / 2, sz / 2,
                                do3d ? 1 : 0, step};
        first_pass(f);
    } else {
        SymbolFetch<uint32_t> f{lowpass, (const uint32_t* const*)dd_highpass, lpRow, lpSlice, sx / 2, sy / 2, sz / 2,
                                do3d ? 1 : 0, step};
        first_pass(f);
    }
    CCA_LAUNCH_CHECK();
    if (do3d) {
        CCA_LAUNCH("dwt97_inv_y", (uint64_t)sx * sy * sz * 8, (inv_s<Inv97, PlainFetch>), s_grid(sx, sy, sz), dim3(256), 0, st, buf1, buf2, sx, sy, bRow,
                           bSlice, bRow, bSlice, PlainFetch{});
        CCA_LAUNCH_CHECK();
    }
    launch_inv97_x_ch(ch, dst, buf1, sx, sy, sz, dRow, dSlice, (int)bRow, (int)bSlice, st);
    CCA_LAUNCH_CHECK();
}

}  // namespace

void dwt53_forward(int16_t* dst, int16_t* buf, const int16_t* src, int sx, int sy, int sz, int dRow, int dSlice,
                   int sRow, int sSlice, hipStream_t st) {
    check_dims(sx, sy, sz);
    if (dRow <= 0) dRow = sx;
    if (dSlice <= 0) dSlice = sy * dRow;
    if (sRow <= 0) sRow = sx;
    if (sSlice <= 0) sSlice = sy * sRow;
    CCA_REQUIRE(dst && buf && src, CC_ERR_INVALID_ARG, "null DWT buffer");
    const bool do3d = sz > 1;
    // X: src -> (3D ? dst : buf), Y: -> (3D ? buf : dst), Z: buf -> dst (DWTInt.cu:106-186)
    int16_t* xdst = do3d ? dst : buf;
    CCA_LAUNCH("dwt53_fwd_x", (uint64_t)sx * sy * sz * 4, (fwd_x<Fwd53, 1>), x_grid(sy * sz), dim3(256), 0, st, xdst, src, sx, sy * sz, sy, dRow, dSlice,
                       sRow, sSlice);
    CCA_LAUNCH_CHECK();
    int16_t* ydst = do3d ? buf : dst;
    CCA_LAUNCH("dwt53_fwd_y", (uint64_t)sx * sy * sz * 4, (fwd_s<Fwd53>), s_grid(sx, sy, sz), dim3(256), 0, st, ydst, xdst, sx, sy, (long long)dRow,
                       (long long)dSlice);
    CCA_LAUNCH_CHECK();
    if (do3d) {
        CCA_LAUNCH("dwt53_fwd_z", (uint64_t)sx * sy * sz * 4, (fwd_s<Fwd53>), s_grid(sx, sz, sy), dim3(256), 0, st, dst, buf, sx, sz, (long long)dSlice,
                           (long long)dRow);
        CCA_LAUNCH_CHECK();
    }
}

void dwt53_inverse(int16_t* dst, int16_t* buf, const int16_t* src, int sx, int sy, int sz, int dRow, int dSlice,
                   int sRow, int sSlice, hipStream_t st) {
    check_dims(sx, sy, sz);
    if (dRow <= 0) dRow = sx;
    if (dSlice <= 0) dSlice = sy * dRow;
    if (sRow <= 0) sRow = sx;
    if (sSlice <= 0) sSlice = sy * sRow;
    CCA_REQUIRE(dst && buf && src, CC_ERR_INVALID_ARG, "null DWT buffer");
    const bool do3d = sz > 1;
    // Z: src -> dst, Y: (3D ? dst : src) -> buf, X: buf -> dst (DWTInt.cu:218-316), src pitches
    // for the intermediate buffers
    if (do3d) {
        CCA_LAUNCH("dwt53_inv_z", (uint64_t)sx * sy * sz * 4, (inv_s<Inv53, PlainFetch>), s_grid(sx, sz, sy), dim3(256), 0, st, dst, src, sx, sz,
                           (long long)sSlice, (long long)sRow, (long long)sSlice, (long long)sRow, PlainFetch{});
        CCA_LAUNCH_CHECK();
    }
    const int16_t* ysrc = do3d ? dst : src;
    CCA_LAUNCH("dwt53_inv_y", (uint64_t)sx * sy * sz * 4, (inv_s<Inv53, PlainFetch>), s_grid(sx, sy, sz), dim3(256), 0, st, buf, ysrc, sx, sy,
                       (long long)sRow, (long long)sSlice, (long long)sRow, (long long)sSlice, PlainFetch{});
    CCA_LAUNCH_CHECK();
    CCA_LAUNCH("dwt53_inv_x", (uint64_t)sx * sy * sz * 4, (inv_x<Inv53, 1>), x_grid(sy * sz), dim3(256), 0, st, dst, buf, sx, sy * sz, sy, dRow, dSlice,
                       sRow, sSlice);
    CCA_LAUNCH_CHECK();
}

}  // namespace cca
