// Brick-batched wavelet passes: descriptors shared by hip/bricks.hip and host/bricks.cpp.
//
// A batch is a list of dense float bricks of arbitrary (even) sizes.  Every pass of every DWT
// level runs as ONE launch over all bricks: workgroup w finds its brick by binary search in the
// per-pass tile prefix, then processes a 64-column x 32-pair tile of that brick along the pass
// axis.  The forward Z pass quantizes on the fly: highpass octants go straight to the brick's
// symbol bands, the lowpass octant to the next level's input (or, on the coarsest level, to the
// lowpass symbol stream).  The inverse Z pass dequantizes on the fly.  The arithmetic per
// coefficient is the single-volume passes' (same fmaf chains, same extensions), so batched and
// per-brick results are bit-identical.
#pragma once

#include <algorithm>

#include "common.hpp"

namespace cca {

struct BrickLevel {
    const float* in;   // forward: level input, dense lx*ly*lz
    float* a;          // work buffers, >= lx*ly*lz floats
    float* b;
    float* lp;         // forward: next level's input (lowpass octant, dense); inverse: this level's lowpass
    void* hp;          // the 7 highpass symbol bands of this level, dense hx*hy*hz each, back to back
    void* lpsym;       // coarsest level: lowpass symbol stream (hx*hy*hz)
    float* out;        // inverse: this level's output, dense lx*ly*lz
    int lx, ly, lz;
    int zk;            // z-slices per workgroup of the fused x+y passes
    int last;          // coarsest level
    float inv_hp, step_hp, inv_lp, step_lp;
};

constexpr int kBrickSP = 32;  // pairs per tile along the pass axis
constexpr int kBrickZP = 32;  // output pairs per column segment of the streaming z passes
constexpr int kBrickCols = 64;

// tiles of one brick for a pass along axis ax (0 x, 1 y, 2 z)
inline uint32_t brick_tiles(int ax, int lx, int ly, int lz) {
    const int n = ax == 0 ? lx : (ax == 1 ? ly : lz);
    const uint64_t cols = ax == 0 ? (uint64_t)ly * lz : (ax == 1 ? (uint64_t)lx * lz : (uint64_t)lx * ly);
    const int sp = ax == 2 ? kBrickZP : kBrickSP;
    return (uint32_t)(((cols + kBrickCols - 1) / kBrickCols) * ((n / 2 + sp - 1) / sp));
}

// Fused x+y passes, streaming along y: a workgroup of kBrickXYThreads threads serves zk z-slices
// of a brick (thread k of a slice: x pair k, i.e. the columns k and lx/2 + k of the x-filtered
// plane) and a segment of kBrickXYSeg output pairs along y.  Each step brings two rows through
// LDS for the x filter while the y filter windows stay in registers.  A level uses them when
// every brick's rows have at least 8 samples and its x pairs fit a workgroup.
constexpr int kBrickXYThreads = 256;
constexpr int kBrickXYSeg = 32;
constexpr int kBrickXYRow = 1024;  // LDS floats of one staged row of every slice (zk (lx + 8) <= 1024)
inline bool brick_xy_fits(int lx) { return lx >= 8 && lx / 2 <= kBrickXYThreads; }
inline uint32_t brick_xy_slices(int lx) { return (uint32_t)(kBrickXYThreads / (lx / 2)); }
inline uint32_t brick_xy_tiles(int ly, int lz, int zk) {
    return div_up((uint64_t)lz, (uint64_t)zk) * div_up((uint64_t)(ly / 2), (uint64_t)kBrickXYSeg);
}

// One level of the batch, forward (X, Y, Z+quantize) or inverse (Z from symbols, Y, X).
// tiles: [4][nbricks+1] prefix of brick_tiles for axes x, y, z and of the fused x+y
// workgroups (brick_xy_tiles per brick); h_tiles_total[3] == 0 selects the separate x and y passes.
void bricks_forward_level(const BrickLevel* d_levels, const uint32_t* d_tiles, uint32_t nbricks,
                          const uint32_t* h_tiles_total, int symSize, uint64_t level_elems, hipStream_t st);
void bricks_inverse_level(const BrickLevel* d_levels, const uint32_t* d_tiles, uint32_t nbricks,
                          const uint32_t* h_tiles_total, int symSize, uint64_t level_elems, hipStream_t st);

}  // namespace cca
