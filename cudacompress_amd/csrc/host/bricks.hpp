// Batched float bricks (host/bricks.cpp).
#pragma once

#include <functional>
#include <memory>
#include <vector>

#include "../bricks.hpp"
#include "batch.hpp"
#include "instance.hpp"

namespace cca {

struct BrickJob {
    std::vector<cc_brick> bricks;
    uint32_t L = 2;
    int symSize = 2;
    // layout (elements): per-brick work buffers a, b, lowpass buffers [level][brick], symbols
    std::vector<uint64_t> a, b, sym;
    std::vector<std::vector<uint64_t>> lp;
    std::vector<uint64_t> level_elems;
    uint64_t work_floats = 0, sym_elems = 0;
    // device tables
    const BrickLevel* d_levels = nullptr;  // [L][n]
    const uint32_t* d_tiles = nullptr;     // [L][4][n+1]
    std::vector<uint32_t> tiles_total;     // [L][4]
    EncodeJob ej;
    DecodeJob dj;
    BoundJob bj;                  // decode from a resident index (bricks_prepare_bound)
    const DecodeIndex* index = nullptr;
    uint32_t* d_out = nullptr;
    uint64_t out_cap = 0;
};

// decode calls: one bitstream per brick, or when `calls` is null the device results of the
// encode staged in the same context
void bricks_prepare(cc_instance* I, BrickCtx& ctx, BrickJob& job, bool enc, bool dec,
                    const std::vector<CallWords>* calls, hipStream_t st);
// Decode of bricks whose streams are parsed in a resident index: brick i of the job is index
// call index_calls[i] (built by bricks_index_build over the same brick shapes and levels).
void bricks_prepare_bound(cc_instance* I, BrickCtx& ctx, BrickJob& job, const DecodeIndex& X,
                          const std::vector<uint32_t>& index_calls, hipStream_t st);
// Resident index of bricks stored at `calls` (one bitstream per brick).
void bricks_index_build(cc_instance* I, DecodeIndex& X, const std::vector<cc_brick>& bricks, uint32_t L,
                        int symSize, const std::vector<CallWords>& calls, hipStream_t st);
bool bricks_decode_results(const BrickJob& job, const std::vector<DecCallOut>& res, uint32_t* status);
// forward transform and quantization of every brick (the first part of bricks_encode_launch)
void bricks_forward_launch(const BrickJob& job, hipStream_t st);
void bricks_encode_launch(BrickCtx& ctx, const BrickJob& job, hipStream_t st, const LayoutLink* link = nullptr);
void bricks_decode_launch(BrickCtx& ctx, const BrickJob& job, hipStream_t st, bool from_encode);

// A batch split into groups of contiguous bricks, each with its own context and stream (host/bricks.cpp).
struct Groups {
    std::vector<uint32_t> first;  // group g: bricks [first[g], first[g+1])
    std::vector<BrickCtx*> ctx;
    std::vector<BrickJob> jobs;
    void setup(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize, uint32_t* d_out,
               uint64_t cap, uint32_t G, std::vector<std::unique_ptr<BrickCtx>>& pool);
    void fork(hipStream_t from, hipEvent_t ev);
    void join(hipStream_t into);
    void encode_launch(hipStream_t from, hipEvent_t ev);
    void decode_launch(hipStream_t from, hipEvent_t ev, bool from_encode);
    // headroom_div: when nonzero, arenas keep used / headroom_div words of headroom afterwards
    uint64_t encode_run(hipStream_t from, hipEvent_t ev, uint64_t* offsets, uint32_t* status, bool* ok,
                        uint64_t headroom_div);
    bool decode_results(uint32_t* status);
};
uint32_t groups_for(const cc_instance* I, const cc_brick* bricks, uint32_t n);

// offsets: nbricks+1 word offsets into d_out (host); status: per-brick cc_status (may be null).
void compress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t levels, int symSize, uint32_t* d_out,
                     uint64_t cap, uint64_t* offsets, uint32_t* status);
void decompress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t levels, int symSize,
                       const uint32_t* d_in, const uint64_t* offsets, uint32_t* status);

cc_bricks_plan* plan_create(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t levels, int symSize,
                            uint32_t* d_out, uint64_t cap);
void plan_compress(cc_bricks_plan* P, uint64_t* offsets, uint32_t* status);
void plan_decompress(cc_bricks_plan* P, uint32_t* status);
void plan_run(cc_bricks_plan* P, uint64_t* offsets, uint32_t* cstatus, uint32_t* dstatus);
void plan_launch(cc_bricks_plan* P);
void plan_wait(cc_bricks_plan* P, uint64_t* offsets, uint32_t* cstatus, uint32_t* dstatus);
void plan_destroy(cc_bricks_plan* P);
// out[4]: symbols coded (N), zero-run pairs of the RLE streams (M), symbols of the Huffman-only
// streams (N_H), compressed words (C / 4) of the plan's last compression
void plan_coding_stats(cc_bricks_plan* P, uint64_t* out);
// the instance a plan belongs to (its launches are that instance's in the kernel timing)
const void* plan_owner(const cc_bricks_plan* P);

}  // namespace cca
