// Batched float bricks (host/bricks.cpp).
#pragma once

#include "instance.hpp"

namespace cca {

// offsets: nbricks+1 word offsets into d_out (host); status: per-brick cc_status (may be null).
void compress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t levels, int symSize, uint32_t* d_out,
                     uint64_t cap, uint64_t* offsets, uint32_t* status);
void decompress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t levels, int symSize,
                       const uint32_t* d_in, const uint64_t* offsets, uint32_t* status);

}  // namespace cca
