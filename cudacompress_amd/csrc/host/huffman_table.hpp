// Huffman table design, header serialisation and decode-table construction (host).
#pragma once

#include <cstdint>
#include <vector>

#include "../common.hpp"
#include "bitio.hpp"

namespace cca {

struct HeaderFormat {
    int dialect;          // CC_DIALECT_GPU or CC_DIALECT_CPU
    uint32_t count_bits;  // width of the per-length counts and of the symbol count
    uint32_t symbits_bits;// width of the symbolBits field
};
// GPU dialect: log2 bits, 4 (Symbol16) / 5 (Symbol32) bit symbolBits, HuffmanTable.cu:470-485.
// CPU dialect: 14 / 14, src/CPU/HuffmanTableCPU.cpp:108-135.
HeaderFormat header_format(int dialect, uint32_t log2);

struct EncTable {
    std::vector<uint32_t> count_per_length;  // [length-1]
    std::vector<uint32_t> symbols;           // in codeword-index order
    uint32_t symbol_max = 0;                 // stream max symbol (source of symbolBits)
    std::vector<uint32_t> code, len;         // by symbol, size symbol_max+1 (len 0xFFFFFFFF: unused)
};

// Designs the table of one stream from its histogram (symbol_max+1 bins).
// Returns CC_OK or CC_ERR_CODEWORD_TOO_LONG.
cc_status design_table(int dialect, const uint32_t* hist, uint32_t symbol_max, EncTable& t);
void write_table(const HeaderFormat& f, const EncTable& t, BitWriter& w);
// exact number of header bits write_table produces
uint64_t table_bits(const HeaderFormat& f, const EncTable& t);

struct DecTable {
    int32_t first[32], mn[32], mx[32];  // device layout: first[32] | min[32] | max[32]
    std::vector<uint32_t> symbols;
};
// Parses a header and builds first/min/max per length (HuffmanTable.cu:117-151, :205-266).
// Returns false on a malformed header.
bool read_table(const HeaderFormat& f, BitReader& r, DecTable& d);

}  // namespace cca
