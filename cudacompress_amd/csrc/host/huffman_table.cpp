// Huffman table design and headers, src/cudaCompress/HuffmanTable.cu:117-266, :470-586 and
// src/cudaCompress/HuffmanDesign.h:13-65 (GPU dialect); src/CPU/HuffmanTableCPU.cpp:22-135
// (CPU dialect).
//
// The tree is built with std::priority_queue over node pointers compared by count only, and the
// leaves are ordered with std::sort by depth only -- the same standard-library calls the
// reference makes, so ties resolve identically (libstdc++).  The GPU dialect seeds the queue
// with the range constructor (make_heap); the CPU dialect pushes leaves one at a time.
#include "huffman_table.hpp"

#include <algorithm>
#include <queue>

namespace cca {

HeaderFormat header_format(int dialect, uint32_t log2) {
    if (dialect == CC_DIALECT_CPU) return HeaderFormat{CC_DIALECT_CPU, 14u, 14u};
    return HeaderFormat{CC_DIALECT_GPU, log2, log2 > 16 ? 5u : 4u};
}

namespace {

struct TreeNode {
    uint32_t symbol;
    uint32_t count;
    int depth;
    TreeNode* child[2];
};

struct ByCountInverted {  // HuffmanDesign.h:47-55: "greater" so the queue top is the rarest
    bool operator()(const TreeNode* a, const TreeNode* b) const { return b->count < a->count; }
};
struct ByDepth {  // HuffmanDesign.h:57-63
    bool operator()(const TreeNode* a, const TreeNode* b) const { return a->depth < b->depth; }
};

// depth of every node below `root`; returns the deepest leaf depth (HuffmanDesign.h:32-44)
int assign_depths(TreeNode* root) {
    int deepest = 0;
    std::vector<TreeNode*> stack{root};
    root->depth = 0;
    while (!stack.empty()) {
        TreeNode* n = stack.back();
        stack.pop_back();
        deepest = std::max(deepest, n->depth);
        for (TreeNode* c : n->child) {
            if (c) {
                c->depth = n->depth + 1;
                stack.push_back(c);
            }
        }
    }
    return deepest;
}

}  // namespace

cc_status design_table(int dialect, const uint32_t* hist, uint32_t symbol_max, EncTable& t) {
    t.count_per_length.clear();
    t.symbols.clear();
    t.symbol_max = symbol_max;
    const uint32_t nsym_space = symbol_max + 1;
    t.code.assign(nsym_space, 0u);
    t.len.assign(nsym_space, 0xFFFFFFFFu);

    std::vector<TreeNode> pool;
    pool.reserve(2 * (size_t)nsym_space);
    std::vector<TreeNode*> leaves;
    for (uint32_t s = 0; s < nsym_space; s++) {
        if (hist[s] == 0) continue;
        pool.push_back(TreeNode{s, hist[s], -1, {nullptr, nullptr}});
    }
    if (pool.empty()) return CC_OK;  // empty stream: empty table
    leaves.reserve(pool.size());
    for (TreeNode& n : pool) leaves.push_back(&n);

    using Queue = std::priority_queue<TreeNode*, std::vector<TreeNode*>, ByCountInverted>;
    Queue q;
    if (dialect == CC_DIALECT_GPU) {
        q = Queue(leaves.begin(), leaves.end());
    } else {
        for (TreeNode* l : leaves) q.push(l);
    }
    while (q.size() > 1) {
        TreeNode* a = q.top();
        q.pop();
        TreeNode* b = q.top();
        q.pop();
        pool.push_back(TreeNode{0xFFFFFFFFu, a->count + b->count, -1, {a, b}});
        q.push(&pool.back());
    }
    const int deepest = assign_depths(q.top());
    std::sort(leaves.begin(), leaves.end(), ByDepth());
    if (deepest > 31) return CC_ERR_CODEWORD_TOO_LONG;

    t.count_per_length.assign((size_t)deepest, 0u);
    t.symbols.resize(leaves.size());
    for (size_t i = 0; i < leaves.size(); i++) {
        if (leaves[i]->depth > 0) t.count_per_length[leaves[i]->depth - 1]++;
        t.symbols[i] = leaves[i]->symbol;
    }
    // canonical codes in codeword-index order (HuffmanTable.cu:553-580)
    uint32_t code = 0, prev_len = leaves[0]->depth > 0 ? (uint32_t)leaves[0]->depth : 0u;
    for (size_t i = 0; i < leaves.size(); i++) {
        const uint32_t l = (uint32_t)leaves[i]->depth;
        if (i > 0) code = (code + 1) << (l - prev_len);
        t.code[leaves[i]->symbol] = code;
        t.len[leaves[i]->symbol] = l;
        prev_len = l;
    }
    return CC_OK;
}

static uint32_t header_symbol_max(const HeaderFormat& f, const EncTable& t) {
    if (f.dialect == CC_DIALECT_GPU) return t.symbol_max;
    uint32_t m = 0;  // CPU dialect: max over the table's symbols (HuffmanTableCPU.cpp:110-114)
    for (uint32_t s : t.symbols) m = std::max(m, s);
    return m;
}

void write_table(const HeaderFormat& f, const EncTable& t, BitWriter& w) {
    w.bits((uint32_t)t.count_per_length.size(), 5);
    for (uint32_t c : t.count_per_length) w.bits(c, f.count_bits);
    w.bits((uint32_t)t.symbols.size(), f.count_bits);
    const uint32_t sb = required_bits(header_symbol_max(f, t));
    w.bits(sb, f.symbits_bits);  // GPU dialect, Symbol16: 16 is written as 0 (HuffmanTable.cu:481)
    for (uint32_t s : t.symbols) w.bits(s, sb);
}

uint64_t table_bits(const HeaderFormat& f, const EncTable& t) {
    const uint32_t sb = required_bits(header_symbol_max(f, t));
    return 5ull + (uint64_t)f.count_bits * (t.count_per_length.size() + 1) + f.symbits_bits +
           (uint64_t)sb * t.symbols.size();
}

bool read_table(const HeaderFormat& f, BitReader& r, DecTable& d) {
    for (int i = 0; i < 32; i++) d.first[i] = d.mn[i] = d.mx[i] = -1;
    const uint32_t nlen = r.bits(5);
    uint32_t counts[32] = {0};
    for (uint32_t i = 0; i < nlen; i++) counts[i] = r.bits(f.count_bits);
    const uint32_t nsym = r.bits(f.count_bits);
    uint32_t sb = r.bits(f.symbits_bits);
    // HuffmanTable.cu:138-139: a stored 0 with more than one symbol means 16 (Symbol16 only).
    // A single-symbol table whose symbol needs 16 bits is read with 0 bits, as the reference does.
    if (f.dialect == CC_DIALECT_GPU && f.symbits_bits == 4 && sb == 0 && nsym > 1) sb = 16;
    if (sb > 32 || !r.ok()) return false;
    d.symbols.resize(nsym);
    for (uint32_t i = 0; i < nsym; i++) d.symbols[i] = r.bits(sb);
    if (!r.ok()) return false;
    if (nsym == 0 || nlen == 0) return true;
    uint64_t total = 0;
    for (uint32_t i = 0; i < nlen; i++) total += counts[i];
    if (total != nsym) return false;
    // canonical codes; first/min/max per length (HuffmanTable.cu:205-266)
    int32_t code = 0;
    uint32_t prev_len = 0, e = 0;
    for (uint32_t l = 1; l <= nlen; l++) {
        if (counts[l - 1] == 0) continue;
        for (uint32_t k = 0; k < counts[l - 1]; k++, e++) {
            if (e > 0) code = (code + 1) << (l - prev_len);
            else code = 0;
            prev_len = l;
            if (k == 0) { d.first[l - 1] = (int32_t)e; d.mn[l - 1] = code; }
            d.mx[l - 1] = code;
        }
    }
    return true;
}

}  // namespace cca
