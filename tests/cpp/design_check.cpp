// CPU check of the table-design restatement used by the GPU kernels (csrc/hip/huff_design.hpp)
// against the C++ standard library calls the reference makes (std::priority_queue, std::sort),
// on tie-heavy random histograms and adversarial sort inputs.  Built and run by
// tests/test_design_host.py; prints "OK <cases>" on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <queue>
#include <random>
#include <vector>

#include "../../cudacompress_amd/csrc/hip/huff_design.hpp"

using namespace cca::design;

namespace {

struct Node {
    uint32_t sym, count;
    int depth;
    Node* c[2];
};

struct Rarer {
    bool operator()(const Node* a, const Node* b) const { return b->count < a->count; }
};

// The reference's design with the standard library (HuffmanTable.cu:487-586).
struct RefTable {
    std::vector<uint32_t> syms, cpl, code, len;
    uint64_t bits = 0;
    bool too_long = false;
};
RefTable ref_design(const std::vector<uint32_t>& hist, int dialect, uint32_t quirk) {
    RefTable t;
    std::vector<Node> pool;
    pool.reserve(2 * hist.size() + 2);
    std::vector<Node*> leaves;
    for (uint32_t s = 0; s < hist.size(); s++)
        if (hist[s]) pool.push_back(Node{s, hist[s], -1, {nullptr, nullptr}});
    for (auto& n : pool) leaves.push_back(&n);
    t.code.assign(hist.size(), 0);
    t.len.assign(hist.size(), 0xFFFFFFFFu);
    if (leaves.empty()) return t;
    using Q = std::priority_queue<Node*, std::vector<Node*>, Rarer>;
    Q q;
    if (dialect == kDialectGpu) q = Q(leaves.begin(), leaves.end());
    else for (Node* l : leaves) q.push(l);
    while (q.size() > 1) {
        Node* a = q.top(); q.pop();
        Node* b = q.top(); q.pop();
        pool.push_back(Node{0xFFFFFFFFu, a->count + b->count, -1, {a, b}});
        q.push(&pool.back());
    }
    std::vector<Node*> st{q.top()};
    q.top()->depth = 0;
    int deepest = 0;
    while (!st.empty()) {
        Node* n = st.back(); st.pop_back();
        deepest = std::max(deepest, n->depth);
        for (Node* c : n->c) if (c) { c->depth = n->depth + 1; st.push_back(c); }
    }
    std::sort(leaves.begin(), leaves.end(), [](const Node* a, const Node* b) { return a->depth < b->depth; });
    if (deepest > 31) { t.too_long = true; return t; }
    t.cpl.assign(deepest, 0);
    uint32_t code = 0, prev = leaves[0]->depth;
    for (size_t i = 0; i < leaves.size(); i++) {
        const uint32_t l = leaves[i]->depth;
        if (l > 0) t.cpl[l - 1]++;
        if (i > 0) code = (code + 1) << (l - prev);
        prev = l;
        t.syms.push_back(leaves[i]->sym);
        t.code[leaves[i]->sym] = code;
        t.len[leaves[i]->sym] = l;
        t.bits += (uint64_t)(leaves[i]->count - ((quirk && leaves[i]->sym == 65535u) ? 1u : 0u)) * l;
    }
    return t;
}

template <class Idx>
bool check_one(const std::vector<uint32_t>& hist, int dialect, uint32_t quirk, const char* what, int form) {
    const uint32_t nb = (uint32_t)hist.size();
    RefTable r = ref_design(hist, dialect, quirk);
    uint32_t K = 0;
    for (uint32_t h : hist) K += h ? 1 : 0;
    std::vector<unsigned char> heap(heap_region_bytes<Idx>(K) + 16);
    std::vector<Idx> parent(2 * K + 2);
    std::vector<int> stack(3 * kStackRanges);
    Work<Idx> w{reinterpret_cast<HeapEntry*>(heap.data()), parent.data(), stack.data()};
    Ctl ctl{};
    std::vector<uint32_t> code(nb, 0xDEADBEEFu), len(nb, 0xDEADBEEFu), syms(K + 1), cpl(32), info(3);
    TableOut out{code.data(), len.data(), syms.data(), cpl.data(), info.data()};
    uint64_t bits = 0;
    // the one-table-per-thread form serves tables of <= kLaneMaxK symbols with small counts
    uint64_t total = 0;
    for (uint32_t h : hist) total += h;
    if (form == 3 && (K > kLaneMaxK || total >= kLaneMaxCount)) form = 0;
    std::vector<uint32_t> lheap(3 * kLaneMaxK + 3, 0xA5A5A5A5u);
    std::vector<uint16_t> lpar(3 * 2 * kLaneMaxK + 3, 0xA5A5u);
    if (form == 3) {
        // the leaves (symbol, count) in symbol order; stride 3 (the device interleaves the threads
        // of a workgroup), offset 1
        struct U2 { uint32_t x, y; };
        std::vector<U2> leaves;
        for (uint32_t s = 0; s < nb; s++)
            if (hist[s]) leaves.push_back(U2{s, hist[s]});
        std::vector<int> lstack(3 * lane_stack_ints(kLaneMaxK) + 3, -7);
        design_table_lane([&](int r) { return leaves[r]; }, (int)K, dialect, quirk,
                          Strided<uint32_t>{lheap.data() + 1, 3}, Strided<uint16_t>{lpar.data() + 1, 3},
                          Strided<int>{lstack.data() + 1, 3}, out, &bits);
    } else if (form == 2) {
        WalkEx ex;
        design_table(ex, hist.data(), nb, dialect, quirk, w, &ctl, out, &bits);
    } else if (form == 1) {
        PathEx ex;
        design_table(ex, hist.data(), nb, dialect, quirk, w, &ctl, out, &bits);
    } else {
        SerialEx ex;
        design_table(ex, hist.data(), nb, dialect, quirk, w, &ctl, out, &bits);
    }
    if (r.too_long) {
        if (info[2] != 5) { printf("FAIL %s: expected codeword-too-long\n", what); return false; }
        return true;
    }
    bool ok = info[2] == 0 && info[1] == K && info[0] == r.cpl.size() && bits == r.bits;
    for (uint32_t i = 0; ok && i < info[0]; i++) ok = cpl[i] == r.cpl[i];
    for (uint32_t i = 0; ok && i < K; i++) ok = syms[i] == r.syms[i];
    for (uint32_t s = 0; ok && s < nb; s++)
        if (hist[s]) ok = code[s] == r.code[s] && len[s] == r.len[s];
    if (!ok) {
        printf("FAIL %s (form %d): K=%u nb=%u dialect=%d nlen %u/%zu bits %llu/%llu\n", what, form, K, nb, dialect, info[0],
               r.cpl.size(), (unsigned long long)bits, (unsigned long long)r.bits);
        return false;
    }
    return true;
}

// the wave introsort + a stable sort reproduces std::sort for a generic key (exercises the
// depth-limit heap sort with a median-of-3 adversary)
bool check_sort(std::vector<uint32_t> v, uint32_t keymask, const char* what) {
    std::vector<uint32_t> want = v;
    auto key = [keymask](uint32_t e) { return e & keymask; };
    std::sort(want.begin(), want.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    const int n = (int)v.size();
    std::vector<uint32_t> ge(n + 1), le(n + 1);
    std::vector<int> stack(3 * kStackRanges);
    Ctl ctl{};
    SerialEx ex;
    introsort_loop_wave(ex, v.data(), n, key, ge.data(), le.data(), stack.data(), &ctl);
    std::stable_sort(v.begin(), v.end(), [&](uint32_t a, uint32_t b) { return key(a) < key(b); });
    if (v != want) { printf("FAIL sort %s n=%d\n", what, n); return false; }
    return true;
}

// median-of-3 killer (Musser): drives introsort into its depth limit
std::vector<uint32_t> med3_killer(int n) {
    std::vector<uint32_t> a(n);
    const int k = n / 2;
    for (int i = 1; i <= k; i++) {
        if (i % 2 == 1) { a[i - 1] = i; a[i] = k + i; }
        a[k + i - 1] = 2 * i;
    }
    for (int i = 0; i < n; i++) a[i] = (a[i] << 8) | (uint32_t)(i & 0xFF);  // tags below the key
    return a;
}

}  // namespace

int main(int argc, char** argv) {
    const int trials = argc > 1 ? atoi(argv[1]) : 400;
    std::mt19937_64 rng(12345);
    int cases = 0;
    bool ok = true;
    // fixed small cases
    std::vector<std::vector<uint32_t>> fixed = {{5}, {0, 0, 3}, {1, 1}, {1, 1, 2, 2}, {4, 4, 4, 4, 4, 4, 4, 4},
                                                {0, 7, 0, 7, 0, 7}, {1, 2, 4, 8, 16, 32, 64, 128}};
    for (auto& h : fixed)
        for (int d = 0; d < 2; d++) {
            for (int pf : {0, 1, 2, 3}) {
                ok &= check_one<uint16_t>(h, d, 0, "fixed16", pf);
                ok &= check_one<uint32_t>(h, d, 0, "fixed32", pf);
                cases += 2;
            }
        }
    // Fibonacci counts: a codeword longer than 31 bits
    {
        std::vector<uint32_t> h{1, 1};
        while (h.size() < 40) h.push_back(h[h.size() - 1] + h[h.size() - 2]);
        ok &= check_one<uint32_t>(h, 0, 0, "fib", 0) && check_one<uint32_t>(h, 0, 0, "fib", 1) &&
              check_one<uint32_t>(h, 0, 0, "fib", 2) && check_one<uint32_t>(h, 0, 0, "fib", 3);
        cases += 4;
    }
    for (int t = 0; t < trials && ok; t++) {
        const int shape = t % 6;
        uint32_t nb = 1 + (uint32_t)(rng() % (shape == 5 ? 20000 : 3000));
        std::vector<uint32_t> h(nb, 0);
        for (auto& x : h) {
            switch (shape) {
                case 0: x = (uint32_t)(rng() % 3); break;                               // heavy ties
                case 1: x = (rng() % 4 == 0) ? (uint32_t)(rng() % 5) : 0u; break;      // sparse ties
                case 2: x = (uint32_t)(rng() % 1000); break;
                case 3: x = 1u << (rng() % 12); break;                                // powers of two
                case 4: x = (uint32_t)(1 + rng() % 2); break;
                default: x = (uint32_t)(rng() % 40 == 0 ? rng() % 100000 : rng() % 3); break;
            }
        }
        const int dialect = (int)(rng() % 2);
        uint32_t quirk = 0;
        if (nb == 65536 || (t % 50 == 0)) {
            h.resize(65536, 0);
            h[65535] += 1;
            quirk = 1;
        }
        ok &= check_one<uint32_t>(h, dialect, quirk, "rand32", t % 3);
        uint32_t K = 0;
        for (uint32_t x : h) K += x ? 1 : 0;
        if (K <= 32768) ok &= check_one<uint16_t>(h, dialect, quirk, "rand16", (t + 1) % 3);
        cases += 2;
    }
    // small tables (the one-table-per-thread form): up to kLaneMaxK symbols over wide bin ranges,
    // tie-heavy and skewed counts, both dialects, the histogram pad quirk
    for (int t = 0; t < 6 * trials && ok; t++) {
        const uint32_t K = 1 + (uint32_t)(rng() % kLaneMaxK);
        const uint32_t nb = K + (uint32_t)(rng() % (t % 3 == 0 ? 4000 : 300));
        std::vector<uint32_t> h(nb, 0);
        std::vector<uint32_t> pos(nb);
        for (uint32_t i = 0; i < nb; i++) pos[i] = i;
        std::shuffle(pos.begin(), pos.end(), rng);
        for (uint32_t i = 0; i < K; i++) {
            switch (t % 5) {
                case 0: h[pos[i]] = 1 + (uint32_t)(rng() % 3); break;
                case 1: h[pos[i]] = 1 + (uint32_t)(rng() % 2000); break;
                case 2: h[pos[i]] = 1u << (rng() % 14); break;
                case 3: h[pos[i]] = (uint32_t)(1 + std::exp((double)(rng() % 1000) / 100.0)); break;
                default: h[pos[i]] = i < 8 ? 100000 + (uint32_t)(rng() % 9) : 1 + (uint32_t)(rng() % 4); break;
            }
        }
        uint32_t quirk = 0;
        if (t % 40 == 0) {
            h.resize(65536, 0);
            h[65535] += 1;
            quirk = 1;
        }
        const int dialect = (int)(rng() % 2);
        ok &= check_one<uint16_t>(h, dialect, quirk, "lane", 3);
        cases++;
    }
    // sorts: tie-heavy, sorted, reversed, organ pipe, median-of-3 killers
    for (int t = 0; t < 200 && ok; t++) {
        const int n = 1 + (int)(rng() % 5000);
        std::vector<uint32_t> v(n);
        for (int i = 0; i < n; i++) v[i] = ((uint32_t)(rng() % (1 + t % 31)) << 24) | (uint32_t)i;
        ok &= check_sort(v, 0xFF000000u, "ties");
        std::vector<uint32_t> s(n);
        for (int i = 0; i < n; i++) s[i] = ((uint32_t)i << 8) | (uint32_t)(rng() % 256);
        ok &= check_sort(s, 0xFFFFFF00u, "sorted");
        std::reverse(s.begin(), s.end());
        ok &= check_sort(s, 0xFFFFFF00u, "reversed");
        for (int i = 0; i < n; i++) s[i] = ((uint32_t)std::min(i, n - 1 - i) << 8) | (uint32_t)(i & 255);
        ok &= check_sort(s, 0xFFFFFF00u, "organ");
        cases += 4;
    }
    for (int n : {64, 100, 1000, 4096, 20000}) {
        ok &= check_sort(med3_killer(n), 0xFFFFFF00u, "med3killer");
        cases++;
    }
    if (!ok) return 1;
    printf("OK %d\n", cases);
    return 0;
}
