// Huffman table design, reproduced step for step so that ties resolve as in the reference.
//
// The reference designs its tables on the host (src/cudaCompress/HuffmanTable.cu:487-586,
// src/cudaCompress/HuffmanDesign.h:13-65): a std::priority_queue of leaf pointers compared by
// count only (the rarest on top), seeded with the range constructor (std::make_heap) in the GPU
// dialect or by one push per leaf in the CPU dialect (src/CPU/HuffmanTableCPU.cpp:60-62); the two
// rarest nodes are merged until one remains; then std::sort orders the leaves by depth only.
// Neither the heap nor the sort is stable, so the tree shape and the order of equal-length
// codewords -- hence the bitstream -- depend on the exact element moves of libstdc++'s
// std::make_heap / pop_heap / push_heap (stl_heap.h) and std::sort (introsort with median-of-3,
// unguarded partition, a 2*floor(log2 n) depth limit falling back to heap sort, and a final
// insertion sort with threshold 16; stl_algo.h).  Those algorithms are restated here from their
// published steps.
//
// The same code runs on the GPU (one wave per table, tables in LDS) and on the host (the CPU
// test compares it against the standard library on tie-heavy inputs).  Parallel steps are
// written against an executor with a host (serial) and a device (wave64) implementation:
//  * leaves are collected in symbol order;
//  * the merge loop's heap operations are restated over their descent paths and ancestor
//    chains (pop_by_path, push_by_chain, make_by_levels), which the wave evaluates in parallel;
//  * leaf depths are found by walking parent links, one leaf per lane;
//  * sort: every partition is computed by the whole wave from the closed form of the unguarded
//    partition (partition_cut); the final insertion sort equals a stable sort by depth (every
//    element only moves inside its <=16-element block and blocks are ordered), which is done as a
//    stable counting sort over the 32 possible depths.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CCA_HD __host__ __device__ __forceinline__
#else
#define CCA_HD inline
#endif

namespace cca {
namespace design {

constexpr int kSortThreshold = 16;  // std::sort's _S_threshold
constexpr int kMaxDepth = 31;       // codeword lengths must fit the 5-bit header field

struct HeapEntry {
    uint32_t count;
    uint32_t node;
};

// ---------------------------------------------------------------- heap (stl_heap.h semantics)
// Comparator of the reference's queue (HuffmanDesign.h:47-55): a is "less" than b when b is rarer.
struct QueueLess {
    CCA_HD bool operator()(const HeapEntry& a, const HeapEntry& b) const { return b.count < a.count; }
};
// Comparator of the reference's leaf sort (HuffmanDesign.h:57-63), on packed (depth << 24 | symbol).
struct DepthLess {
    CCA_HD bool operator()(uint32_t a, uint32_t b) const { return (a >> 24) < (b >> 24); }
};

// The heap and sort routines take any array-like H (a pointer, or the strided per-thread view of
// the one-table-per-thread design) with elements of type T.
template <class H, class T, class Less>
CCA_HD void push_heap_at(H h, int hole, int top, T v, Less less) {
    int parent = (hole - 1) / 2;
    while (hole > top && less(h[parent], v)) {
        h[hole] = h[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    h[hole] = v;
}

template <class H, class T, class Less>
CCA_HD void adjust_heap(H h, int hole, int len, T v, Less less) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (less(h[child], h[child - 1])) child--;
        h[hole] = h[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        h[hole] = h[child - 1];
        hole = child - 1;
    }
    push_heap_at(h, hole, top, v, less);
}

template <class H, class Less>
CCA_HD void make_heap(H h, int len, Less less) {
    if (len < 2) return;
    for (int parent = (len - 2) / 2;; parent--) {
        auto v = h[parent];
        adjust_heap(h, parent, len, v, less);
        if (parent == 0) return;
    }
}

// std::pop_heap over h[0, len): the top moves to h[len-1].
template <class H, class Less>
CCA_HD void pop_heap(H h, int len, Less less) {
    if (len > 1) {
        auto v = h[len - 1];
        h[len - 1] = h[0];
        adjust_heap(h, 0, len - 1, v, less);
    }
}

// std::push_heap over h[0, len) whose new element is h[len-1].
template <class H, class Less>
CCA_HD void push_heap(H h, int len, Less less) {
    auto v = h[len - 1];
    push_heap_at(h, len - 1, 0, v, less);
}

// std::partial_sort(first, last, last): heap select (make_heap) then sort_heap.
template <class H, class Less>
CCA_HD void heap_sort(H a, int len, Less less) {
    make_heap(a, len, less);
    while (len > 1) {
        len--;
        auto v = a[len];
        a[len] = a[0];
        adjust_heap(a, 0, len, v, less);
    }
}

// ---------------------------------------------------------------- sort (stl_algo.h semantics)
template <class H>
CCA_HD void swap_elems(H a, int i, int j) {
    auto t = a[i];
    a[i] = a[j];
    a[j] = t;
}

template <class H, class Less>
CCA_HD void move_median_to_first(H a, int result, int x, int y, int z, Less less) {
    if (less(a[x], a[y])) {
        if (less(a[y], a[z])) swap_elems(a, result, y);
        else if (less(a[x], a[z])) swap_elems(a, result, z);
        else swap_elems(a, result, x);
    } else if (less(a[x], a[z])) {
        swap_elems(a, result, x);
    } else if (less(a[y], a[z])) {
        swap_elems(a, result, z);
    } else {
        swap_elems(a, result, y);
    }
}

template <class H, class Less>
CCA_HD int unguarded_partition(H a, int first, int last, int pivot, Less less) {
    for (;;) {
        while (less(a[first], a[pivot])) first++;
        last--;
        while (less(a[pivot], a[last])) last--;
        if (!(first < last)) return first;
        swap_elems(a, first, last);
        first++;
    }
}

// __unguarded_partition_pivot: median of (first+1, mid, last-1) moved to first.
template <class H, class Less>
CCA_HD void place_pivot(H a, int first, int last, Less less) {
    const int mid = first + (last - first) / 2;
    move_median_to_first(a, first, first + 1, mid, last - 1, less);
}

CCA_HD int floor_log2(uint32_t n) {
    int r = 0;
    while (n >>= 1) r++;
    return r;
}

// The serial introsort loop (without the final insertion sort) over [first, last) with the given
// remaining depth.  stack: room for 2*floor(log2 n)+2 ranges (3 ints each).
template <class H, class Less, class Stk>
CCA_HD void introsort_loop(H a, int first, int last, int depth, Less less, Stk stack) {
    int sp = 0;
    for (;;) {
        while (last - first > kSortThreshold) {
            if (depth == 0) {
                heap_sort(a + first, last - first, less);
                break;
            }
            depth--;
            place_pivot(a, first, last, less);
            const int cut = unguarded_partition(a, first + 1, last, first, less);
            stack[sp++] = cut;  // right part [cut, last) with the decremented depth, later
            stack[sp++] = last;
            stack[sp++] = depth;
            last = cut;
        }
        if (sp == 0) return;
        depth = stack[--sp];
        last = stack[--sp];
        first = stack[--sp];
    }
}

// Closed form of unguarded_partition(a, first+1, last, pivot=first) given, in increasing
// position order, the positions ge[0..G) of elements not less than the pivot and le[0..L) of
// elements not greater than it.  The k-th swap exchanges ge[k] with the k-th element of le
// counted from the right, while ge[k] lies left of it; the scan then stops at the next ge
// position or at the last swapped right position, whichever comes first.  nswap = the number
// of k with ge[k] < le[L-1-k]; returns the cut position.
template <class Idx>
CCA_HD int partition_cut(const Idx* ge, int G, const Idx* le, int L, int last, int nswap) {
    int cut = nswap == 0 ? last : (int)le[L - nswap];  // (k-1)-th from the right == le[L-1-(k-1)]
    if (nswap < G && (int)ge[nswap] < cut) cut = (int)ge[nswap];
    return cut;
}

// ---------------------------------------------------------------- canonical codes
// Codes from the count per length, in codeword-index order: c_0 = 0 and
// c_i = (c_{i-1} + 1) << (l_i - l_{i-1}) (HuffmanTable.cu:553-580), i.e. the first code of length
// l is (first(l-1) + count(l-1)) << 1.
CCA_HD void first_codes(const uint32_t* count_per_length, int nlen, uint32_t* first_code, uint32_t* first_index) {
    uint32_t code = 0, idx = 0;
    for (int l = 1; l <= nlen; l++) {
        first_code[l - 1] = code;
        first_index[l - 1] = idx;
        code = (code + count_per_length[l - 1]) << 1;
        idx += count_per_length[l - 1];
    }
}

// ---------------------------------------------------------------- the design driver
// Memory of one design (LDS on the device).  The heap region holds K HeapEntry during the merge
// loop and is then reused for the sort: K packed elements followed by the two position lists.
template <class Idx>
struct Work {
    HeapEntry* heap;   // K entries (8K bytes)
    Idx* parent;       // 2K-1 entries
    int* stack;        // 3 * kStackRanges ints
    // views of the heap region after the merge loop
    CCA_HD uint32_t* elems() const { return reinterpret_cast<uint32_t*>(heap); }
    CCA_HD Idx* ge(int K) const { return reinterpret_cast<Idx*>(elems() + K); }
    CCA_HD Idx* le(int K) const { return ge(K) + K; }
};
// pending ranges never exceed 2*floor(log2 n) + 2 <= 50 for n < 2^24 (their depths strictly decrease)
constexpr int kStackRanges = 64;

// bytes of the heap region needed for K leaves with index type Idx
template <class Idx>
CCA_HD uint64_t heap_region_bytes(uint32_t K) {
    const uint64_t a = 8ull * K, b = 4ull * K + 2ull * sizeof(Idx) * K;
    return a > b ? a : b;
}

// Shared scalars of a design (LDS on the device).
struct Ctl {
    int K, size, root, deepest, err;
    int first, last, depth, sp;
    int G, L, T;
};

// Result of one design.
struct TableOut {
    uint32_t* code;   // by symbol, smax+1 entries (only used symbols are written)
    uint32_t* len;    // by symbol
    uint32_t* syms;   // symbols in codeword-index order (K entries)
    uint32_t* cpl;    // count per length, 32 entries (first nlen valid)
    uint32_t* info;   // [0] nlen, [1] nsym, [2] error (0 ok, 5 codeword too long)
};

enum { kDialectGpu = 0, kDialectCpu = 1 };

template <class KeyF>
struct KeyLess {
    KeyF key;
    template <class T>
    CCA_HD bool operator()(const T& a, const T& b) const { return key(a) < key(b); }
};

// std::sort's introsort loop (without the final insertion sort) over el[0, n), comparing key(e).
// Partitions are computed by the whole executor from the closed form above; ranges whose depth
// budget runs out are heap-sorted by lane 0.  ge/le: n entries each; stack: 3*kStackRanges ints.
template <class Ex, class T, class KeyF, class Idx>
CCA_HD void introsort_loop_wave(Ex& ex, T* el, int n, KeyF key, Idx* ge, Idx* le, int* stack, Ctl* ctl) {
    if (n <= kSortThreshold) return;
    const int lane = ex.lane(), nl = ex.lanes();
    const KeyLess<KeyF> less{key};
    if (lane == 0) {
        stack[0] = 0; stack[1] = n; stack[2] = 2 * floor_log2((uint32_t)n);
        ctl->sp = 1;
    }
    ex.barrier();
    for (;;) {
        if (lane == 0) {
            const int sp = --ctl->sp;
            ctl->first = stack[3 * sp];
            ctl->last = stack[3 * sp + 1];
            ctl->depth = stack[3 * sp + 2];
            if (ctl->last - ctl->first > kSortThreshold) {
                if (ctl->depth == 0) {
                    heap_sort(el + ctl->first, ctl->last - ctl->first, less);
                    ctl->last = ctl->first;  // this range is finished
                } else {
                    ctl->depth--;
                    place_pivot(el, ctl->first, ctl->last, less);
                }
            }
        }
        ex.barrier();
        const int first = ctl->first, last = ctl->last, depth = ctl->depth;
        if (last - first > kSortThreshold) {
            const auto p = key(el[first]);
            const int G = ex.compact(first + 1, last, [&](int i) { return !(key(el[i]) < p); },
                                     [&](int i, int r) { ge[r] = (Idx)i; });
            const int L = ex.compact(first + 1, last, [&](int i) { return !(p < key(el[i])); },
                                     [&](int i, int r) { le[r] = (Idx)i; });
            ex.barrier();
            const int mn = G < L ? G : L;
            const int nsw = ex.count(0, mn, [&](int t) { return (int)ge[t] < (int)le[L - 1 - t]; });
            for (int t = lane; t < nsw; t += nl) swap_elems(el, (int)ge[t], (int)le[L - 1 - t]);
            ex.barrier();
            if (lane == 0) {
                const int cut = partition_cut(ge, G, le, L, last, nsw);
                int sp = ctl->sp;
                stack[3 * sp] = cut; stack[3 * sp + 1] = last; stack[3 * sp + 2] = depth; sp++;
                stack[3 * sp] = first; stack[3 * sp + 1] = cut; stack[3 * sp + 2] = depth; sp++;
                ctl->sp = sp;
            }
            ex.barrier();
        }
        const int sp = ctl->sp;
        ex.barrier();
        if (sp == 0) return;
    }
}

// Designs one table.  Every lane of the executor calls this with the same arguments.
// Returns nothing; results in `out`, Σ hist*len (minus one codeword when `quirk`) in *bits (lane 0).
template <class Ex, class Idx>
CCA_HD void design_table(Ex& ex, const uint32_t* hist, uint32_t nbins, int dialect, uint32_t quirk, Work<Idx> w,
                         Ctl* ctl, TableOut out, uint64_t* bits) {
    const int lane = ex.lane(), nl = ex.lanes();
    // (1) leaves in symbol order
    int K = ex.compact(0, (int)nbins, [&](int s) { return hist[s] != 0u; },
                       [&](int s, int rank) { w.heap[rank] = HeapEntry{hist[s], (uint32_t)rank}; });
    ex.barrier();
    if (K == 0) {
        if (lane == 0) {
            out.info[0] = 0; out.info[1] = 0; out.info[2] = 0;
            *bits = 0;
        }
        return;
    }
    // (2) merge loop (every lane runs it; the executor parallelises each heap operation)
    {
        if (dialect == kDialectGpu) {
            ex.heap_make(w.heap, K);
        } else {
            for (int i = 1; i < K; i++) ex.heap_push(w.heap, i, w.heap[i], HeapEntry{0u, 0u});
        }
        int size = K, next = K;
        ex.barrier();
        HeapEntry top = w.heap[0];  // the executor's pop and push return the new top
        while (size > 1) {
            const HeapEntry a = top;
            top = ex.heap_pop(w.heap, size--);
            const HeapEntry b = top;
            top = ex.heap_pop(w.heap, size--);
            if (lane == 0) {
                w.parent[a.node] = (Idx)next;
                w.parent[b.node] = (Idx)next;
            }
            top = ex.heap_push(w.heap, size++, HeapEntry{a.count + b.count, (uint32_t)next}, top);
            next++;
        }
        if (lane == 0) {
            ctl->root = next - 1;
            ctl->deepest = 0;
            ctl->err = 0;
        }
    }
    ex.barrier();
    // (3) leaf depths, packed with the symbol into the (now free) heap region
    const int root = ctl->root;
    uint32_t* el = w.elems();
    int deepest = 0;
    ex.compact(0, (int)nbins, [&](int s) { return hist[s] != 0u; }, [&](int s, int rank) {
        int d = 0;
        for (int x = rank; x != root; x = (int)w.parent[x]) d++;
        deepest = d > deepest ? d : deepest;
        el[rank] = ((uint32_t)(d > kMaxDepth ? kMaxDepth : d) << 24) | (uint32_t)s;
    });
    deepest = ex.max_all(deepest);
    ex.barrier();
    if (deepest > kMaxDepth) {
        if (lane == 0) { out.info[0] = 0; out.info[1] = 0; out.info[2] = 5; *bits = 0; }
        return;
    }
    // (4) introsort loop over the depths
    introsort_loop_wave(ex, el, K, [](uint32_t e) { return e >> 24; }, w.ge(K), w.le(K), w.stack, ctl);
    // (5) final insertion sort == stable sort by depth: counting sort into out.syms; (6) codes
    uint32_t* cpl = out.cpl;
    for (int l = lane; l < 32; l += nl) cpl[l] = 0;
    ex.barrier();
    uint64_t b = 0;
    if (deepest == 0) {  // a single leaf: no codeword bits
        if (lane == 0) {
            out.syms[0] = el[0] & 0xFFFFFFu;
            out.code[el[0] & 0xFFFFFFu] = 0;
            out.len[el[0] & 0xFFFFFFu] = 0;
        }
    } else {
        ex.stable_rank_by_depth(el, K, cpl, [&](int i, uint32_t e, uint32_t rank_in_depth, const uint32_t* first_idx,
                                                 const uint32_t* first_code) {
            const uint32_t d = e >> 24, s = e & 0xFFFFFFu;
            const uint32_t dest = first_idx[d - 1] + rank_in_depth;
            out.syms[dest] = s;
            out.code[s] = first_code[d - 1] + rank_in_depth;
            out.len[s] = d;
            // the pad symbol counted by the histogram quirk is not encoded
            b += (uint64_t)(hist[s] - ((quirk && s == 65535u) ? 1u : 0u)) * d;
        });
    }
    b = ex.sum_all(b);
    if (lane == 0) {
        out.info[0] = (uint32_t)deepest;
        out.info[1] = (uint32_t)K;
        out.info[2] = 0;
        *bits = b;
    }
}

// ---------------------------------------------------------------- one table per thread
// Tables of at most kLaneMaxK symbols whose sequence has fewer than kLaneMaxCount symbols are
// designed by one thread each, with the serial steps above (the same moves as the reference's
// libstdc++ calls): the heap holds packed 32-bit entries count << kLaneNodeBits | node (a count
// never exceeds the sequence length, a node index stays below 2 kLaneMaxK), the parent links are
// 16-bit node indices, and both arrays are strided views, which the device lays out interleaved
// across the threads of a workgroup (element j of thread t at j * threads + t) so that threads at
// the same step of their tables touch distinct LDS banks.
constexpr uint32_t kLaneMaxK = 256;
constexpr int kLaneNodeBits = 9;
constexpr uint32_t kLaneMaxCount = 1u << (32 - kLaneNodeBits);
static_assert(2 * kLaneMaxK <= (1u << kLaneNodeBits), "node indices fit the packed entry");

template <class T>
struct Strided {
    T* base;
    int stride;
    CCA_HD T& operator[](int j) const { return base[j * stride]; }
    CCA_HD Strided operator+(int j) const { return Strided{base + j * stride, stride}; }
};

// QueueLess on packed entries: the counts decide, ties are equal.
struct PackedQueueLess {
    CCA_HD bool operator()(uint32_t a, uint32_t b) const { return (b >> kLaneNodeBits) < (a >> kLaneNodeBits); }
};

// Designs one table (same results as design_table) from its K leaves: leaf(r) is the (symbol,
// count) of the r-th used symbol in symbol order (K <= kLaneMaxK, counts summing below
// kLaneMaxCount); heap has room for kLaneMaxK entries, parent for 2 kLaneMaxK, stack for
// lane_stack_ints(K) ints; out.cpl receives 32 entries.  Leaves are read in batches of kLaneBatch
// so that their loads overlap; no array is indexed by a per-thread value other than these.
constexpr int kLaneBatch = 8;
// ints of the introsort stack for tables of at most k symbols: pending ranges <= 2 floor(log2 k) + 2
CCA_HD constexpr int lane_stack_ints(uint32_t k) {
    int l = 0;
    while ((2u << l) <= k) l++;
    return 3 * (2 * l + 2);
}
template <class LeafF, class Stk>
CCA_HD void design_table_lane(LeafF leaf, int K, int dialect, uint32_t quirk, Strided<uint32_t> heap,
                              Strided<uint16_t> parent, Stk stack, TableOut out, uint64_t* bits) {
    constexpr uint32_t nodemask = (1u << kLaneNodeBits) - 1;
    for (int l = 0; l < 32; l++) out.cpl[l] = 0;
    *bits = 0;
    out.info[0] = 0;
    out.info[1] = (uint32_t)K;
    out.info[2] = 0;
    if (K == 0) return;
    // (1) leaves in symbol order
    for (int r0 = 0; r0 < K; r0 += kLaneBatch) {
        uint32_t c[kLaneBatch];
        for (int j = 0; j < kLaneBatch; j++) c[j] = r0 + j < K ? leaf(r0 + j).y : 0u;
        for (int j = 0; j < kLaneBatch; j++)
            if (r0 + j < K) heap[r0 + j] = (c[j] << kLaneNodeBits) | (uint32_t)(r0 + j);
    }
    // (2) merge loop: pop, pop, push of the merged node
    const PackedQueueLess qless;
    if (dialect == kDialectGpu) {
        make_heap(heap, K, qless);
    } else {
        for (int i = 1; i < K; i++) push_heap(heap, i + 1, qless);
    }
    int size = K, next = K;
    while (size > 1) {
        const uint32_t a = heap[0];
        pop_heap(heap, size, qless);
        size--;
        const uint32_t b = heap[0];
        pop_heap(heap, size, qless);
        size--;
        parent[(int)(a & nodemask)] = (uint16_t)next;
        parent[(int)(b & nodemask)] = (uint16_t)next;
        heap[size] = (((a >> kLaneNodeBits) + (b >> kLaneNodeBits)) << kLaneNodeBits) | (uint32_t)next;
        size++;
        push_heap(heap, size, qless);
        next++;
    }
    // (3) depths: every node's parent has a larger index, so one pass from the root down turns
    // the parent links into depths
    const int root = next - 1;
    parent[root] = 0;
    for (int x = root - 1; x >= 0; x--) parent[x] = (uint16_t)(parent[(int)parent[x]] + 1);
    // (4) leaves packed as depth << 24 | rank in symbol order, in the (now free) heap; the sort
    // compares depths only, so carrying the rank instead of the symbol moves nothing differently
    int deepest = 0;
    for (int r = 0; r < K; r++) {
        const int d = parent[r];
        deepest = d > deepest ? d : deepest;
        heap[r] = ((uint32_t)(d > kMaxDepth ? kMaxDepth : d) << 24) | (uint32_t)r;
    }
    if (deepest > kMaxDepth) {
        out.info[1] = 0;
        out.info[2] = 5;
        return;
    }
    if (deepest == 0) {  // a single leaf: no codeword bits
        const uint32_t s0 = leaf(0).x;
        out.syms[0] = s0;
        out.code[s0] = 0;
        out.len[s0] = 0;
        return;
    }
    // (5) std::sort by depth: the introsort loop, then its final insertion sort (a stable
    // insertion sort gives the same order: elements only move past strictly deeper ones)
    introsort_loop(heap, 0, K, 2 * floor_log2((uint32_t)K), DepthLess(), stack);
    for (int i = 1; i < K; i++) {
        const uint32_t v = heap[i];
        int j = i;
        for (; j > 0 && (heap[j - 1] >> 24) > (v >> 24); j--) heap[j] = heap[j - 1];
        heap[j] = v;
    }
    // (6) canonical codes in sorted order, c_0 = 0, c_i = (c_{i-1} + 1) << (l_i - l_{i-1})
    // (HuffmanTable.cu:553-580), and the count per length as runs of equal depth
    uint64_t b = 0;
    uint32_t code = 0, prev = 0, run = 0;
    for (int i0 = 0; i0 < K; i0 += kLaneBatch) {
        uint32_t e[kLaneBatch], sym[kLaneBatch], cnt[kLaneBatch];
        for (int j = 0; j < kLaneBatch; j++) e[j] = i0 + j < K ? heap[i0 + j] : 0u;
        for (int j = 0; j < kLaneBatch; j++) {
            if (i0 + j < K) {
                const auto lv = leaf((int)(e[j] & 0xFFFFFFu));
                sym[j] = lv.x;
                cnt[j] = lv.y;
            } else {
                sym[j] = cnt[j] = 0u;
            }
        }
        for (int j = 0; j < kLaneBatch; j++) {
            if (i0 + j >= K) break;
            const uint32_t d = e[j] >> 24, s = sym[j];
            if (i0 + j > 0) code = (code + 1) << (d - prev);
            if (d != prev && run) {
                out.cpl[prev - 1] = run;
                run = 0;
            }
            run++;
            prev = d;
            out.syms[i0 + j] = s;
            out.code[s] = code;
            out.len[s] = d;
            b += (uint64_t)(cnt[j] - ((quirk && s == 65535u) ? 1u : 0u)) * d;
        }
    }
    out.cpl[prev - 1] = run;
    out.info[0] = (uint32_t)deepest;
    *bits = b;
}

// ---------------------------------------------------------------- heap operations by path
// The same element moves as pop_heap / push_heap / make_heap, restated so that each operation
// becomes a few data-parallel steps (the wave executor's form):
//  * pop: the hole left at the root descends along the path on which every node continues to
//    its left child iff that child's count is strictly smaller (ties go right), down to the last
//    node with two children, plus the lone left child of the last internal node when the heap
//    size is even; counts along that path are non-decreasing, so the former last element v
//    lands at path position i* = m - #{k in 1..m : v.count < count(p_k)} and p_0..p_{i*-1} take
//    the values of p_1..p_{i*}.
//  * push: the ancestors of the new position have non-increasing counts going up, so the new
//    element rises past exactly the j ancestors with a larger count.
//  * make_heap: nodes of one level have disjoint subtrees, so each level's adjustments are
//    independent of their order.
CCA_HD int descend_path(const HeapEntry* h, int len, int start, int* path) {
    int m = 0, x = start;
    path[0] = start;
    while (x < (len - 1) / 2) {
        int r = 2 * (x + 1);
        if (h[r - 1].count < h[r].count) r--;
        path[++m] = r;
        x = r;
    }
    if ((len & 1) == 0 && x == (len - 2) / 2) path[++m] = 2 * x + 1;
    return m;
}

// The descent in the wave's form: rounds of six levels below the round's root `cur`.  Inside a
// round the walk goes from local node j (0 = cur; global index (cur << depth) + j) to local child
// 2j + 1 + r(j), with r(j) = 1 iff the node has two children and its left child's count is not
// smaller -- unconditionally, for all six levels.  The real path positions are those whose parent
// has two children, a prefix of the walk (indices grow along it, and a node has two children iff
// its index is below (n-1)/2).  Where the path ends at the last internal node of an even-sized
// heap, the walk's next step (r = 0 there) is that node's lone left child.  Same path as
// descend_path; path needs room for the last round's six positions.
CCA_HD int descend_walk(const HeapEntry* h, int n, int* path) {
    const int lim = (n - 1) / 2, lastint = (n - 2) / 2;
    int base = 0, cur = 0;
    path[0] = 0;
    for (;;) {
        int j = 0;
        for (int d = 0; d < 6; d++) {
            const int g = (cur << d) + j;
            const int r = (g < lim && !(h[2 * g + 1].count < h[2 * g + 2].count)) ? 1 : 0;
            j = 2 * j + 1 + r;
            path[base + d + 1] = (cur << (d + 1)) + j;
        }
        int s = 0;
        while (s < 6 && ((path[base + s + 1] - 1) >> 1) < lim) s++;
        if (s < 6) {
            int m = base + s;
            if ((n & 1) == 0 && path[m] == lastint) m++;
            return m;
        }
        base += 6;
        cur = path[base];
    }
}

// pop over h[0, len): afterwards h[0, len-1) is the heap pop_heap leaves (h[len-1] not written)
CCA_HD void pop_by_walk(HeapEntry* h, int len, int* path) {
    if (len <= 1) return;
    const HeapEntry v = h[len - 1];
    const int m = descend_walk(h, len - 1, path);
    int larger = 0;
    for (int k = 1; k <= m; k++) larger += (v.count < h[path[k]].count) ? 1 : 0;
    const int is = m - larger;
    for (int i = 0; i < is; i++) h[path[i]] = h[path[i + 1]];
    h[path[is]] = v;
}

// pop over h[0, len): afterwards h[0, len-1) is the heap pop_heap leaves (h[len-1] not written)
CCA_HD void pop_by_path(HeapEntry* h, int len, int* path) {
    if (len <= 1) return;
    const HeapEntry v = h[len - 1];
    const int m = descend_path(h, len - 1, 0, path);
    int larger = 0;
    for (int k = 1; k <= m; k++) larger += (v.count < h[path[k]].count) ? 1 : 0;
    const int is = m - larger;
    for (int i = 0; i < is; i++) h[path[i]] = h[path[i + 1]];
    h[path[is]] = v;
}

// push of v at position pos over the heap h[0, pos)
CCA_HD void push_by_chain(HeapEntry* h, int pos, HeapEntry v) {
    int T = 0;
    while (((pos + 1) >> (T + 1)) > 0) T++;
    int j = 0;
    for (int t = 1; t <= T; t++) j += (v.count < h[((pos + 1) >> t) - 1].count) ? 1 : 0;
    for (int t = 1; t <= j; t++) h[((pos + 1) >> (t - 1)) - 1] = h[((pos + 1) >> t) - 1];
    h[((pos + 1) >> j) - 1] = v;
}

CCA_HD void make_by_levels(HeapEntry* h, int len) {
    if (len < 2) return;
    const int last = (len - 2) / 2;
    int d = 0;
    while ((1 << (d + 1)) - 1 <= last) d++;
    for (; d >= 0; d--) {
        const int lo = (1 << d) - 1, hi = ((1 << (d + 1)) - 2) < last ? ((1 << (d + 1)) - 2) : last;
        for (int x = lo; x <= hi; x++) adjust_heap(h, x, len, h[x], QueueLess());
    }
}

// ---------------------------------------------------------------- serial executor (host)
struct SerialEx {
    CCA_HD int lane() const { return 0; }
    CCA_HD int lanes() const { return 1; }
    CCA_HD void barrier() const {}
    CCA_HD void heap_make(HeapEntry* h, int len) const { make_heap(h, len, QueueLess()); }
    // pop / push return the new top (h[0])
    CCA_HD HeapEntry heap_pop(HeapEntry* h, int len) const {
        pop_heap(h, len, QueueLess());
        return h[0];
    }
    CCA_HD HeapEntry heap_push(HeapEntry* h, int pos, HeapEntry v, HeapEntry) const {
        h[pos] = v;
        push_heap(h, pos + 1, QueueLess());
        return h[0];
    }
    template <class P, class S>
    CCA_HD int compact(int lo, int hi, P pred, S sink) const {
        int r = 0;
        for (int i = lo; i < hi; i++)
            if (pred(i)) sink(i, r++);
        return r;
    }
    template <class P>
    CCA_HD int count(int lo, int hi, P pred) const {
        int r = 0;
        for (int i = lo; i < hi; i++) r += pred(i) ? 1 : 0;
        return r;
    }
    CCA_HD int max_all(int v) const { return v; }
    CCA_HD uint64_t sum_all(uint64_t v) const { return v; }
    // counts per depth into cpl, then visits elements in order with their rank among equal depths
    template <class F>
    CCA_HD void stable_rank_by_depth(const uint32_t* el, int K, uint32_t* cpl, F f) const {
        for (int i = 0; i < K; i++) cpl[(el[i] >> 24) - 1]++;
        uint32_t fi[32], fc[32], seen[32];
        first_codes(cpl, 31, fc, fi);
        for (int l = 0; l < 32; l++) seen[l] = 0;
        for (int i = 0; i < K; i++) {
            const uint32_t d = el[i] >> 24;
            f(i, el[i], seen[d - 1]++, fi, fc);
        }
    }
};

// Serial executor using the path formulations of the heap operations (what the wave executor
// computes in parallel); the host check runs both executors against the standard library.
struct PathEx : SerialEx {
    mutable int path[96];
    CCA_HD void heap_make(HeapEntry* h, int len) const { make_by_levels(h, len); }
    CCA_HD HeapEntry heap_pop(HeapEntry* h, int len) const {
        pop_by_path(h, len, path);
        return h[0];
    }
    CCA_HD HeapEntry heap_push(HeapEntry* h, int pos, HeapEntry v, HeapEntry) const {
        push_by_chain(h, pos, v);
        return h[0];
    }
};

// Serial executor using the wave's six-level walk for the descent.
struct WalkEx : PathEx {
    CCA_HD HeapEntry heap_pop(HeapEntry* h, int len) const {
        pop_by_walk(h, len, path);
        return h[0];
    }
};

}  // namespace design
}  // namespace cca
