// Device-resident decode with no allocation: cc_decode_device's kernels (hip/decoder.hip) with
// every buffer they need beyond the caller's bitstream and output streams -- the call, item and
// wave tables, the sequence records, the job state and the decode arena -- carved from one
// file-scope device array.  For callers that cannot allocate device memory (tight serving loops,
// code that must not call hipMalloc).  The array is per device and shared by every call on that
// device: calls that overlap on one device race.
#include <mutex>

#include "../coder.hpp"
#include "../host/batch.hpp"
#include "../profile.hpp"

namespace cca {

constexpr uint64_t kNoAllocScratchBytes = 32ull << 20;  // per device
constexpr uint32_t kNoAllocMaxStreams = 128;             // output streams of one call

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

struct NoAllocArgs {
    const uint32_t* words;
    uint64_t nwords;
    uint32_t n, nstreams, rl, waves_per_seq, chunks_per_item;
    DecCall* calls;
    DecItem* items;
    uint32_t* wave_seq;
    uint32_t* wave_first;
    uint32_t* item_first;
    void* outs[kNoAllocMaxStreams];
};

// Writes the job tables of one call of nstreams equal items: item s decodes sequence s*q (+1
// for its zero counts, q = 2 for RLE items), each sequence's blocks in workgroups of
// kDecGroupBlocks (a short sequence: one wave), its zero-run
// expansion chunks from s * chunks_per_item.
__global__ void k_noalloc_setup(NoAllocArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t q = a.rl ? 2u : 1u;
    if (t == 0) a.calls[0] = DecCall{a.words, a.nwords, 0, a.nstreams};
    if (t < a.nstreams) a.items[t] = DecItem{a.outs[t], a.n, a.rl, t * q, 0};
    if (t <= a.nstreams) a.item_first[t] = t * a.chunks_per_item;
    const uint32_t nwaves = a.nstreams * q * a.waves_per_seq;
    if (t < nwaves) {
        a.wave_seq[t] = t / a.waves_per_seq;
        a.wave_first[t] = (t % a.waves_per_seq) * kDecGroupBlocks;  // one entry (first block 0) when short
    }
}

std::mutex g_mu;
unsigned char* g_ptr[64] = {};

// The array's device address on the current device, looked up once per device.
unsigned char* scratch_address(int device) {
    CCA_REQUIRE(device >= 0 && device < 64, CC_ERR_INVALID_ARG, "device index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            throw Error{CC_ERR_DEVICE, std::string("hipGetSymbolAddress of the decode scratch failed: ") +
                                           hipGetErrorString(e)};
        }
        g_ptr[device] = static_cast<unsigned char*>(p);
    }
    return g_ptr[device];
}

}  // namespace

uint64_t decode_device_noalloc(cc_instance* I, bool rl, int symSize, const uint32_t* d_words, uint64_t wordCount,
                               void* const* outs, uint32_t nstreams, uint32_t n, uint32_t* err) {
    check_symbols(I, symSize);
    CCA_REQUIRE(d_words != nullptr || wordCount == 0, CC_ERR_INVALID_ARG, "null bitstream");
    CCA_REQUIRE(nstreams <= kNoAllocMaxStreams, CC_ERR_INVALID_ARG,
                "more output streams than the no-allocation decode takes (128)");
    for (uint32_t s = 0; s < nstreams; s++)
        CCA_REQUIRE(outs[s] != nullptr || n == 0, CC_ERR_INVALID_ARG, "null output stream");
    DeviceGuard guard(I->device);
    unsigned char* base = scratch_address(I->device);
    const hipStream_t st = I->stream;
    const CoderParams p = params_of(I);
    const uint32_t q = rl ? 2u : 1u;
    const uint32_t nblk = div_up(n, p.cbs);
    const uint32_t waves_per_seq = dec_entries(nblk);
    const bool lng = dec_long(nblk);
    const uint32_t nseq = nstreams * q;
    const uint64_t nwaves = (uint64_t)nseq * waves_per_seq;
    // the arena bound of the allocating decode (block offsets, lookup and symbol tables, the
    // decoded pairs of RLE sequences)
    const uint64_t arena_words =
        64 + (uint64_t)nseq * ((uint64_t)nblk + 16 + (1u << kDecLutBits) + n + 17 +
                               (rl ? ((uint64_t)n * symSize + 3) / 4 + 17 : 0));
    const uint64_t o_items = 256;
    const uint64_t o_seqs = o_items + align_up(sizeof(DecItem) * kNoAllocMaxStreams, 256);
    const uint64_t o_state = o_seqs + align_up(sizeof(DecSeq) * 2 * kNoAllocMaxStreams, 256);
    const uint64_t o_wseq = o_state + 256;  // CoderState, then the call's DecCallOut
    const uint64_t o_wfirst = o_wseq + align_up(4 * std::max<uint64_t>(nwaves, 1), 256);
    const uint32_t cpi = rle_chunks(n, rl);
    const uint64_t nchunks = (uint64_t)cpi * nstreams;
    const uint64_t o_ifirst = o_wfirst + align_up(4 * std::max<uint64_t>(nwaves, 1), 256);
    const uint64_t o_sums = o_ifirst + align_up(4 * ((uint64_t)nstreams + 1), 256);
    const uint64_t o_arena = o_sums + align_up(4 * std::max<uint64_t>(nchunks, 1), 256);
    CCA_REQUIRE(o_arena + 4 * arena_words <= kNoAllocScratchBytes, CC_ERR_INVALID_ARG,
                "call too large for the no-allocation decode scratch (32 MiB per device)");
    static_assert(sizeof(CoderState) + sizeof(DecCallOut) <= 256, "state and call result share 256 bytes");

    NoAllocArgs a{};
    a.words = d_words;
    a.nwords = wordCount;
    a.n = n;
    a.nstreams = nstreams;
    a.rl = rl ? 1u : 0u;
    a.waves_per_seq = waves_per_seq;
    a.chunks_per_item = cpi;
    a.calls = reinterpret_cast<DecCall*>(base);
    a.items = reinterpret_cast<DecItem*>(base + o_items);
    a.wave_seq = reinterpret_cast<uint32_t*>(base + o_wseq);
    a.wave_first = reinterpret_cast<uint32_t*>(base + o_wfirst);
    a.item_first = reinterpret_cast<uint32_t*>(base + o_ifirst);
    for (uint32_t s = 0; s < nstreams; s++) a.outs[s] = outs[s];
    const uint64_t setup_threads = std::max<uint64_t>({nwaves, (uint64_t)nstreams + 1, 1});
    CCA_LAUNCH("noalloc_setup", 0, k_noalloc_setup, dim3(div_up(setup_threads, 256)), dim3(256), 0, st, a);
    CCA_LAUNCH_CHECK();

    CoderState* d_state = reinterpret_cast<CoderState*>(base + o_state);
    DecCallOut* d_out = reinterpret_cast<DecCallOut*>(base + o_state + sizeof(CoderState));
    coder_decode(a.calls, 1, a.items, nstreams, reinterpret_cast<DecSeq*>(base + o_seqs), nseq, p, symSize,
                 reinterpret_cast<uint32_t*>(base + o_arena), arena_words, d_state, d_out, a.wave_seq, a.wave_first,
                 (uint32_t)nwaves, lng ? (uint32_t)nwaves : 0u, (uint64_t)nstreams * n,
                 RleChunks{a.item_first, (uint32_t)nchunks, reinterpret_cast<uint32_t*>(base + o_sums)}, st);
    struct {
        CoderState s;
        DecCallOut o;
    } r{};
    CCA_HIP(hipMemcpyAsync(&r, d_state, sizeof(r), hipMemcpyDeviceToHost, st));
    CCA_HIP(hipStreamSynchronize(st));
    *err = r.o.err ? r.o.err : r.s.err;
    return r.o.words;
}

}  // namespace cca
