// Batched float bricks: compressVolumeFloat / decompressVolumeFloat (src/examples/
// CompressVolume.cpp:313-516) for many bricks of different sizes at once.  Each brick's
// bitstream is word for word what the single-volume call produces for it: the same streams in
// the same order (lowpass with encodeHuff, then the 7 highpass bands of every level, coarsest
// first, with encodeRLHuff), the same per-level quantization steps (q0 / 2^level, lowpass
// q0 / 2^(levels-1)), coded by the device-resident coder with one call per brick.
//
// A job is prepared once (layout, level tables, coder items: one pinned upload each) and then
// launched as kernels only, so a plan can capture the whole compression and the whole
// decompression into two hipGraphs and replay them.
#include "bricks.hpp"

#include <algorithm>
#include <cstring>

#include "../bricks.hpp"

namespace cca {

namespace {

uint64_t elems(const cc_brick& b) { return (uint64_t)b.sizeX * b.sizeY * b.sizeZ; }

void check_bricks(const cc_brick* bricks, uint32_t n, uint32_t L) {
    CCA_REQUIRE(L >= 1 && L <= 10, CC_ERR_INVALID_ARG, "numLevels must be in [1, 10]");
    CCA_REQUIRE(n == 0 || bricks, CC_ERR_INVALID_ARG, "null brick array");
    const uint32_t f = 1u << L;
    for (uint32_t i = 0; i < n; i++) {
        const cc_brick& b = bricks[i];
        CCA_REQUIRE(b.sizeX % f == 0 && b.sizeY % f == 0 && b.sizeZ % f == 0 && b.sizeX && b.sizeY && b.sizeZ > 1,
                    CC_ERR_INVALID_ARG, "brick sizes must be divisible by 2^numLevels (3D)");
        CCA_REQUIRE(elems(b) < (1ull << 31), CC_ERR_INVALID_ARG, "brick too large");
    }
}

// symbol offset (elements, from the brick's symbol base) of stream k in bitstream order:
// k = 0 lowpass, then 7 bands per level from the coarsest level down
uint64_t stream_off(uint64_t N, uint32_t L, uint32_t k) {
    if (k == 0) return 0;
    uint64_t off = N >> (3 * L);
    uint32_t kk = 1;
    for (int l = (int)L - 1; l >= 0; l--) {
        const uint64_t nb = N >> (3 * (l + 1));
        for (uint32_t band = 0; band < 7; band++, kk++) {
            if (kk == k) return off;
            off += nb;
        }
    }
    return off;
}

void make_layout(BrickJob& job) {
    const uint32_t n = (uint32_t)job.bricks.size(), L = job.L;
    job.a.resize(n);
    job.b.resize(n);
    job.sym.resize(n);
    job.lp.assign(L, std::vector<uint64_t>(n, 0));
    job.level_elems.assign(L, 0);
    uint64_t w = 0, s = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t N = elems(job.bricks[i]);
        job.a[i] = w;
        w += align_up(N, 64);
        job.b[i] = w;
        w += align_up(N, 64);
        for (uint32_t l = 0; l + 1 < L; l++) {
            job.lp[l][i] = w;
            w += align_up(N >> (3 * (l + 1)), 64);
        }
        for (uint32_t l = 0; l < L; l++) job.level_elems[l] += N >> (3 * l);
        job.sym[i] = s;
        s += align_up(N, 64);
    }
    job.work_floats = w;
    job.sym_elems = s;
}

// coder items of all bricks in bitstream order (one call per brick)
std::vector<ItemSpec> brick_items(const BrickJob& job, unsigned char* sym) {
    std::vector<ItemSpec> items;
    const uint32_t L = job.L;
    for (uint32_t i = 0; i < job.bricks.size(); i++) {
        const uint64_t N = elems(job.bricks[i]);
        auto ptr = [&](uint32_t k) {
            return sym ? (void*)(sym + (job.sym[i] + stream_off(N, L, k)) * job.symSize) : nullptr;
        };
        items.push_back(ItemSpec{ptr(0), (uint32_t)(N >> (3 * L)), false, i});
        uint32_t k = 1;
        for (int l = (int)L - 1; l >= 0; l--) {
            const uint32_t nb = (uint32_t)(N >> (3 * (l + 1)));
            for (uint32_t band = 0; band < 7; band++) items.push_back(ItemSpec{ptr(k++), nb, true, i});
        }
    }
    return items;
}

void upload_levels(BrickCtx& ctx, BrickJob& job, float* work, unsigned char* sym, const cc_tuning& tune,
                   hipStream_t st) {
    const uint32_t n = (uint32_t)job.bricks.size(), L = job.L;
    const size_t lv_bytes = sizeof(BrickLevel) * (size_t)L * n;
    const size_t tb_off = align_up(std::max<size_t>(lv_bytes, 1), 256);
    const size_t bytes = tb_off + 4ull * L * 4 * (n + 1);
    unsigned char* h = (unsigned char*)ctx.h_blevels.get(bytes);
    BrickLevel* hl = (BrickLevel*)h;
    uint32_t* ht = (uint32_t*)(h + tb_off);
    job.tiles_total.assign(4 * L, 0);
    for (uint32_t l = 0; l < L; l++) {
        // the fused x+y passes serve a level when every brick's rows fit them (and the tuning
        // keeps them on)
        bool fused = tune.fusedXY != 0;
        for (uint32_t i = 0; i < n; i++) fused = fused && brick_xy_fits((int)(job.bricks[i].sizeX >> l));
        uint32_t acc[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < n; i++) {
            const cc_brick& b = job.bricks[i];
            const uint64_t N = elems(b);
            BrickLevel& e = hl[(size_t)l * n + i];
            e.lx = (int)(b.sizeX >> l);
            e.ly = (int)(b.sizeY >> l);
            e.lz = (int)(b.sizeZ >> l);
            e.last = l == L - 1;
            e.in = l == 0 ? b.src : work + job.lp[l - 1][i];
            e.a = work + job.a[i];
            e.b = work + job.b[i];
            e.lp = (l + 1 < L) ? work + job.lp[l][i] : nullptr;
            // highpass bands of level l start at stream 1 + 7 (L-1-l) in bitstream order
            e.hp = sym + (job.sym[i] + stream_off(N, L, 1 + 7 * (L - 1 - l))) * job.symSize;
            e.lpsym = sym + job.sym[i] * job.symSize;
            e.out = l == 0 ? b.dst : work + job.lp[l - 1][i];
            const float q0 = b.quantStep;
            e.step_hp = q0 / float(1u << l);  // CompressVolume.cpp:366 (q0 / (f/2), f = 2^(l+1))
            e.inv_hp = 1.0f / e.step_hp;      // Quantize.cu:94
            e.step_lp = q0 / float(1u << (L - 1));
            e.inv_lp = 1.0f / e.step_lp;
            uint32_t zk = fused ? std::min<uint32_t>((uint32_t)e.lz, brick_xy_slices(e.lx)) : 0u;
            if (fused && tune.xySlicesMax) zk = std::min(zk, tune.xySlicesMax);
            e.zk = (int)zk;
            for (int ax = 0; ax < 4; ax++) {
                ht[((size_t)l * 4 + ax) * (n + 1) + i] = acc[ax];
                acc[ax] += ax < 3 ? brick_tiles(ax, e.lx, e.ly, e.lz) : (fused ? brick_xy_tiles(e.ly, e.lz, e.zk) : 0u);
            }
        }
        for (int ax = 0; ax < 4; ax++) {
            ht[((size_t)l * 4 + ax) * (n + 1) + n] = acc[ax];
            job.tiles_total[4 * l + ax] = acc[ax];
        }
    }
    unsigned char* d = (unsigned char*)ctx.d_blevels.get(bytes);
    CCA_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    job.d_levels = (const BrickLevel*)d;
    job.d_tiles = (const uint32_t*)(d + tb_off);
}

cc_status status_of_encode(uint32_t e) {
    return e == 0 ? CC_OK
           : e == kErrOverflow ? CC_ERR_SYMBOL_OVERFLOW
           : e == kErrTooLong ? CC_ERR_CODEWORD_TOO_LONG
           : e == kErrOutput ? CC_ERR_INVALID_ARG : CC_ERR_OUT_OF_MEMORY;
}

}  // namespace

void bricks_prepare(cc_instance* I, BrickCtx& ctx, BrickJob& job, bool enc, bool dec,
                    const std::vector<CallWords>* calls, hipStream_t st) {
    const uint32_t n = (uint32_t)job.bricks.size();
    check_bricks(job.bricks.data(), n, job.L);
    for (uint32_t i = 0; i < n; i++) {
        CCA_REQUIRE(!enc || job.bricks[i].src, CC_ERR_INVALID_ARG, "null brick source");
        CCA_REQUIRE(!dec || job.bricks[i].dst, CC_ERR_INVALID_ARG, "null brick destination");
    }
    make_layout(job);
    float* work = ctx.d_bwork.as<float>(std::max<uint64_t>(job.work_floats, 64));
    unsigned char* sym = (unsigned char*)ctx.d_bsym.get(std::max<uint64_t>(job.sym_elems, 64) * job.symSize);
    upload_levels(ctx, job, work, sym, I->tuning, st);
    const std::vector<ItemSpec> items = brick_items(job, sym);
    if (enc) encode_prepare(I, ctx.coder, items, n, job.symSize, job.ej, st);
    if (dec) {
        if (calls) {
            CCA_REQUIRE(calls->size() == n, CC_ERR_INVALID_ARG, "one bitstream per brick");
            decode_prepare(I, ctx.coder, items, *calls, job.symSize, job.dj, st);
        } else {  // positions filled on the device from the encoder's results
            decode_prepare(I, ctx.coder, items, std::vector<CallWords>(n, CallWords{nullptr, 0}), job.symSize, job.dj,
                           st);
        }
    }
}

void bricks_prepare_bound(cc_instance* I, BrickCtx& ctx, BrickJob& job, const DecodeIndex& X,
                          const std::vector<uint32_t>& index_calls, hipStream_t st) {
    const uint32_t n = (uint32_t)job.bricks.size();
    CCA_REQUIRE(index_calls.size() == n, CC_ERR_INVALID_ARG, "one index call per brick");
    check_bricks(job.bricks.data(), n, job.L);
    for (uint32_t i = 0; i < n; i++)
        CCA_REQUIRE(job.bricks[i].dst, CC_ERR_INVALID_ARG, "null brick destination");
    make_layout(job);
    float* work = ctx.d_bwork.as<float>(std::max<uint64_t>(job.work_floats, 64));
    unsigned char* sym = (unsigned char*)ctx.d_bsym.get(std::max<uint64_t>(job.sym_elems, 64) * job.symSize);
    upload_levels(ctx, job, work, sym, I->tuning, st);
    decode_prepare_bound(I, ctx.coder, X, brick_items(job, sym), index_calls, job.symSize, job.bj, st);
    job.index = &X;
}

void bricks_index_build(cc_instance* I, DecodeIndex& X, const std::vector<cc_brick>& bricks, uint32_t L,
                        int symSize, const std::vector<CallWords>& calls, hipStream_t st) {
    BrickJob job;
    job.bricks = bricks;
    job.L = L;
    job.symSize = symSize;
    check_bricks(job.bricks.data(), (uint32_t)bricks.size(), L);
    decode_index_build(I, X, brick_items(job, nullptr), calls, symSize, st);
}

void bricks_forward_launch(const BrickJob& job, hipStream_t st) {
    const uint32_t n = (uint32_t)job.bricks.size();
    for (uint32_t l = 0; l < job.L; l++)
        bricks_forward_level(job.d_levels + (size_t)l * n, job.d_tiles + (size_t)l * 4 * (n + 1), n,
                             &job.tiles_total[4 * l], job.symSize, job.level_elems[l], st);
}

void bricks_encode_launch(BrickCtx& ctx, const BrickJob& job, hipStream_t st, const LayoutLink* link) {
    bricks_forward_launch(job, st);
    encode_launch(ctx.coder, job.ej, job.d_out, job.out_cap, st, link);
}

void bricks_decode_launch(BrickCtx& ctx, const BrickJob& job, hipStream_t st, bool from_encode) {
    const uint32_t n = (uint32_t)job.bricks.size();
    if (job.index) {
        decode_launch_bound(ctx.coder, *job.index, job.bj, st);
    } else {
        if (from_encode)
            decode_calls_from_encode(ctx.coder, job.dj, (const CallOut*)ctx.coder.d_calls.p, job.d_out, job.out_cap,
                                     st);
        decode_launch(ctx.coder, job.dj, st);
    }
    for (int l = (int)job.L - 1; l >= 0; l--)
        bricks_inverse_level(job.d_levels + (size_t)l * n, job.d_tiles + (size_t)l * 4 * (n + 1), n,
                             &job.tiles_total[4 * l], job.symSize, job.level_elems[l], st);
}

bool bricks_encode_results(const BrickJob& job, const std::vector<CallOut>& res, uint64_t total, uint64_t* offsets,
                           uint32_t* status) {
    const uint32_t n = (uint32_t)job.bricks.size();
    bool ok = true;
    for (uint32_t i = 0; i < n; i++) {
        offsets[i] = res[i].word_begin;
        const cc_status s = status_of_encode(res[i].err);
        if (status) status[i] = (uint32_t)s;
        ok = ok && s == CC_OK;
    }
    offsets[n] = total;
    return ok;
}

bool bricks_decode_results(const BrickJob& job, const std::vector<DecCallOut>& res, uint32_t* status) {
    bool ok = true;
    for (uint32_t i = 0; i < job.bricks.size(); i++) {
        const cc_status s = res[i].err == 0 ? CC_OK
                            : (res[i].err == kDecCorrupt ? CC_ERR_CORRUPT_STREAM : CC_ERR_OUT_OF_MEMORY);
        if (status) status[i] = (uint32_t)s;
        ok = ok && s == CC_OK;
    }
    return ok;
}

// ------------------------------------------------------------------------------------- groups
// A batch runs as groups: contiguous brick ranges of about equal voxel count, each a pipeline of
// its own (context, stream).  The groups' kernels overlap on the device -- in particular the
// latency-bound table design of one group with the bandwidth-bound passes of the others -- and
// their layouts are chained so that the bitstreams land back to back in brick order, exactly
// where a single batch puts them.

// Groups are opt-in (cc_set_instance_brick_groups): each group adds a stream plus its coder's
// design streams, and with the runtime's four hardware queues per process those streams share
// queues, so one group's long table designs hold back the next group's passes queued behind them.
uint32_t groups_for(const cc_instance* I, const cc_brick* bricks, uint32_t n) {
    (void)bricks;
    return I->brick_groups ? std::min<uint32_t>(I->brick_groups, std::max<uint32_t>(n, 1)) : 1u;
}

// group boundaries: group g holds bricks [first[g], first[g+1]), never empty
std::vector<uint32_t> split_groups(const cc_brick* bricks, uint32_t n, uint32_t G) {
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) total += elems(bricks[i]);
    std::vector<uint32_t> first{0};
    uint64_t acc = 0;
    for (uint32_t i = 0; i + 1 < n; i++) {
        acc += elems(bricks[i]);
        if (first.size() < G && acc * G >= total * first.size()) first.push_back(i + 1);
    }
    first.push_back(n);
    return first;
}

void Groups::setup(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize, uint32_t* d_out,
                   uint64_t cap, uint32_t G, std::vector<std::unique_ptr<BrickCtx>>& pool) {
    first = split_groups(bricks, n, std::max<uint32_t>(G, 1));
    const uint32_t ng = (uint32_t)first.size() - 1;
    while (pool.size() < ng) pool.emplace_back(new BrickCtx());
    ctx.assign(ng, nullptr);
    jobs.assign(ng, BrickJob());
    for (uint32_t g = 0; g < ng; g++) {
        ctx[g] = pool[g].get();
        ctx[g]->ensure_stream();
        BrickJob& job = jobs[g];
        job.bricks.assign(bricks + first[g], bricks + first[g + 1]);
        job.L = L;
        job.symSize = symSize;
        job.d_out = d_out;
        job.out_cap = cap;
    }
    (void)I;
}

void Groups::fork(hipStream_t from, hipEvent_t ev) {
    CCA_HIP(hipEventRecord(ev, from));
    for (BrickCtx* c : ctx) CCA_HIP(hipStreamWaitEvent(c->st, ev, 0));
}

void Groups::join(hipStream_t into) {
    for (BrickCtx* c : ctx) {
        CCA_HIP(hipEventRecord(c->ev_done, c->st));
        CCA_HIP(hipStreamWaitEvent(into, c->ev_done, 0));
    }
}

void Groups::encode_launch(hipStream_t from, hipEvent_t ev) {
    fork(from, ev);
    for (size_t g = 0; g < ctx.size(); g++) {
        const LayoutLink link{g ? (const CoderState*)ctx[g - 1]->coder.d_state.p : nullptr,
                              g ? ctx[g - 1]->ev_layout : nullptr, ctx[g]->ev_layout};
        bricks_encode_launch(*ctx[g], jobs[g], ctx[g]->st, &link);
    }
    join(from);
}

void Groups::decode_launch(hipStream_t from, hipEvent_t ev, bool from_encode) {
    fork(from, ev);
    for (size_t g = 0; g < ctx.size(); g++) bricks_decode_launch(*ctx[g], jobs[g], ctx[g]->st, from_encode);
    join(from);
}

// Encodes (retrying with larger arenas when a group's runs out) and reads the per-brick results.
// Returns the total word count.
uint64_t Groups::encode_run(hipStream_t from, hipEvent_t ev, uint64_t* offsets, uint32_t* status, bool* ok,
                            uint64_t headroom_div) {
    const size_t ng = ctx.size();
    std::vector<std::vector<CallOut>> res(ng);
    std::vector<CoderState> s(ng);
    for (int attempt = 0;; attempt++) {
        encode_launch(from, ev);
        bool retry = false;
        for (size_t g = 0; g < ng; g++) {
            s[g] = encode_finish(ctx[g]->coder, jobs[g].ej, &res[g], ctx[g]->st);
            if (s[g].err == kErrScratch && attempt < 4) {
                encode_reserve(ctx[g]->coder, jobs[g].ej,
                               std::max<uint64_t>(2 * jobs[g].ej.scratch_cap, s[g].scratch_used + s[g].scratch_used / 2));
                retry = true;
            }
        }
        if (!retry) break;
    }
    if (headroom_div)
        for (size_t g = 0; g < ng; g++)
            encode_reserve(ctx[g]->coder, jobs[g].ej,
                           std::max<uint64_t>(jobs[g].ej.scratch_cap, s[g].scratch_used + s[g].scratch_used / headroom_div));
    bool all = true;
    for (size_t g = 0; g < ng; g++) {
        CCA_REQUIRE(s[g].err != kErrScratch, CC_ERR_OUT_OF_MEMORY, "coder arena exhausted");
        const uint64_t end = s[g].out_base + s[g].out_words;
        all = bricks_encode_results(jobs[g], res[g], end, offsets + first[g], status ? status + first[g] : nullptr) && all;
    }
    *ok = all;
    return s[ng - 1].out_base + s[ng - 1].out_words;
}

bool Groups::decode_results(uint32_t* status) {
    bool all = true;
    for (size_t g = 0; g < ctx.size(); g++) {
        std::vector<DecCallOut> res;
        decode_finish(ctx[g]->coder, jobs[g].dj, &res, ctx[g]->st);
        all = bricks_decode_results(jobs[g], res, status ? status + first[g] : nullptr) && all;
    }
    return all;
}

void compress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize, uint32_t* d_out,
                     uint64_t cap, uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(I->device);
    check_bricks(bricks, n, L);
    if (n == 0) {
        offsets[0] = 0;
        return;
    }
    if (!I->group_fork) CCA_HIP(hipEventCreateWithFlags(&I->group_fork, hipEventDisableTiming));
    Groups gr;
    gr.setup(I, bricks, n, L, symSize, d_out, cap, groups_for(I, bricks, n), I->groups);
    for (size_t g = 0; g < gr.ctx.size(); g++)
        bricks_prepare(I, *gr.ctx[g], gr.jobs[g], true, false, nullptr, gr.ctx[g]->st);
    bool ok = false;
    const uint64_t total = gr.encode_run(I->stream, I->group_fork, offsets, status, &ok, 0);
    CCA_REQUIRE(total <= cap, CC_ERR_INVALID_ARG, "output buffer too small for the compressed bricks");
    CCA_REQUIRE(ok, CC_ERR_SYMBOL_OVERFLOW, "some bricks failed to encode (see the per-brick status)");
}

void decompress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize,
                       const uint32_t* d_in, const uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(I->device);
    check_bricks(bricks, n, L);
    if (n == 0) return;
    for (uint32_t i = 0; i < n; i++)
        CCA_REQUIRE(offsets[i + 1] >= offsets[i], CC_ERR_INVALID_ARG, "brick offsets must be nondecreasing");
    if (!I->group_fork) CCA_HIP(hipEventCreateWithFlags(&I->group_fork, hipEventDisableTiming));
    Groups gr;
    gr.setup(I, bricks, n, L, symSize, nullptr, 0, groups_for(I, bricks, n), I->groups);
    for (size_t g = 0; g < gr.ctx.size(); g++) {
        std::vector<CallWords> calls;
        for (uint32_t i = gr.first[g]; i < gr.first[g + 1]; i++)
            calls.push_back(CallWords{d_in + offsets[i], offsets[i + 1] - offsets[i]});
        bricks_prepare(I, *gr.ctx[g], gr.jobs[g], false, true, &calls, gr.ctx[g]->st);
    }
    gr.decode_launch(I->stream, I->group_fork, false);
    CCA_REQUIRE(gr.decode_results(status), CC_ERR_CORRUPT_STREAM,
                "some bricks failed to decode (see the per-brick status)");
}

}  // namespace cca

// ------------------------------------------------------------------------------------- plans
// A plan captures its batch once and replays it.  With one group (the default) the compression
// and the decompression are two hipGraphs on the group's stream.  With several groups
// (cc_set_instance_brick_groups), each group -- a contiguous range of bricks with its own context
// and stream -- is captured on its own as three graphs (forward passes, coder, decode); a replay
// starts group g's forward passes once group g-1's are done, so the groups run staggered: one
// group's latency-bound table designs overlap the next group's bandwidth-bound passes and the
// previous group's decode.  Each group lays its bitstream out from word 0 of its own staging
// region and decodes it from there; once every group has been coded, one kernel concatenates the
// regions into the caller's output, where the words are exactly those of the one-group plan.
struct cc_bricks_plan {
    cc_instance* inst = nullptr;
    std::vector<std::unique_ptr<cca::BrickCtx>> pool;  // one context (and stream) per group
    cca::Groups gr;
    // one group: the group context's stream (capture origin and replay stream); several groups:
    // the plan's own stream (ordering, concatenation, results)
    hipStream_t st = nullptr;
    bool own_stream = false;
    hipGraph_t g_enc = nullptr, g_dec = nullptr;
    hipGraphExec_t x_enc = nullptr, x_dec = nullptr;
    // several groups: per group its forward, coder and decode graphs and the events between them
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> gx_fwd, gx_code, gx_dec;
    std::vector<hipEvent_t> ev_fwd, ev_code, ev_dec;
    cca::DevBuf d_stage;   // the groups' bitstreams, group g from word stage_off[g]
    cca::DevBuf d_concat;  // ConcatGroup records of the concatenation kernel
    std::vector<uint64_t> stage_off, stage_cap;
    uint32_t* d_out = nullptr;
    uint64_t out_cap = 0;
    hipEvent_t order = nullptr;  // orders replays after the instance stream's earlier work
    bool encoded = false;
    bool launched = false;     // a plan_launch not yet waited for
    ~cc_bricks_plan();
};

cc_bricks_plan::~cc_bricks_plan() {
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (inst) (void)hipSetDevice(inst->device);
    if (st) (void)hipStreamSynchronize(st);
    for (auto* c : gr.ctx)
        if (c && c->st) (void)hipStreamSynchronize(c->st);
    if (x_enc) (void)hipGraphExecDestroy(x_enc);
    if (x_dec) (void)hipGraphExecDestroy(x_dec);
    if (g_enc) (void)hipGraphDestroy(g_enc);
    if (g_dec) (void)hipGraphDestroy(g_dec);
    for (auto* v : {&gx_fwd, &gx_code, &gx_dec})
        for (hipGraphExec_t x : *v)
            if (x) (void)hipGraphExecDestroy(x);
    for (hipGraph_t g : graphs)
        if (g) (void)hipGraphDestroy(g);
    for (auto* v : {&ev_fwd, &ev_code, &ev_dec})
        for (hipEvent_t e : *v)
            if (e) (void)hipEventDestroy(e);
    if (order) (void)hipEventDestroy(order);
    d_stage.release();
    d_concat.release();
    if (own_stream && st) (void)hipStreamDestroy(st);
    for (auto& c : pool) c->release();  // also destroys the group streams
    if (prev >= 0) (void)hipSetDevice(prev);
}

namespace cca {

namespace {
void capture(hipStream_t st, hipGraph_t* g, hipGraphExec_t* x, const std::function<void()>& body) {
    CCA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(st, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    CCA_HIP(hipStreamEndCapture(st, g));
    CCA_HIP(hipGraphInstantiate(x, *g, nullptr, nullptr, 0));
}

bool grouped(const cc_bricks_plan* P) { return P->gr.ctx.size() > 1; }

// Several groups: an eager grouped compression (chained layouts into the caller's output) sizes
// every group's arena and bitstream; then each group gets a staging region with headroom and is
// captured as forward, coder and decode graphs.
void plan_create_groups(cc_bricks_plan* P, cc_instance* I) {
    Groups& gr = P->gr;
    const size_t ng = gr.ctx.size();
    CCA_HIP(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
    P->own_stream = true;
    CCA_HIP(hipEventRecord(P->order, I->stream));
    CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
    for (size_t g = 0; g < ng; g++) {
        CCA_HIP(hipStreamWaitEvent(gr.ctx[g]->st, P->order, 0));
        bricks_prepare(I, *gr.ctx[g], gr.jobs[g], true, true, nullptr, gr.ctx[g]->st);
    }
    const uint32_t n = gr.first.back();
    std::vector<uint64_t> offs(n + 1);
    bool ok = false;
    hipEvent_t fork = nullptr;
    CCA_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    P->ev_fwd.push_back(fork);  // destroyed with the plan
    gr.encode_run(P->st, fork, offs.data(), nullptr, &ok, 4);
    P->stage_off.assign(ng, 0);
    P->stage_cap.assign(ng, 0);
    uint64_t total = 0;
    for (size_t g = 0; g < ng; g++) {
        const uint64_t w = offs[gr.first[g + 1]] - offs[gr.first[g]];
        P->stage_off[g] = total;
        P->stage_cap[g] = w + w / 4 + 4096;
        total += align_up(P->stage_cap[g], 64);
    }
    uint32_t* stage = P->d_stage.as<uint32_t>(total);
    std::vector<ConcatGroup> cg(ng);
    for (size_t g = 0; g < ng; g++) {
        gr.jobs[g].d_out = stage + P->stage_off[g];
        gr.jobs[g].out_cap = P->stage_cap[g];
        cg[g] = ConcatGroup{gr.jobs[g].d_out, (const CoderState*)gr.ctx[g]->coder.d_state.p};
    }
    CCA_HIP(hipMemcpy(P->d_concat.get(sizeof(ConcatGroup) * ng), cg.data(), sizeof(ConcatGroup) * ng,
                      hipMemcpyHostToDevice));
    P->ev_fwd.pop_back();
    CCA_HIP(hipEventDestroy(fork));
    P->gx_fwd.assign(ng, nullptr);
    P->gx_code.assign(ng, nullptr);
    P->gx_dec.assign(ng, nullptr);
    P->ev_fwd.assign(ng, nullptr);
    P->ev_code.assign(ng, nullptr);
    P->ev_dec.assign(ng, nullptr);
    for (size_t g = 0; g < ng; g++) {
        for (auto* v : {&P->ev_fwd, &P->ev_code, &P->ev_dec})
            CCA_HIP(hipEventCreateWithFlags(&(*v)[g], hipEventDisableTiming));
        BrickCtx& ctx = *gr.ctx[g];
        const BrickJob& job = gr.jobs[g];
        hipGraph_t a = nullptr, b = nullptr, c = nullptr;
        capture(ctx.st, &a, &P->gx_fwd[g], [&] { bricks_forward_launch(job, ctx.st); });
        P->graphs.push_back(a);
        capture(ctx.st, &b, &P->gx_code[g],
                [&] { encode_launch(ctx.coder, job.ej, job.d_out, job.out_cap, ctx.st, nullptr); });
        P->graphs.push_back(b);
        capture(ctx.st, &c, &P->gx_dec[g], [&] { bricks_decode_launch(ctx, job, ctx.st, true); });
        P->graphs.push_back(c);
    }
}

// Several groups: replays the groups' graphs, staggered by their forward passes; the plan stream
// waits for every group's coder, concatenates, and (with decode) waits for every group's decode.
void launch_groups(cc_bricks_plan* P, bool enc, bool dec) {
    Groups& gr = P->gr;
    const size_t ng = gr.ctx.size();
    CCA_HIP(hipEventRecord(P->order, P->inst->stream));
    CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
    for (size_t g = 0; g < ng; g++) {
        const hipStream_t s = gr.ctx[g]->st;
        CCA_HIP(hipStreamWaitEvent(s, P->order, 0));
        if (enc) {
            if (g > 0) CCA_HIP(hipStreamWaitEvent(s, P->ev_fwd[g - 1], 0));
            CCA_HIP(hipGraphLaunch(P->gx_fwd[g], s));
            CCA_HIP(hipEventRecord(P->ev_fwd[g], s));
            CCA_HIP(hipGraphLaunch(P->gx_code[g], s));
            CCA_HIP(hipEventRecord(P->ev_code[g], s));
        }
        if (dec) {
            CCA_HIP(hipGraphLaunch(P->gx_dec[g], s));
            CCA_HIP(hipEventRecord(P->ev_dec[g], s));
        }
    }
    if (enc) {
        for (size_t g = 0; g < ng; g++) CCA_HIP(hipStreamWaitEvent(P->st, P->ev_code[g], 0));
        coder_concat((const ConcatGroup*)P->d_concat.p, (uint32_t)ng, P->d_out, P->out_cap, 0, P->st);
    }
    if (dec)
        for (size_t g = 0; g < ng; g++) CCA_HIP(hipStreamWaitEvent(P->st, P->ev_dec[g], 0));
}
}  // namespace

cc_bricks_plan* plan_create(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize,
                            uint32_t* d_out, uint64_t cap) {
    DeviceGuard guard(I->device);
    check_bricks(bricks, n, L);
    CCA_REQUIRE(n > 0, CC_ERR_INVALID_ARG, "a plan needs at least one brick");
    std::unique_ptr<cc_bricks_plan> P(new cc_bricks_plan());
    P->inst = I;
    P->d_out = d_out;
    P->out_cap = cap;
    CCA_HIP(hipEventCreateWithFlags(&P->order, hipEventDisableTiming));
    Groups& gr = P->gr;
    gr.setup(I, bricks, n, L, symSize, d_out, cap, groups_for(I, bricks, n), P->pool);
    if (grouped(P.get())) {
        plan_create_groups(P.get(), I);
        return P.release();
    }
    // One group: captured on the group context's stream (the capture origin), with the
    // table-design classes forked to that context's side streams.
    BrickCtx& ctx = *gr.ctx[0];
    BrickJob& job = gr.jobs[0];
    P->st = ctx.st;
    CCA_HIP(hipEventRecord(P->order, I->stream));
    CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
    bricks_prepare(I, ctx, job, true, true, nullptr, P->st);
    // one eager pass sizes the arena (with headroom for other data of the same geometry) and
    // allocates every buffer before capture
    std::vector<CallOut> res;
    for (int attempt = 0;; attempt++) {
        bricks_encode_launch(ctx, job, P->st);
        const CoderState s = encode_finish(ctx.coder, job.ej, &res, P->st);
        if (s.err != kErrScratch || attempt >= 4) {
            encode_reserve(ctx.coder, job.ej, std::max<uint64_t>(job.ej.scratch_cap, s.scratch_used + s.scratch_used / 4));
            break;
        }
        encode_reserve(ctx.coder, job.ej, std::max<uint64_t>(2 * job.ej.scratch_cap, s.scratch_used + s.scratch_used / 2));
    }
    capture(P->st, &P->g_enc, &P->x_enc, [&] { bricks_encode_launch(ctx, job, P->st); });
    capture(P->st, &P->g_dec, &P->x_dec, [&] { bricks_decode_launch(ctx, job, P->st, true); });

    return P.release();
}

namespace {

// Reads the results of the plan's last compression (synchronises its streams).
void plan_encode_results(cc_bricks_plan* P, uint64_t* offsets, uint32_t* status) {
    Groups& gr = P->gr;
    P->encoded = true;
    const size_t ng = gr.ctx.size();
    uint64_t prefix = 0;
    bool ok = true;
    for (size_t g = 0; g < ng; g++) {
        std::vector<CallOut> res;
        const hipStream_t s = grouped(P) ? gr.ctx[g]->st : P->st;
        const CoderState cs = encode_finish(gr.ctx[g]->coder, gr.jobs[g].ej, &res, s);
        CCA_REQUIRE(cs.err != kErrScratch, CC_ERR_OUT_OF_MEMORY, "plan arena too small for this data; create a new plan");
        const uint64_t total = cs.out_base + cs.out_words;
        CCA_REQUIRE(total <= gr.jobs[g].out_cap, CC_ERR_INVALID_ARG,
                    grouped(P) ? "plan staging too small for this data; create a new plan"
                               : "output buffer too small for the compressed bricks");
        const uint32_t f = gr.first[g], nb = gr.first[g + 1] - f;
        std::vector<uint64_t> o(nb + 1);
        ok = bricks_encode_results(gr.jobs[g], res, total, o.data(), status ? status + f : nullptr) && ok;
        for (uint32_t k = 0; k <= nb; k++) offsets[f + k] = prefix + o[k];
        prefix += total;
    }
    CCA_REQUIRE(prefix <= P->out_cap, CC_ERR_INVALID_ARG, "output buffer too small for the compressed bricks");
    CCA_REQUIRE(ok, CC_ERR_SYMBOL_OVERFLOW, "some bricks failed to encode (see the per-brick status)");
}

void plan_decode_results(cc_bricks_plan* P, uint32_t* status) {
    Groups& gr = P->gr;
    bool ok = true;
    for (size_t g = 0; g < gr.ctx.size(); g++) {
        std::vector<DecCallOut> res;
        decode_finish(gr.ctx[g]->coder, gr.jobs[g].dj, &res, grouped(P) ? gr.ctx[g]->st : P->st);
        ok = bricks_decode_results(gr.jobs[g], res, status ? status + gr.first[g] : nullptr) && ok;
    }
    CCA_REQUIRE(ok, CC_ERR_CORRUPT_STREAM, "some bricks failed to decode (see the per-brick status)");
}

}  // namespace

void plan_compress(cc_bricks_plan* P, uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(P->inst->device);
    if (grouped(P)) {
        launch_groups(P, true, false);
        CCA_HIP(hipStreamSynchronize(P->st));
    } else {
        CCA_HIP(hipEventRecord(P->order, P->inst->stream));
        CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
        CCA_HIP(hipGraphLaunch(P->x_enc, P->st));
    }
    plan_encode_results(P, offsets, status);
}

void plan_launch(cc_bricks_plan* P) {
    DeviceGuard guard(P->inst->device);
    if (grouped(P)) {
        launch_groups(P, true, true);
    } else {
        CCA_HIP(hipEventRecord(P->order, P->inst->stream));
        CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
        CCA_HIP(hipGraphLaunch(P->x_enc, P->st));
        CCA_HIP(hipGraphLaunch(P->x_dec, P->st));  // stream positions from the device-side results
    }
    P->launched = true;
}

void plan_wait(cc_bricks_plan* P, uint64_t* offsets, uint32_t* cstatus, uint32_t* dstatus) {
    DeviceGuard guard(P->inst->device);
    CCA_REQUIRE(P->launched, CC_ERR_INVALID_ARG, "the plan has no launch to wait for");
    P->launched = false;
    CCA_HIP(hipStreamSynchronize(P->st));
    plan_encode_results(P, offsets, cstatus);
    plan_decode_results(P, dstatus);
}

void plan_run(cc_bricks_plan* P, uint64_t* offsets, uint32_t* cstatus, uint32_t* dstatus) {
    plan_launch(P);
    plan_wait(P, offsets, cstatus, dstatus);
}

void plan_decompress(cc_bricks_plan* P, uint32_t* status) {
    DeviceGuard guard(P->inst->device);
    CCA_REQUIRE(P->encoded, CC_ERR_INVALID_ARG, "the plan has not compressed anything yet");
    if (grouped(P)) {
        launch_groups(P, false, true);
    } else {
        CCA_HIP(hipEventRecord(P->order, P->inst->stream));
        CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
        CCA_HIP(hipGraphLaunch(P->x_dec, P->st));
    }
    CCA_HIP(hipStreamSynchronize(P->st));
    plan_decode_results(P, status);
}

void plan_coding_stats(cc_bricks_plan* P, uint64_t* out) {
    DeviceGuard guard(P->inst->device);
    CCA_REQUIRE(P->encoded, CC_ERR_INVALID_ARG, "the plan has not compressed anything yet");
    CCA_HIP(hipStreamSynchronize(P->st));
    uint64_t n = 0, m = 0, nh = 0, words = 0;
    for (size_t g = 0; g < P->gr.ctx.size(); g++) {
        const BrickJob& job = P->gr.jobs[g];
        const DeviceCoder& D = P->gr.ctx[g]->coder;
        if (grouped(P)) CCA_HIP(hipStreamSynchronize(P->gr.ctx[g]->st));
        std::vector<CodeItem> items(job.ej.nitems);
        std::vector<SeqInfo> seqs(job.ej.nseq);
        CCA_HIP(hipMemcpy(items.data(), D.d_items.p, items.size() * sizeof(CodeItem), hipMemcpyDeviceToHost));
        CCA_HIP(hipMemcpy(seqs.data(), D.d_seqs.p, seqs.size() * sizeof(SeqInfo), hipMemcpyDeviceToHost));
        for (const CodeItem& it : items) {
            n += it.n;
            if (it.rl) m += seqs[it.seq].count;  // pairs: symbols and zero counts are one pair each
            else nh += it.n;
        }
        CoderState s{};
        CCA_HIP(hipMemcpy(&s, D.d_state.p, sizeof(CoderState), hipMemcpyDeviceToHost));
        words += s.out_base + s.out_words;
    }
    out[0] = n;
    out[1] = m;
    out[2] = nh;
    out[3] = words;
}

void plan_destroy(cc_bricks_plan* P) { delete P; }
const void* plan_owner(const cc_bricks_plan* P) { return P ? P->inst : nullptr; }

}  // namespace cca
