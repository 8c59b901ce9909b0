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
        auto ptr = [&](uint32_t k) { return (void*)(sym + (job.sym[i] + stream_off(N, L, k)) * job.symSize); };
        items.push_back(ItemSpec{ptr(0), (uint32_t)(N >> (3 * L)), false, i});
        uint32_t k = 1;
        for (int l = (int)L - 1; l >= 0; l--) {
            const uint32_t nb = (uint32_t)(N >> (3 * (l + 1)));
            for (uint32_t band = 0; band < 7; band++) items.push_back(ItemSpec{ptr(k++), nb, true, i});
        }
    }
    return items;
}

void upload_levels(BrickCtx& ctx, BrickJob& job, float* work, unsigned char* sym, hipStream_t st) {
    const uint32_t n = (uint32_t)job.bricks.size(), L = job.L;
    const size_t lv_bytes = sizeof(BrickLevel) * (size_t)L * n;
    const size_t tb_off = align_up(std::max<size_t>(lv_bytes, 1), 256);
    const size_t bytes = tb_off + 4ull * L * 3 * (n + 1);
    unsigned char* h = (unsigned char*)ctx.h_blevels.get(bytes);
    BrickLevel* hl = (BrickLevel*)h;
    uint32_t* ht = (uint32_t*)(h + tb_off);
    job.tiles_total.assign(3 * L, 0);
    for (uint32_t l = 0; l < L; l++) {
        uint32_t acc[3] = {0, 0, 0};
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
            for (int ax = 0; ax < 3; ax++) {
                ht[((size_t)l * 3 + ax) * (n + 1) + i] = acc[ax];
                acc[ax] += brick_tiles(ax, e.lx, e.ly, e.lz);
            }
        }
        for (int ax = 0; ax < 3; ax++) {
            ht[((size_t)l * 3 + ax) * (n + 1) + n] = acc[ax];
            job.tiles_total[3 * l + ax] = acc[ax];
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
    upload_levels(ctx, job, work, sym, st);
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

void bricks_encode_launch(BrickCtx& ctx, const BrickJob& job, hipStream_t st) {
    const uint32_t n = (uint32_t)job.bricks.size();
    for (uint32_t l = 0; l < job.L; l++)
        bricks_forward_level(job.d_levels + (size_t)l * n, job.d_tiles + (size_t)l * 3 * (n + 1), n,
                             &job.tiles_total[3 * l], job.symSize, job.level_elems[l], st);
    encode_launch(ctx.coder, job.ej, job.d_out, job.out_cap, st);
}

void bricks_decode_launch(BrickCtx& ctx, const BrickJob& job, hipStream_t st, bool from_encode) {
    const uint32_t n = (uint32_t)job.bricks.size();
    if (from_encode)
        decode_calls_from_encode(ctx.coder, job.dj, (const CallOut*)ctx.coder.d_calls.p, job.d_out, st);
    decode_launch(ctx.coder, job.dj, st);
    for (int l = (int)job.L - 1; l >= 0; l--)
        bricks_inverse_level(job.d_levels + (size_t)l * n, job.d_tiles + (size_t)l * 3 * (n + 1), n,
                             &job.tiles_total[3 * l], job.symSize, job.level_elems[l], st);
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

void compress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize, uint32_t* d_out,
                     uint64_t cap, uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    BrickJob job;
    job.bricks.assign(bricks, bricks + n);
    job.L = L;
    job.symSize = symSize;
    job.d_out = d_out;
    job.out_cap = cap;
    bricks_prepare(I, I->batch, job, true, false, nullptr, st);
    std::vector<CallOut> res;
    CoderState s{};
    for (int attempt = 0;; attempt++) {
        bricks_encode_launch(I->batch, job, st);
        s = encode_finish(I->batch.coder, job.ej, &res, st);
        if (s.err != kErrScratch || attempt >= 4) break;
        encode_reserve(I->batch.coder, job.ej, std::max<uint64_t>(2 * job.ej.scratch_cap, s.scratch_used + s.scratch_used / 2));
    }
    CCA_REQUIRE(s.out_words <= cap, CC_ERR_INVALID_ARG, "output buffer too small for the compressed bricks");
    const bool ok = bricks_encode_results(job, res, s.out_words, offsets, status);
    CCA_REQUIRE(ok, CC_ERR_SYMBOL_OVERFLOW, "some bricks failed to encode (see the per-brick status)");
}

void decompress_bricks(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize,
                       const uint32_t* d_in, const uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(I->device);
    hipStream_t st = I->stream;
    BrickJob job;
    job.bricks.assign(bricks, bricks + n);
    job.L = L;
    job.symSize = symSize;
    std::vector<CallWords> calls(n);
    for (uint32_t i = 0; i < n; i++) {
        CCA_REQUIRE(offsets[i + 1] >= offsets[i], CC_ERR_INVALID_ARG, "brick offsets must be nondecreasing");
        calls[i] = CallWords{d_in + offsets[i], offsets[i + 1] - offsets[i]};
    }
    bricks_prepare(I, I->batch, job, false, true, &calls, st);
    bricks_decode_launch(I->batch, job, st, false);
    std::vector<DecCallOut> res;
    decode_finish(I->batch.coder, job.dj, &res, st);
    CCA_REQUIRE(bricks_decode_results(job, res, status), CC_ERR_CORRUPT_STREAM,
                "some bricks failed to decode (see the per-brick status)");
}

}  // namespace cca

// ------------------------------------------------------------------------------------- plans
struct cc_bricks_plan {
    cc_instance* inst = nullptr;
    cca::BrickCtx ctx;
    cca::BrickJob job;
    hipStream_t st = nullptr;
    hipGraph_t g_enc = nullptr, g_dec = nullptr;
    hipGraphExec_t x_enc = nullptr, x_dec = nullptr;
    hipEvent_t order = nullptr;  // orders replays after the instance stream's earlier work
    bool encoded = false;
    ~cc_bricks_plan();
};

cc_bricks_plan::~cc_bricks_plan() {
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (inst) (void)hipSetDevice(inst->device);
    if (st) (void)hipStreamSynchronize(st);
    if (x_enc) (void)hipGraphExecDestroy(x_enc);
    if (x_dec) (void)hipGraphExecDestroy(x_dec);
    if (g_enc) (void)hipGraphDestroy(g_enc);
    if (g_dec) (void)hipGraphDestroy(g_dec);
    if (order) (void)hipEventDestroy(order);
    ctx.release();
    if (st) (void)hipStreamDestroy(st);
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
}  // namespace

cc_bricks_plan* plan_create(cc_instance* I, const cc_brick* bricks, uint32_t n, uint32_t L, int symSize,
                            uint32_t* d_out, uint64_t cap) {
    DeviceGuard guard(I->device);
    std::unique_ptr<cc_bricks_plan> P(new cc_bricks_plan());
    P->inst = I;
    CCA_HIP(hipStreamCreateWithFlags(&P->st, hipStreamNonBlocking));
    CCA_HIP(hipEventCreateWithFlags(&P->order, hipEventDisableTiming));
    CCA_HIP(hipEventRecord(P->order, I->stream));
    CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
    BrickJob& job = P->job;
    job.bricks.assign(bricks, bricks + n);
    job.L = L;
    job.symSize = symSize;
    job.d_out = d_out;
    job.out_cap = cap;
    bricks_prepare(I, P->ctx, job, true, true, nullptr, P->st);
    // one eager pass sizes the arena (with headroom for other data of the same geometry) and
    // allocates every buffer before capture
    std::vector<CallOut> res;
    for (int attempt = 0;; attempt++) {
        bricks_encode_launch(P->ctx, job, P->st);
        const CoderState s = encode_finish(P->ctx.coder, job.ej, &res, P->st);
        if (s.err != kErrScratch || attempt >= 4) {
            encode_reserve(P->ctx.coder, job.ej, std::max<uint64_t>(job.ej.scratch_cap, s.scratch_used + s.scratch_used / 4));
            break;
        }
        encode_reserve(P->ctx.coder, job.ej, std::max<uint64_t>(2 * job.ej.scratch_cap, s.scratch_used + s.scratch_used / 2));
    }
    capture(P->st, &P->g_enc, &P->x_enc, [&] { bricks_encode_launch(P->ctx, job, P->st); });
    capture(P->st, &P->g_dec, &P->x_dec, [&] { bricks_decode_launch(P->ctx, job, P->st, true); });
    return P.release();
}

void plan_compress(cc_bricks_plan* P, uint64_t* offsets, uint32_t* status) {
    DeviceGuard guard(P->inst->device);
    CCA_HIP(hipEventRecord(P->order, P->inst->stream));
    CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
    CCA_HIP(hipGraphLaunch(P->x_enc, P->st));
    std::vector<CallOut> res;
    const CoderState s = encode_finish(P->ctx.coder, P->job.ej, &res, P->st);
    P->encoded = true;
    CCA_REQUIRE(s.err != kErrScratch, CC_ERR_OUT_OF_MEMORY,
                "plan arena too small for this data; create a new plan");
    CCA_REQUIRE(s.out_words <= P->job.out_cap, CC_ERR_INVALID_ARG, "output buffer too small for the compressed bricks");
    const bool ok = bricks_encode_results(P->job, res, s.out_words, offsets, status);
    CCA_REQUIRE(ok, CC_ERR_SYMBOL_OVERFLOW, "some bricks failed to encode (see the per-brick status)");
}

void plan_decompress(cc_bricks_plan* P, uint32_t* status) {
    DeviceGuard guard(P->inst->device);
    CCA_REQUIRE(P->encoded, CC_ERR_INVALID_ARG, "the plan has not compressed anything yet");
    CCA_HIP(hipEventRecord(P->order, P->inst->stream));
    CCA_HIP(hipStreamWaitEvent(P->st, P->order, 0));
    CCA_HIP(hipGraphLaunch(P->x_dec, P->st));
    std::vector<DecCallOut> res;
    decode_finish(P->ctx.coder, P->job.dj, &res, P->st);
    CCA_REQUIRE(bricks_decode_results(P->job, res, status), CC_ERR_CORRUPT_STREAM,
                "some bricks failed to decode (see the per-brick status)");
}

void plan_destroy(cc_bricks_plan* P) { delete P; }

}  // namespace cca
