// Decode serving loop: compressed bricks stay resident in device memory; requests (lists of
// brick ids with caller-owned output bricks) are admitted brick by brick as they arrive, batched
// by a decode-workspace budget into one brick batch per free slot, and launched asynchronously on
// that slot's stream, so batches of different slots overlap on the device and a large request
// spreads over several slots.  Completion is observed with an event per slot; a request completes
// when the batch holding its last brick does.
//
// No reference counterpart (the reference decodes one volume per call).  Each slot owns a brick
// batch context (host/bricks.cpp), so a request decodes to exactly the values of a batch that
// holds it alone.  At creation every brick's headers are parsed and its decode tables and block
// offsets built once into a resident index (the per-call work of src/cudaCompress/Encode.cpp:
// 522-621); a batch then binds its bricks' index entries and starts at the block decode.
#include "server.hpp"

#include <algorithm>
#include <chrono>
#include <deque>
#include <new>

#include "bricks.hpp"

namespace {

// A request's bricks are admitted one by one: they may land in batches of different slots, and
// the request completes when its last brick's batch does.
struct Request {
    uint64_t id;
    uint32_t remaining;  // bricks not yet decoded
    uint32_t err;        // first decoder error among its bricks
};

struct WorkItem {        // one brick of a request, waiting for a slot
    uint32_t req;        // index into the server's request table
    uint32_t brick;
    float* out;
    uint64_t workspace;
};

struct Slot {
    cca::BrickCtx ctx;
    cca::BrickJob job;
    hipStream_t st = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<uint32_t> brick_req;  // request of every brick of the batch, in batch order
};

uint64_t elems3(const uint32_t* d) { return (uint64_t)d[0] * d[1] * d[2]; }

}  // namespace

struct cc_server {
    cc_instance* inst = nullptr;
    const uint32_t* words = nullptr;
    std::vector<uint64_t> begin, end;   // per brick word range
    std::vector<uint32_t> dims;         // per brick x, y, z
    std::vector<float> steps;
    cca::DecodeIndex index;             // parsed headers, tables and block offsets of every brick
    uint32_t levels = 2;
    int symSize = 2;
    uint64_t slot_budget = 0;
    std::vector<Slot> slots;
    std::deque<WorkItem> pending;       // bricks in arrival order
    std::vector<Request> reqs;          // request table (entries reused through free_reqs)
    std::vector<uint32_t> free_reqs;
    uint32_t live_reqs = 0;             // submitted and not yet finished
    std::vector<std::pair<uint64_t, double>> completed;
    std::vector<std::pair<uint64_t, uint32_t>> failed;
    ~cc_server();
};

cc_server::~cc_server() {
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (inst) (void)hipSetDevice(inst->device);
    for (Slot& s : slots) {
        if (s.st) (void)hipStreamSynchronize(s.st);
        s.ctx.release();
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.st) (void)hipStreamDestroy(s.st);
    }
    if (prev >= 0) (void)hipSetDevice(prev);
}

namespace cca {

namespace {

// decode workspace of one brick: work buffers (2 N + N/7 floats), symbols, coder arena
uint64_t brick_workspace(uint64_t N, int symSize) { return N * 4 * 2 + N * 4 / 7 + N * symSize + N * 6 + 64 * 1024; }

// One brick of request r finished (err: its decoder status).
void finish_brick(cc_server* S, uint32_t r, uint32_t err, double now) {
    Request& q = S->reqs[r];
    if (err && !q.err) q.err = err;
    if (--q.remaining) return;
    if (q.err) S->failed.emplace_back(q.id, q.err);
    else S->completed.emplace_back(q.id, now);
    S->free_reqs.push_back(r);
    S->live_reqs--;
}

// Every slot's buffers sized up front for the largest batches admission can form: the slot's
// budget filled with copies of the largest brick, and with copies of the smallest (the most
// bricks; at most kWarmBricks).  A buffer that grew while serving would free its old allocation,
// and hipFree waits for the whole device, every other slot's batch with it.  Best effort: a warm
// batch that does not fit the device leaves the slot to grow on demand.
constexpr uint64_t kWarmBricks = 4096;
void warm_slots(cc_server* S) {
    const uint32_t n = (uint32_t)S->begin.size();
    if (n == 0) return;
    uint32_t lo = 0, hi = 0;
    for (uint32_t i = 1; i < n; i++) {
        const uint64_t e = elems3(&S->dims[3 * i]);
        if (e < elems3(&S->dims[3 * lo])) lo = i;
        if (e > elems3(&S->dims[3 * hi])) hi = i;
    }
    for (const uint32_t pick : {hi, lo}) {
        const uint32_t* d = &S->dims[3 * pick];
        const uint64_t w = brick_workspace(elems3(d), S->symSize);
        const uint32_t k = (uint32_t)std::min<uint64_t>(kWarmBricks, std::max<uint64_t>(1, S->slot_budget / w));
        const std::vector<uint32_t> ids(k, pick);
        for (Slot& s : S->slots) {
            BrickJob job;
            job.L = S->levels;
            job.symSize = S->symSize;
            // destinations are recorded, never written: nothing is launched
            job.bricks.assign(k, cc_brick{nullptr, reinterpret_cast<float*>(const_cast<uint32_t*>(S->words)), d[0], d[1],
                                          d[2], S->steps[pick]});
            try {
                bricks_prepare_bound(S->inst, s.ctx, job, S->index, ids, s.st);
            } catch (const Error& e) {
                if (e.status != CC_ERR_OUT_OF_MEMORY) throw;
            }
            CCA_HIP(hipStreamSynchronize(s.st));
        }
    }
}

void reap(cc_server* S, Slot& s, double now) {
    if (!s.busy) return;
    const hipError_t q = hipEventQuery(s.done);
    if (q == hipErrorNotReady) return;
    CCA_HIP(q);
    // per-brick statuses were copied to the pinned result buffer before the event
    const DecCallOut* r = (const DecCallOut*)((const unsigned char*)s.ctx.coder.h_dresult.p + sizeof(CoderState));
    for (size_t k = 0; k < s.brick_req.size(); k++) finish_brick(S, s.brick_req[k], r[k].err, now);
    s.busy = false;
    s.brick_req.clear();
}

void launch(cc_server* S, Slot& s, double now) {
    // admit bricks in arrival order while the batch fits the slot's workspace budget (at least
    // one); the batch is built aside and becomes the slot's only once it is on the device
    std::vector<uint32_t> ids, breq;
    BrickJob job;
    job.L = S->levels;
    job.symSize = S->symSize;
    uint64_t ws = 0;
    while (!S->pending.empty()) {
        const WorkItem& w = S->pending.front();
        if (!ids.empty() && ws + w.workspace > S->slot_budget) break;
        ws += w.workspace;
        const uint32_t* d = &S->dims[3 * w.brick];
        job.bricks.push_back(cc_brick{nullptr, w.out, d[0], d[1], d[2], S->steps[w.brick]});
        ids.push_back(w.brick);
        breq.push_back(w.req);
        S->pending.pop_front();
    }
    if (ids.empty()) return;
    try {
        s.job = std::move(job);
        bricks_prepare_bound(S->inst, s.ctx, s.job, S->index, ids, s.st);
        bricks_decode_launch(s.ctx, s.job, s.st, false);
        // statuses to the pinned buffer, then the completion event
        DeviceCoder& D = s.ctx.coder;
        const uint32_t ncalls = s.job.bj.ncalls;  // one call per brick of the batch
        unsigned char* r = (unsigned char*)D.h_dresult.get(sizeof(CoderState) + sizeof(DecCallOut) * ncalls);
        CCA_HIP(hipMemcpyAsync(r + sizeof(CoderState), D.d_dout.p, sizeof(DecCallOut) * ncalls,
                               hipMemcpyDeviceToHost, s.st));
        CCA_HIP(hipEventRecord(s.done, s.st));
    } catch (const Error& e) {
        // the admitted bricks fail their requests with the launch's status; the slot stays free
        for (uint32_t r : breq) finish_brick(S, r, (uint32_t)e.status, now);
        s.job = BrickJob();
        throw;
    } catch (const std::bad_alloc&) {
        for (uint32_t r : breq) finish_brick(S, r, (uint32_t)CC_ERR_OUT_OF_MEMORY, now);
        s.job = BrickJob();
        throw;
    }
    s.brick_req = std::move(breq);
    s.busy = true;
}

}  // namespace

cc_server* server_create(cc_instance* I, const uint32_t* d_words, uint32_t nbricks, const uint64_t* ranges,
                         const uint32_t* dims, const float* steps, uint32_t levels, int symSize, uint32_t nslots,
                         uint64_t workspace) {
    DeviceGuard guard(I->device);
    CCA_REQUIRE(nslots >= 1 && nslots <= 16, CC_ERR_INVALID_ARG, "slots must be in [1, 16]");
    CCA_REQUIRE(d_words && ranges && dims && steps, CC_ERR_INVALID_ARG, "null argument");
    std::unique_ptr<cc_server> S(new cc_server());
    S->inst = I;
    S->words = d_words;
    S->levels = levels;
    S->symSize = symSize;
    S->begin.resize(nbricks);
    S->end.resize(nbricks);
    for (uint32_t i = 0; i < nbricks; i++) {
        S->begin[i] = ranges[2 * i];
        S->end[i] = ranges[2 * i + 1];
        CCA_REQUIRE(S->end[i] >= S->begin[i], CC_ERR_INVALID_ARG, "bad brick word range");
    }
    S->dims.assign(dims, dims + 3 * (size_t)nbricks);
    S->steps.assign(steps, steps + nbricks);
    S->slot_budget = workspace / nslots;
    S->slots.resize(nslots);
    for (Slot& s : S->slots) {
        CCA_HIP(hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking));
        CCA_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    }
    std::vector<cc_brick> shapes(nbricks);
    std::vector<CallWords> calls(nbricks);
    for (uint32_t i = 0; i < nbricks; i++) {
        shapes[i] = cc_brick{nullptr, nullptr, dims[3 * i], dims[3 * i + 1], dims[3 * i + 2], steps[i]};
        calls[i] = CallWords{d_words + S->begin[i], S->end[i] - S->begin[i]};
    }
    bricks_index_build(I, S->index, shapes, levels, symSize, calls, S->slots[0].st);
    warm_slots(S.get());
    return S.release();
}

void server_submit(cc_server* S, uint64_t id, const uint32_t* bricks, uint32_t count, float* const* outs,
                   double arrival) {
    (void)arrival;
    CCA_REQUIRE(count >= 1 && bricks && outs, CC_ERR_INVALID_ARG, "a request names at least one brick");
    for (uint32_t i = 0; i < count; i++) {
        CCA_REQUIRE(bricks[i] < S->begin.size(), CC_ERR_INVALID_ARG, "brick id out of range");
        CCA_REQUIRE(outs[i], CC_ERR_INVALID_ARG, "null output brick");
    }
    uint32_t r;
    if (!S->free_reqs.empty()) {
        r = S->free_reqs.back();
        S->free_reqs.pop_back();
    } else {
        r = (uint32_t)S->reqs.size();
        S->reqs.emplace_back();
    }
    S->reqs[r] = Request{id, count, 0u};
    S->live_reqs++;
    for (uint32_t i = 0; i < count; i++)
        S->pending.push_back(WorkItem{r, bricks[i], outs[i], brick_workspace(elems3(&S->dims[3 * bricks[i]]), S->symSize)});
}

uint32_t server_poll(cc_server* S, double now) {
    DeviceGuard guard(S->inst->device);
    uint32_t launched = 0;
    for (Slot& s : S->slots) reap(S, s, now);
    for (Slot& s : S->slots) {
        if (S->pending.empty()) break;
        if (s.busy) continue;
        launch(S, s, now);
        launched++;
    }
    return launched;
}

void server_destroy(cc_server* S) { delete S; }
const void* server_owner(const cc_server* S) { return S ? S->inst : nullptr; }

uint32_t server_busy(const cc_server* S) {
    uint32_t b = 0;
    for (const Slot& s : S->slots) b += s.busy ? 1 : 0;
    return b + S->live_reqs;
}

uint32_t server_take_completed(cc_server* S, uint64_t* ids, double* times, uint32_t max) {
    const uint32_t n = (uint32_t)std::min<size_t>(max, S->completed.size());
    for (uint32_t i = 0; i < n; i++) {
        if (ids) ids[i] = S->completed[i].first;
        if (times) times[i] = S->completed[i].second;
    }
    S->completed.erase(S->completed.begin(), S->completed.begin() + n);
    return n;
}

uint32_t server_take_failed(cc_server* S, uint64_t* ids, uint32_t* status, uint32_t max) {
    const uint32_t n = (uint32_t)std::min<size_t>(max, S->failed.size());
    for (uint32_t i = 0; i < n; i++) {
        if (ids) ids[i] = S->failed[i].first;
        if (status) status[i] = S->failed[i].second == kDecCorrupt ? (uint32_t)CC_ERR_CORRUPT_STREAM
                                                                  : (uint32_t)CC_ERR_OUT_OF_MEMORY;
    }
    S->failed.erase(S->failed.begin(), S->failed.begin() + n);
    return n;
}

void server_run(cc_server* S, uint32_t nreq, const double* arrivals, const uint32_t* counts, const uint32_t* ids,
                float* const* outs, double* latency, double* elapsed) {
    CCA_REQUIRE(server_busy(S) == 0 && S->completed.empty() && S->failed.empty(), CC_ERR_INVALID_ARG,
                "the server must be idle with no unclaimed results");
    CCA_REQUIRE(nreq == 0 || (arrivals && counts && ids && outs && latency), CC_ERR_INVALID_ARG, "null argument");
    std::vector<uint64_t> first(nreq + 1, 0);
    for (uint32_t r = 0; r < nreq; r++) {
        CCA_REQUIRE(r == 0 || arrivals[r] >= arrivals[r - 1], CC_ERR_INVALID_ARG, "arrivals must be sorted");
        CCA_REQUIRE(counts[r] >= 1, CC_ERR_INVALID_ARG, "a request names at least one brick");
        first[r + 1] = first[r] + counts[r];
    }
    // the whole trace is checked before anything is submitted
    for (uint64_t k = 0; k < first[nreq]; k++) {
        CCA_REQUIRE(ids[k] < S->begin.size(), CC_ERR_INVALID_ARG, "brick id out of range");
        CCA_REQUIRE(outs[k], CC_ERR_INVALID_ARG, "null output brick");
    }
    for (uint32_t r = 0; r < nreq; r++) latency[r] = -1.0;
    using clk = std::chrono::steady_clock;
    const clk::time_point t0 = clk::now();
    auto now = [&] { return std::chrono::duration<double>(clk::now() - t0).count(); };
    uint32_t next = 0, done = 0;
    double t_last = 0.0;
    while (done < nreq) {
        double t = now();
        while (next < nreq && arrivals[next] <= t) {
            server_submit(S, next, ids + first[next], counts[next], outs + first[next], arrivals[next]);
            next++;
        }
        server_poll(S, now());
        for (const auto& c : S->completed) {
            latency[c.first] = c.second - arrivals[c.first];
            t_last = std::max(t_last, c.second);
            done++;
        }
        S->completed.clear();
        for (const auto& f : S->failed) {
            latency[f.first] = -1.0;
            done++;
        }
        S->failed.clear();
    }
    if (elapsed) *elapsed = t_last;
}

}  // namespace cca
