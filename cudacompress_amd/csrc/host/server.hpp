// Decode serving loop (host/server.cpp).
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "instance.hpp"

namespace cca {

// ranges: 2 per brick (begin, end word); dims: 3 per brick (x, y, z)
cc_server* server_create(cc_instance* I, const uint32_t* d_words, uint32_t nbricks, const uint64_t* ranges,
                         const uint32_t* dims, const float* steps, uint32_t levels, int symSize, uint32_t nslots,
                         uint64_t workspace);
void server_submit(cc_server* S, uint64_t id, const uint32_t* bricks, uint32_t count, float* const* outs, double arrival);
// reaps finished batches (completion time `now`) and launches pending requests on free slots
uint32_t server_poll(cc_server* S, double now);
void server_destroy(cc_server* S);
// the instance a server belongs to (its launches are that instance's in the kernel timing)
const void* server_owner(const cc_server* S);
// slots in flight + requests waiting
uint32_t server_busy(const cc_server* S);
uint32_t server_take_completed(cc_server* S, uint64_t* ids, double* times, uint32_t max);
uint32_t server_take_failed(cc_server* S, uint64_t* ids, uint32_t* status, uint32_t max);
// drives an idle server through a request trace in real time (seconds from the call's start);
// latency[r] = completion - arrival, -1 for a failed request; elapsed = last completion
void server_run(cc_server* S, uint32_t nreq, const double* arrivals, const uint32_t* counts, const uint32_t* ids,
                float* const* outs, double* latency, double* elapsed);

}  // namespace cca
