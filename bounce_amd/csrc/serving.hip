// Serving loop: many worlds' narrowphase steps on one context, continuously batched.
//
// The reference runs one b3World per b3ContactManager::UpdateContacts call
// (contact_manager.cpp:167-231); a server runs the narrowphase of many such worlds on one
// GPU.  A world keeps its contact list, feature caches and manifolds in HBM; a request is
// one step of one world (its body transforms, optionally sweeps).  A worker thread packs
// whole pending requests, oldest first and at most one per world, into one device batch of
// up to max_batch_pairs contacts; a gather kernel expands each request's body transforms
// into per-contact inputs, the batch kernels (cache, convex, SAT, contact) run with each
// contact's cache and manifold slot in the world's slab range, and the request completes
// when the batch's event has fired.  Up to three batches are in flight, each on its own
// lane (streams, device inputs and scratch), so they overlap on the device while the next
// is packed; a batch holding a world's step first waits for the batch holding that
// world's previous step (an event wait on the device), which keeps each world's cache
// chain in order.
//
// Results do not depend on the batching: every contact is computed from its own inputs
// and cache slot only, so a world's manifolds are bit-identical alone or in any mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bounce_amd.h"
#include "context.h"
#include "kernels.h"

namespace b3amd {

// One request inside a batch (device copy).
struct ReqDesc {
    uint32_t pairOffset;  // first batch contact of the request
    uint32_t pairCount;
    uint32_t bodyOffset;  // first body of the request in the batch body arrays
    uint32_t slotBase;    // the world's first cache/manifold slot
    const uint32_t* bodyShapes;
    const uint32_t* bodyPairs;
};

struct GatherArgs {
    uint32_t n;              // contacts in the batch
    uint32_t requests;
    const ReqDesc* descs;    // sorted by pairOffset
    const float* bodyXfs;    // [bodies][12]
    const float* bodySweeps; // [bodies][8] (orientation0, orientation) or null
    uint32_t* pairs;         // [n][2] out
    float* xfs;              // [n][24] out
    float* sweeps;           // [n][16] out, or null
    uint32_t* slots;         // [n] out
};

// Per contact: find its request, then read the world's contact (body indices) and the
// bodies' shapes and transforms.
__global__ __launch_bounds__(256) void gather_kernel(GatherArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    uint32_t lo = 0, hi = a.requests;  // last desc with pairOffset <= k
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.descs[mid].pairOffset <= k) lo = mid;
        else hi = mid;
    }
    const ReqDesc d = a.descs[lo];
    const uint32_t i = k - d.pairOffset;
    const uint32_t ba = d.bodyPairs[2 * i], bb = d.bodyPairs[2 * i + 1];
    a.pairs[2 * k] = d.bodyShapes[ba];
    a.pairs[2 * k + 1] = d.bodyShapes[bb];
    const float* xa = a.bodyXfs + 12 * (size_t)(d.bodyOffset + ba);
    const float* xb = a.bodyXfs + 12 * (size_t)(d.bodyOffset + bb);
    float* xo = a.xfs + 24 * (size_t)k;
#pragma unroll
    for (int t = 0; t < 12; ++t) {
        xo[t] = xa[t];
        xo[12 + t] = xb[t];
    }
    if (a.sweeps) {
        float* so = a.sweeps + 16 * (size_t)k;
        if (a.bodySweeps) {
            const float* sa = a.bodySweeps + 8 * (size_t)(d.bodyOffset + ba);
            const float* sb = a.bodySweeps + 8 * (size_t)(d.bodyOffset + bb);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                so[t] = sa[t];
                so[8 + t] = sb[t];
            }
        }
    }
    a.slots[k] = d.slotBase + i;
}

// One lane per contact.
__global__ __launch_bounds__(256) void body_gather_kernel(BodyGatherArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const uint32_t ba = a.contacts[2 * k], bb = a.contacts[2 * k + 1];
    const bool ok = ba < a.bodyCount && bb < a.bodyCount;
    const uint2 sh = ok ? make_uint2(a.bodyShapes[ba], a.bodyShapes[bb]) : make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
    reinterpret_cast<uint2*>(a.pairs)[k] = sh;
}

hipError_t launch_body_gather(const BodyGatherArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    body_gather_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd

namespace {

using Clock = std::chrono::steady_clock;

struct World {
    bool live = false;
    bool closing = false;  // b3amd_world_destroy is waiting: no further submissions
    uint32_t bodyCount = 0, pairCount = 0;
    uint32_t* dBodyShapes = nullptr;
    uint32_t* dBodyPairs = nullptr;
    uint32_t slotBase = 0;
    uint32_t active = 0;    // requests queued or in flight
    uint64_t lastBatch = 0; // sequence number of the in-flight batch holding its latest request (0: none)
};

struct Request {
    uint64_t ticket = 0;
    uint32_t world = 0;
    const float* xfs = nullptr;     // [bodies][12]
    const float* sweeps = nullptr;  // [bodies][8] or null
    std::vector<float> ownXfs, ownSweeps;
    Clock::time_point submit;
    double latencyMs = 0.0;
    bool done = false;
};

// One pipeline lane.  A batch in flight owns a lane: its two streams (default priority),
// its device inputs and scratch and its pinned host staging, so batches on different lanes
// run concurrently on the device.
struct Lane {
    BatchStreams streams{nullptr, nullptr, nullptr, nullptr};
    float* dBodyXfs = nullptr;
    float* dBodySweeps = nullptr;
    ReqDesc* dDescs = nullptr;
    uint32_t* dPairs = nullptr;
    float* dXfs = nullptr;
    float* dSweeps = nullptr;
    uint32_t* dSlots = nullptr;
    BatchScratch scratch{nullptr, nullptr, nullptr};
    float* hBodyXfs = nullptr;
    float* hBodySweeps = nullptr;
    ReqDesc* hDescs = nullptr;
    // Recorded around each batch the lane runs: created once with the lane (event creation
    // and destruction per batch cost milliseconds now and then in the runtime).  A lane runs
    // one batch at a time, and a later batch's wait on `done` is enqueued while the batch that
    // recorded it is still in flight.
    hipEvent_t start = nullptr, done = nullptr;
    bool busy = false;
};

struct Batch {
    std::vector<Request*> requests;
    std::vector<uint32_t> bodyCounts;  // per request
    std::vector<hipEvent_t> waits;     // done events of in-flight batches holding an earlier step of one of its worlds
    uint32_t pairs = 0, bodies = 0;
    uint64_t seq = 0;
    hipEvent_t start = nullptr, done = nullptr;  // around the batch's copies and kernels
    int lane = 0;
};

constexpr int kMaxInflight = 3;

}  // namespace

struct b3amd_server {
    b3amd_context* ctx = nullptr;
    uint32_t maxPairs = 0, maxBodies = 0, maxRequests = 0, slabSlots = 0;

    // Device state shared by all lanes: the worlds' cache and manifold slots.
    ConvexCache* dCaches = nullptr;  // [slabSlots]
    Manifold* dManifolds = nullptr;  // [slabSlots]
    Lane lanes[kMaxInflight];
    uint32_t batchPairs = 0;  // current cap on contacts per batch (<= maxPairs)
    uint32_t maxInflight = 2;

    // Worlds and their slot ranges (first fit over free ranges).
    std::vector<World> worlds;
    std::vector<std::pair<uint32_t, uint32_t>> freeRanges;  // (offset, length), sorted
    std::deque<uint64_t> admitQueue;                         // worlds waiting for slots, in arrival order
    uint64_t nextAdmit = 1;

    // Requests.
    std::mutex mu;
    std::condition_variable cvWork, cvDone, cvAdmit;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, Request*> requests;
    std::deque<Batch> inflight;
    uint64_t nextTicket = 1, nextBatch = 1;
    bool stop = false;
    std::thread worker;
    int workerError = B3AMD_OK;
    std::string workerMessage;

    // Statistics.
    b3amd_server_stats stats{};
};

namespace {

// Empty b3ConvexCache records (b3SimplexCache zeroed, feature state empty) for a slot range.
__global__ void init_slots_kernel(ConvexCache* caches, Manifold* manifolds, uint32_t base, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    ConvexCache c;
    for (int t = 0; t < 5; ++t) c.simplex[t] = 0u;
    c.feature.state = kCacheEmpty;
    c.feature.type = kFeatureEdge1;
    c.feature.index1 = 0u;
    c.feature.index2 = 0u;
    caches[base + i] = c;
    manifolds[base + i].pointCount = 0u;
    manifolds[base + i].flags = 0u;
}

bool fits(const b3amd_server* s, uint32_t n) {
    for (const auto& r : s->freeRanges)
        if (r.second >= n) return true;
    return false;
}

bool take_range(b3amd_server* s, uint32_t n, uint32_t* base) {
    for (size_t i = 0; i < s->freeRanges.size(); ++i) {
        auto& r = s->freeRanges[i];
        if (r.second >= n) {
            *base = r.first;
            r.first += n;
            r.second -= n;
            if (r.second == 0) s->freeRanges.erase(s->freeRanges.begin() + (long)i);
            return true;
        }
    }
    return false;
}

void give_range(b3amd_server* s, uint32_t base, uint32_t n) {
    auto& fr = s->freeRanges;
    size_t i = 0;
    while (i < fr.size() && fr[i].first < base) ++i;
    fr.insert(fr.begin() + (long)i, {base, n});
    // Merge with the neighbours.
    if (i + 1 < fr.size() && fr[i].first + fr[i].second == fr[i + 1].first) {
        fr[i].second += fr[i + 1].second;
        fr.erase(fr.begin() + (long)i + 1);
    }
    if (i > 0 && fr[i - 1].first + fr[i - 1].second == fr[i].first) {
        fr[i - 1].second += fr[i].second;
        fr.erase(fr.begin() + (long)i);
    }
}

// Takes the oldest pending requests, at most one per world, up to the batch caps, and
// writes their descriptors into the lane's staging buffer (lock held: world records are
// only read and updated here).  A world whose previous step is still in flight makes the
// batch wait for that batch's done event, so each world's steps run in order on its slots.
// Returns the number taken.
uint32_t select_batch(b3amd_server* s, Batch& batch) {
    Lane& lane = s->lanes[batch.lane];
    std::vector<uint32_t> seen;  // worlds already taken or passed over in this batch
    uint32_t pairs = 0, bodies = 0, nreq = 0;
    for (auto it = s->pending.begin(); it != s->pending.end();) {
        Request* r = *it;
        World& w = s->worlds[r->world];
        bool dup = false;
        for (uint32_t v : seen)
            if (v == r->world) dup = true;
        if (dup) {  // keep the world's requests in order: one per batch
            ++it;
            continue;
        }
        seen.push_back(r->world);
        if (nreq == s->maxRequests || (nreq > 0 && pairs + w.pairCount > s->batchPairs) ||
            pairs + w.pairCount > s->maxPairs || bodies + w.bodyCount > s->maxBodies)
            break;
        ReqDesc& d = lane.hDescs[nreq];
        d.pairOffset = pairs;
        d.pairCount = w.pairCount;
        d.bodyOffset = bodies;
        d.slotBase = w.slotBase;
        d.bodyShapes = w.dBodyShapes;
        d.bodyPairs = w.dBodyPairs;
        if (w.lastBatch != 0) {
            for (const Batch& b : s->inflight) {
                if (b.seq != w.lastBatch) continue;
                bool have = false;
                for (hipEvent_t e : batch.waits) have = have || e == b.done;
                if (!have) batch.waits.push_back(b.done);
            }
        }
        w.lastBatch = batch.seq;
        batch.bodyCounts.push_back(w.bodyCount);
        pairs += w.pairCount;
        bodies += w.bodyCount;
        ++nreq;
        batch.requests.push_back(r);
        it = s->pending.erase(it);
    }
    batch.pairs = pairs;
    batch.bodies = bodies;
    return nreq;
}

// Copies the batch's transforms into its lane's staging buffer and enqueues the lane's
// copies and kernels (lock not held: the requests' inputs and the worlds' device arrays
// stay unchanged until the batch completes, and only the worker touches the lane).
void launch_batch(b3amd_server* s, Batch& batch) {
    const Clock::time_point t0 = Clock::now();
    Lane& L = s->lanes[batch.lane];
    const uint32_t pairs = batch.pairs, bodies = batch.bodies;
    const uint32_t nreq = (uint32_t)batch.requests.size();
    bool anySweeps = false;
    for (uint32_t q = 0; q < nreq; ++q) {
        const Request* r = batch.requests[q];
        const size_t off = L.hDescs[q].bodyOffset, nb = batch.bodyCounts[q];
        std::memcpy(L.hBodyXfs + 12 * off, r->xfs, 48 * nb);
        if (r->sweeps) {
            std::memcpy(L.hBodySweeps + 8 * off, r->sweeps, 32 * nb);
            anySweeps = true;
        }
    }
    if (anySweeps) {
        // Requests without sweeps get identity orientations (what a null sweep array means).
        for (uint32_t q = 0; q < nreq; ++q) {
            if (batch.requests[q]->sweeps) continue;
            float* o = L.hBodySweeps + 8 * (size_t)L.hDescs[q].bodyOffset;
            for (uint32_t b = 0; b < batch.bodyCounts[q]; ++b, o += 8) {
                o[0] = o[1] = o[2] = 0.0f, o[3] = 1.0f;
                o[4] = o[5] = o[6] = 0.0f, o[7] = 1.0f;
            }
        }
    }
    b3amd_context* ctx = s->ctx;
    hipStream_t stream = L.streams.main;
    bool ok = true;
    for (hipEvent_t e : batch.waits) ok = ok && hipStreamWaitEvent(stream, e, 0) == hipSuccess;
    ok = ok && hipEventRecord(batch.start, stream) == hipSuccess &&
         hipMemcpyAsync(L.dBodyXfs, L.hBodyXfs, 48ull * bodies, hipMemcpyHostToDevice, stream) == hipSuccess &&
         (!anySweeps ||
          hipMemcpyAsync(L.dBodySweeps, L.hBodySweeps, 32ull * bodies, hipMemcpyHostToDevice, stream) == hipSuccess) &&
         hipMemcpyAsync(L.dDescs, L.hDescs, sizeof(ReqDesc) * nreq, hipMemcpyHostToDevice, stream) == hipSuccess;
    if (ok) {
        GatherArgs g;
        g.n = pairs;
        g.requests = nreq;
        g.descs = L.dDescs;
        g.bodyXfs = L.dBodyXfs;
        g.bodySweeps = anySweeps ? L.dBodySweeps : nullptr;
        g.pairs = L.dPairs;
        g.xfs = L.dXfs;
        g.sweeps = anySweeps ? L.dSweeps : nullptr;
        g.slots = L.dSlots;
        gather_kernel<<<(pairs + 255) / 256, 256, 0, stream>>>(g);
        ok = hipGetLastError() == hipSuccess;
    }
    if (ok)
        ok = enqueue_batch(ctx, pairs, L.dPairs, L.dXfs, anySweeps ? L.dSweeps : nullptr, s->dCaches, s->dManifolds,
                           L.scratch, L.dSlots, nullptr, &L.streams) == B3AMD_OK;
    if (ok) ok = hipEventRecord(batch.done, stream) == hipSuccess;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!ok && s->workerError == B3AMD_OK) {
        s->workerError = B3AMD_ERR_DEVICE;
        s->workerMessage = b3amd_last_error();
    }
    s->stats.batches += 1;
    s->stats.requests += nreq;
    s->stats.pairs += pairs;
    if (pairs > s->stats.max_batch_pairs) s->stats.max_batch_pairs = pairs;
    s->stats.pack_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// Lock held.
void retire(b3amd_server* s, std::deque<Batch>::iterator it) {
    Batch& b = *it;
    const Clock::time_point now = Clock::now();
    for (Request* r : b.requests) {
        r->latencyMs = std::chrono::duration<double, std::milli>(now - r->submit).count();
        r->done = true;
        World& w = s->worlds[r->world];
        w.active -= 1;
        if (w.lastBatch == b.seq) w.lastBatch = 0;
    }
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, b.start, b.done) == hipSuccess) s->stats.gpu_ms += ms;
    s->lanes[b.lane].busy = false;
    s->inflight.erase(it);
}

void worker_main(b3amd_server* s) {
    hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cvWork.wait(lk, [&] { return s->stop || !s->pending.empty() || !s->inflight.empty(); });
        if (s->stop && s->pending.empty() && s->inflight.empty()) break;
        // Retire every finished batch, in any order (lanes complete independently).
        bool retired = false;
        for (auto it = s->inflight.begin(); it != s->inflight.end();) {
            if (hipEventQuery(it->done) == hipSuccess) {
                retire(s, it);
                it = s->inflight.begin();
                retired = true;
            } else {
                ++it;
            }
        }
        if (retired) s->cvDone.notify_all();
        if (!s->pending.empty() && s->inflight.size() < s->maxInflight && s->workerError == B3AMD_OK) {
            Batch b;
            while (s->lanes[b.lane].busy) ++b.lane;  // a free lane exists: inflight < maxInflight <= lanes
            b.seq = s->nextBatch++;
            b.start = s->lanes[b.lane].start;
            b.done = s->lanes[b.lane].done;
            if (select_batch(s, b) == 0) continue;
            s->lanes[b.lane].busy = true;
            s->inflight.push_back(b);
            Batch& queued = s->inflight.back();  // deque: stays valid while others are pushed or erased
            lk.unlock();
            launch_batch(s, queued);
            lk.lock();
            continue;
        }
        if (!s->inflight.empty()) {
            // Nothing to pack (or every lane busy): poll the lanes, waking early for new work.
            s->cvWork.wait_for(lk, std::chrono::microseconds(20));
            continue;
        }
        if (s->workerError != B3AMD_OK) {
            // Device failure: fail everything queued so waiters return.
            for (Request* r : s->pending) {
                r->done = true;
                s->worlds[r->world].active -= 1;
            }
            s->pending.clear();
            s->cvDone.notify_all();
        }
    }
}

int alloc_lane(b3amd_server* s, Lane& L) {
    b3amd_context* ctx = s->ctx;
    const uint64_t P = s->maxPairs, B = s->maxBodies;
    if (hipStreamCreateWithFlags(&L.streams.main, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&L.streams.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&L.streams.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&L.streams.join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&L.start, 0) != hipSuccess || hipEventCreateWithFlags(&L.done, 0) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "stream or event creation failed");
    int rc = B3AMD_OK;
    auto A = [&](uint64_t bytes, void** p) {
        if (rc == B3AMD_OK) rc = ctx->alloc(bytes, p);
    };
    A(48 * B, (void**)&L.dBodyXfs);
    A(32 * B, (void**)&L.dBodySweeps);
    A(sizeof(ReqDesc) * (uint64_t)s->maxRequests, (void**)&L.dDescs);
    A(8 * P, (void**)&L.dPairs);
    A(96 * P, (void**)&L.dXfs);
    A(64 * P, (void**)&L.dSweeps);
    A(4 * P, (void**)&L.dSlots);
    A(4 * scratch_words(P), (void**)&L.scratch.stage);
    A(sizeof(SatResult) * P, (void**)&L.scratch.sat);
    A(4 * kCountWords, (void**)&L.scratch.counter);
    if (rc != B3AMD_OK) return rc;
    if (hipHostMalloc((void**)&L.hBodyXfs, 48 * B, 0) != hipSuccess ||
        hipHostMalloc((void**)&L.hBodySweeps, 32 * B, 0) != hipSuccess ||
        hipHostMalloc((void**)&L.hDescs, sizeof(ReqDesc) * s->maxRequests, 0) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "hipHostMalloc of the staging buffers failed");
    return B3AMD_OK;
}

void free_lane(b3amd_server* s, Lane& L) {
    b3amd_context* ctx = s->ctx;
    if (L.streams.main) hipStreamSynchronize(L.streams.main);
    if (L.streams.side) hipStreamSynchronize(L.streams.side);
    void* bufs[] = {L.dBodyXfs, L.dBodySweeps, L.dDescs, L.dPairs, L.dXfs, L.dSweeps, L.dSlots,
                    L.scratch.stage, L.scratch.sat, L.scratch.counter};
    for (void* p : bufs) ctx->release(p);
    if (L.hBodyXfs) hipHostFree(L.hBodyXfs);
    if (L.hBodySweeps) hipHostFree(L.hBodySweeps);
    if (L.hDescs) hipHostFree(L.hDescs);
    if (L.streams.fork) hipEventDestroy(L.streams.fork);
    if (L.streams.join) hipEventDestroy(L.streams.join);
    if (L.start) hipEventDestroy(L.start);
    if (L.done) hipEventDestroy(L.done);
    if (L.streams.main) hipStreamDestroy(L.streams.main);
    if (L.streams.side) hipStreamDestroy(L.streams.side);
    L = Lane();
}

// Admits a world: waits (FIFO among admissions) until the slab has pair_count free slots
// or until `timeout` passes, then uploads the world's arrays and empty caches.
int admit_world(b3amd_server* s, uint32_t body_count, const uint32_t* body_shapes, uint32_t pair_count,
                const uint32_t* body_pairs, double timeout_ms, uint32_t* world) {
    if (!s || !world || !body_shapes || !body_pairs || !body_count || !pair_count)
        return fail(B3AMD_ERR_ARGUMENT, "null argument or empty world");
    b3amd_context* ctx = s->ctx;
    if (pair_count > s->maxPairs || body_count > s->maxBodies)
        return fail(B3AMD_ERR_ARGUMENT, "world of %u contacts / %u bodies exceeds the batch capacity %u / %u",
                    pair_count, body_count, s->maxPairs, s->maxBodies);
    if (pair_count > s->slabSlots)
        return fail(B3AMD_ERR_BUDGET, "world of %u contacts exceeds the slab of %u slots", pair_count, s->slabSlots);
    for (uint32_t b = 0; b < body_count; ++b)
        if (body_shapes[b] >= ctx->shapeCount)
            return fail(B3AMD_ERR_ARGUMENT, "body %u: shape %u of %u", b, body_shapes[b], ctx->shapeCount);
    for (uint32_t i = 0; i < pair_count; ++i) {
        const uint32_t a = body_pairs[2 * i], b = body_pairs[2 * i + 1];
        if (a >= body_count || b >= body_count)
            return fail(B3AMD_ERR_ARGUMENT, "contact %u refers to body %u/%u of %u", i, a, b, body_count);
        if (ctx->shapes[body_shapes[a]].type > ctx->shapes[body_shapes[b]].type)
            return fail(B3AMD_ERR_ARGUMENT, "contact %u: typeA > typeB (contact_manager.cpp:238-242)", i);
    }
    std::unique_lock<std::mutex> lk(s->mu);
    const uint64_t id = s->nextAdmit++;
    s->admitQueue.push_back(id);
    const auto deadline =
        Clock::now() + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::milli>(timeout_ms));
    const bool room = s->cvAdmit.wait_until(lk, deadline, [&] {
        return s->stop || (s->admitQueue.front() == id && fits(s, pair_count));
    }) && !s->stop;
    for (auto it = s->admitQueue.begin(); it != s->admitQueue.end(); ++it)
        if (*it == id) {
            s->admitQueue.erase(it);
            break;
        }
    s->cvAdmit.notify_all();  // the next waiter may be at the front now
    uint32_t base = 0;
    if (!room || !take_range(s, pair_count, &base))
        return fail(B3AMD_ERR_BUDGET, "no room for %u contact slots in the server's slab", pair_count);
    World w;
    int rc = ctx->alloc(4ull * body_count, (void**)&w.dBodyShapes);
    if (rc == B3AMD_OK) rc = ctx->alloc(8ull * pair_count, (void**)&w.dBodyPairs);
    if (rc != B3AMD_OK) {
        std::string msg = b3amd_last_error();
        ctx->release(w.dBodyShapes);
        give_range(s, base, pair_count);
        s->cvAdmit.notify_all();
        return fail(rc, "%s", msg.c_str());
    }
    // Uploads on the context stream, complete before any batch can use the world.
    HIP_TRY(hipMemcpyAsync(w.dBodyShapes, body_shapes, 4ull * body_count, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(w.dBodyPairs, body_pairs, 8ull * pair_count, hipMemcpyHostToDevice, ctx->stream));
    init_slots_kernel<<<(pair_count + 255) / 256, 256, 0, ctx->stream>>>(s->dCaches, s->dManifolds, base, pair_count);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    w.live = true;
    w.bodyCount = body_count;
    w.pairCount = pair_count;
    w.slotBase = base;
    uint32_t wid = 0;
    while (wid < s->worlds.size() && s->worlds[wid].live) ++wid;
    if (wid == s->worlds.size()) s->worlds.push_back(w);
    else s->worlds[wid] = w;
    *world = wid;
    return B3AMD_OK;
}

// Waits (lock held) until none of the world's requests is queued or in flight.
int quiesce_world(b3amd_server* s, std::unique_lock<std::mutex>& lk, uint32_t world) {
    if (world >= s->worlds.size() || !s->worlds[world].live) return fail(B3AMD_ERR_ARGUMENT, "world %u", world);
    s->cvDone.wait(lk, [&] { return s->worlds[world].active == 0; });
    return B3AMD_OK;
}

}  // namespace

extern "C" {

int b3amd_server_create(b3amd_context* ctx, uint32_t max_batch_pairs, uint32_t max_batch_bodies,
                        uint32_t slab_slots, b3amd_server** out) {
    if (!out) return fail(B3AMD_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (!ctx || !max_batch_pairs || !max_batch_bodies || !slab_slots)
        return fail(B3AMD_ERR_ARGUMENT, "null context or zero capacity");
    b3amd_server* s = new b3amd_server();
    s->ctx = ctx;
    s->maxPairs = max_batch_pairs;
    s->maxBodies = max_batch_bodies;
    s->maxRequests = max_batch_pairs < 65536u ? max_batch_pairs : 65536u;
    s->slabSlots = slab_slots;
    int rc = ctx->alloc(sizeof(ConvexCache) * (uint64_t)slab_slots, (void**)&s->dCaches);
    if (rc == B3AMD_OK) rc = ctx->alloc(sizeof(Manifold) * (uint64_t)slab_slots, (void**)&s->dManifolds);
    for (int i = 0; i < kMaxInflight && rc == B3AMD_OK; ++i) rc = alloc_lane(s, s->lanes[i]);
    if (rc != B3AMD_OK) {
        std::string msg = b3amd_last_error();
        b3amd_server_destroy(s);
        return fail(rc, "%s", msg.c_str());
    }
    s->freeRanges.push_back({0u, slab_slots});
    s->batchPairs = max_batch_pairs;
    s->worker = std::thread(worker_main, s);
    *out = s;
    return B3AMD_OK;
}

int b3amd_server_destroy(b3amd_server* s) {
    if (!s) return B3AMD_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cvAdmit.notify_all();
    if (s->worker.joinable()) {
        s->cvWork.notify_all();
        s->worker.join();
    }
    b3amd_context* ctx = s->ctx;
    for (Lane& L : s->lanes) free_lane(s, L);
    hipStreamSynchronize(ctx->stream);
    for (World& w : s->worlds) {
        if (!w.live) continue;
        ctx->release(w.dBodyShapes);
        ctx->release(w.dBodyPairs);
    }
    ctx->release(s->dCaches);
    ctx->release(s->dManifolds);
    for (auto& kv : s->requests) delete kv.second;
    delete s;
    return B3AMD_OK;
}

int b3amd_world_create(b3amd_server* s, uint32_t body_count, const uint32_t* body_shapes, uint32_t pair_count,
                       const uint32_t* body_pairs, uint32_t* world) {
    return admit_world(s, body_count, body_shapes, pair_count, body_pairs, 0.0, world);
}

int b3amd_world_admit(b3amd_server* s, uint32_t body_count, const uint32_t* body_shapes, uint32_t pair_count,
                      const uint32_t* body_pairs, double timeout_ms, uint32_t* world) {
    if (!(timeout_ms >= 0.0)) return fail(B3AMD_ERR_ARGUMENT, "negative or NaN timeout");
    return admit_world(s, body_count, body_shapes, pair_count, body_pairs, timeout_ms, world);
}

int b3amd_world_destroy(b3amd_server* s, uint32_t world) {
    if (!s) return fail(B3AMD_ERR_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (world >= s->worlds.size() || !s->worlds[world].live || s->worlds[world].closing)
        return fail(B3AMD_ERR_ARGUMENT, "world %u", world);
    s->worlds[world].closing = true;
    int rc = quiesce_world(s, lk, world);
    if (rc) return rc;
    World& w = s->worlds[world];
    s->ctx->release(w.dBodyShapes);
    s->ctx->release(w.dBodyPairs);
    give_range(s, w.slotBase, w.pairCount);
    w = World();
    s->cvAdmit.notify_all();
    return B3AMD_OK;
}

// Queues one step of a world.  copy: keep a private copy of the transforms (otherwise the
// caller keeps them alive until the request completes); at: arrival time (null: now).
static int submit(b3amd_server* s, uint32_t world, const float* xfs, const float* sweeps, bool copy,
                  const Clock::time_point* at, uint64_t* ticket) {
    if (!s || !xfs || !ticket) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    Request* r = new Request();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->workerError != B3AMD_OK) {
            delete r;
            return fail(s->workerError, "server failed: %s", s->workerMessage.c_str());
        }
        if (world >= s->worlds.size() || !s->worlds[world].live || s->worlds[world].closing) {
            delete r;
            return fail(B3AMD_ERR_ARGUMENT, "world %u", world);
        }
        World& w = s->worlds[world];
        if (copy) {
            r->ownXfs.assign(xfs, xfs + 12 * (size_t)w.bodyCount);
            r->xfs = r->ownXfs.data();
            if (sweeps) {
                r->ownSweeps.assign(sweeps, sweeps + 8 * (size_t)w.bodyCount);
                r->sweeps = r->ownSweeps.data();
            }
        } else {
            r->xfs = xfs;
            r->sweeps = sweeps;
        }
        r->ticket = s->nextTicket++;
        r->world = world;
        r->submit = at ? *at : Clock::now();
        w.active += 1;
        s->requests[r->ticket] = r;
        s->pending.push_back(r);
        *ticket = r->ticket;
    }
    s->cvWork.notify_one();
    return B3AMD_OK;
}

int b3amd_world_submit(b3amd_server* s, uint32_t world, const b3amd_transform* body_xfs,
                       const b3amd_quat* body_sweeps, uint64_t* ticket) {
    return submit(s, world, reinterpret_cast<const float*>(body_xfs), reinterpret_cast<const float*>(body_sweeps),
                  true, nullptr, ticket);
}

int b3amd_server_wait(b3amd_server* s, uint64_t ticket, double* latency_ms) {
    if (!s) return fail(B3AMD_ERR_ARGUMENT, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(ticket);
    if (it == s->requests.end()) return fail(B3AMD_ERR_ARGUMENT, "unknown or already collected ticket %llu",
                                             (unsigned long long)ticket);
    Request* r = it->second;
    s->cvDone.wait(lk, [&] { return r->done; });
    if (latency_ms) *latency_ms = r->latencyMs;
    s->requests.erase(ticket);
    delete r;
    if (s->workerError != B3AMD_OK) return fail(s->workerError, "server failed: %s", s->workerMessage.c_str());
    return B3AMD_OK;
}

// Copies `count` records of the world's slots after every step queued so far has run.
static int read_slots(b3amd_server* s, uint32_t world, void* out, bool caches) {
    if (!s || !out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    uint32_t base, count;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        int rc = quiesce_world(s, lk, world);
        if (rc) return rc;
        base = s->worlds[world].slotBase;
        count = s->worlds[world].pairCount;
    }
    if (caches)
        HIP_TRY(hipMemcpyAsync(out, s->dCaches + base, sizeof(ConvexCache) * (uint64_t)count, hipMemcpyDeviceToHost,
                               s->ctx->stream));
    else
        HIP_TRY(hipMemcpyAsync(out, s->dManifolds + base, sizeof(Manifold) * (uint64_t)count, hipMemcpyDeviceToHost,
                               s->ctx->stream));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream));
    return B3AMD_OK;
}

int b3amd_world_read_manifolds(b3amd_server* s, uint32_t world, b3amd_manifold* out) {
    return read_slots(s, world, out, false);
}

int b3amd_world_read_caches(b3amd_server* s, uint32_t world, b3amd_convex_cache* out) {
    return read_slots(s, world, out, true);
}

int b3amd_server_replay(b3amd_server* s, uint32_t count, const uint32_t* worlds, const double* times_s,
                        const b3amd_transform* const* body_xfs, double* latency_ms) {
    if (!s || !worlds || !times_s || !body_xfs || !latency_ms) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::vector<uint64_t> tickets(count);
    std::vector<Clock::time_point> due(count);
    {
        // Every request stays in the ticket map until the waits at the end: room for all of
        // them now, so that no insertion rehashes the map under the lock (milliseconds with
        // a few hundred thousand entries, while the worker cannot take the lock to launch).
        std::lock_guard<std::mutex> lk(s->mu);
        s->requests.reserve(s->requests.size() + count);
    }
    const Clock::time_point t0 = Clock::now();
    for (uint32_t i = 0; i < count; ++i) {
        due[i] = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(times_s[i]));
        while (Clock::now() < due[i]) {
        }
        // Latency counts from the scheduled arrival, so submission lag is included.
        int rc = submit(s, worlds[i], reinterpret_cast<const float*>(body_xfs[i]), nullptr, false, &due[i],
                        &tickets[i]);
        if (rc) {
            for (uint32_t k = 0; k < i; ++k) b3amd_server_wait(s, tickets[k], &latency_ms[k]);
            return rc;
        }
    }
    int rc = B3AMD_OK;
    for (uint32_t i = 0; i < count; ++i) {
        int r = b3amd_server_wait(s, tickets[i], &latency_ms[i]);
        if (r && rc == B3AMD_OK) rc = r;
    }
    return rc;
}

int b3amd_server_configure(b3amd_server* s, uint32_t batch_pairs, uint32_t max_inflight) {
    if (!s) return fail(B3AMD_ERR_ARGUMENT, "null server");
    if (max_inflight < 1 || max_inflight > (uint32_t)kMaxInflight)
        return fail(B3AMD_ERR_ARGUMENT, "max_inflight %u outside 1..%d", max_inflight, kMaxInflight);
    std::lock_guard<std::mutex> lk(s->mu);
    s->batchPairs = batch_pairs == 0 || batch_pairs > s->maxPairs ? s->maxPairs : batch_pairs;
    s->maxInflight = max_inflight;
    return B3AMD_OK;
}

int b3amd_server_get_stats(b3amd_server* s, b3amd_server_stats* out) {
    if (!s || !out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return B3AMD_OK;
}

}  // extern "C"
