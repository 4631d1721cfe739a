// Internal host-side state of a narrowphase context, shared by the C ABI (api.cpp) and the
// serving loop (server.cpp).  Not part of the public interface.
#ifndef B3AMD_CONTEXT_H
#define B3AMD_CONTEXT_H

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "bounce_amd.h"
#include "device_hull.h"
#include "kernels.h"
#include "static_tree.h"

namespace b3amd {

// Records the message for b3amd_last_error (per thread) and returns `code`.
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess) return ::b3amd::fail(B3AMD_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

struct KernelStat {
    uint32_t launches = 0;
    double total_ms = 0.0;
};

struct PendingTiming {
    std::string name;
    hipEvent_t start, stop;
};

// Body-level inputs of a world step (b3amd_update_contacts): contact i joins bodies
// bodies[2i], bodies[2i + 1]; the kernels read the body transforms and sweeps directly.
struct BodyRefs {
    const uint32_t* bodies;  // [n][2]
    const float* xfs;        // [B][12]
    const float* sweeps;     // [B][8] or null
};

// hullOffsets entry of a freed registry slot.
constexpr uint32_t kFreeHull = 0xFFFFFFFFu;

// The streams a batch runs on: the hull kernels on `main`; sphere/capsule pairs forked to
// `side` after `fork` is recorded on main, joined back through `join`.  All at the default
// priority.
struct BatchStreams {
    hipStream_t main, side;
    hipEvent_t fork, join;
};

// Per-batch device scratch: scratch_words(capacity) words (stage, SAT list, rebuild lists),
// the SAT result per pair, and the counter words (kCountWords).
struct BatchScratch {
    uint32_t* stage;
    SatResult* sat;
    uint32_t* counter;
};

}  // namespace b3amd

using namespace b3amd;  // internal header: included by the library's .cpp files only

struct b3amd_context {
    int device = 0;
    hipStream_t stream = nullptr;
    // Sphere/capsule pairs run here beside the hull kernels: forked after the batch's inputs
    // are on `stream` and joined back before its end (captured graphs keep the fork).
    hipStream_t side = nullptr;
    hipEvent_t forkEvent = nullptr, joinEvent = nullptr;
    uint64_t budget = 0;
    uint64_t inUse = 0;
    std::unordered_map<void*, uint64_t> allocs;

    // Hull registry: one device arena of 16-byte units.  A record goes first-fit into a
    // freed range, else at the tail (the arena doubles when full).  Slot h's record starts at
    // hullOffsets[h]; a freed slot holds kFreeHull and is reused by the next add.
    float4* dPool = nullptr;
    uint64_t poolCap = 0, poolUsed = 0;  // units
    uint32_t* dOffsets = nullptr;
    uint32_t offsetsCap = 0;
    uint32_t hullCount = 0;              // slots, live or free
    std::vector<HullLayout> layouts;
    std::vector<uint32_t> hullOffsets;   // first 16-byte unit of each slot's record
    std::vector<uint32_t> hullUnits;     // record length of each slot, in units
    std::vector<uint32_t> freeSlots;
    std::vector<std::pair<uint64_t, uint64_t>> freeRanges;  // (first unit, units), sorted

    // Shape table.
    ShapeRec* dShapes = nullptr;
    uint32_t shapeCount = 0;
    bool roundShapes = false;  // the table holds spheres or capsules
    std::vector<ShapeRec> shapes;

    // Per-batch scratch and host-API staging.
    uint32_t* dStage = nullptr;
    SatResult* dSat = nullptr;
    uint32_t* dCounter = nullptr;  // tail counter of the persistent SAT kernel
    uint32_t scratchCap = 0;
    SatTuning tuning{0u, 0u};
    Counters* dCounters = nullptr;  // statistics totals (b3amd_get_counters)
    void* noAllocScratch = nullptr; // this device's copy of the no-allocation scratch array
    bool counting = false;
    void* dIn = nullptr;
    uint64_t dInCap = 0;
    // Status words of entry points that check device-side inputs (16 words, on demand).
    uint32_t* dStatus = nullptr;
    // Pinned host words for the counts that host code reads back between launches
    // (kHostWords words; a copy into pageable memory goes through a staging copy).
    static constexpr uint32_t kHostWords = 64;
    uint32_t* hWords = nullptr;
    // Shape pairs of the contacts of a world step (b3amd_update_contacts).
    void* dGather = nullptr;
    uint32_t gatherCap = 0;

    // Meshes (b3Mesh with its static tree) and mesh-contact scratch.
    struct MeshRecord {
        float4* verts = nullptr;
        uint4* tris = nullptr;
        TreeNode* nodes = nullptr;
        uint32_t vertexCount = 0, triangleCount = 0, nodeCount = 0;
    };
    std::vector<MeshRecord> meshes;
    void* dMeshScratch = nullptr;
    uint64_t meshScratchCap = 0;

    // Bumped whenever the hull pool or shape table is replaced: graphs captured before
    // hold the old device pointers and refuse to launch.
    uint64_t generation = 0;
    std::vector<b3amd_graph*> graphs;

    // Profiling.
    bool profiling = false;
    std::vector<hipEvent_t> freeEvents;
    std::vector<PendingTiming> pending;
    std::map<std::string, KernelStat> stats;

    int alloc(uint64_t bytes, void** p) {
        if (bytes == 0) bytes = 16;
        if (inUse + bytes > budget)
            return fail(B3AMD_ERR_BUDGET, "HBM budget exceeded: %llu in use + %llu requested > %llu",
                        (unsigned long long)inUse, (unsigned long long)bytes, (unsigned long long)budget);
        HIP_TRY(hipSetDevice(device));
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(B3AMD_ERR_BUDGET, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
        allocs[q] = bytes;
        inUse += bytes;
        *p = q;
        return B3AMD_OK;
    }
    int release(void* p) {
        if (!p) return B3AMD_OK;
        auto it = allocs.find(p);
        if (it == allocs.end()) return fail(B3AMD_ERR_ARGUMENT, "pointer %p not owned by this context", p);
        HIP_TRY(hipSetDevice(device));
        HIP_TRY(hipFree(p));
        inUse -= it->second;
        allocs.erase(it);
        return B3AMD_OK;
    }
    hipEvent_t event() {
        if (!freeEvents.empty()) {
            hipEvent_t e = freeEvents.back();
            freeEvents.pop_back();
            return e;
        }
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Copies `bytes` (at most 4 * kHostWords) of device memory to `h` after the work queued
    // on the stream so far, through the pinned words.
    int read_back(const void* d, void* h, size_t bytes) {
        if (bytes > 4ull * kHostWords) return fail(B3AMD_ERR_ARGUMENT, "read-back of %zu bytes", bytes);
        HIP_TRY(hipMemcpyAsync(hWords, d, bytes, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        std::memcpy(h, hWords, bytes);
        return B3AMD_OK;
    }
    int resolve_timings() {
        if (pending.empty()) return B3AMD_OK;
        HIP_TRY(hipStreamSynchronize(stream));
        for (PendingTiming& t : pending) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, t.start, t.stop));
            KernelStat& s = stats[t.name];
            s.launches += 1;
            s.total_ms += ms;
            freeEvents.push_back(t.start);
            freeEvents.push_back(t.stop);
        }
        pending.clear();
        return B3AMD_OK;
    }
};

struct b3amd_graph {
    b3amd_context* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t generation = 0;
    BatchScratch scratch{nullptr, nullptr, nullptr};
};


namespace b3amd {

// Brackets one kernel launch (launch() returns its hipError_t) with events while profiling
// is on; the timings go to b3amd_kernel_stats.
template <class F>
int timed(b3amd_context* ctx, hipStream_t s, const char* name, F launch) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->profiling) {
        a = ctx->event();
        b = ctx->event();
        if (!a || !b) return fail(B3AMD_ERR_DEVICE, "hipEventCreate failed");
        HIP_TRY(hipEventRecord(a, s));
    }
    hipError_t e = launch();
    if (e != hipSuccess) return fail(B3AMD_ERR_DEVICE, "launch of %s: %s", name, hipGetErrorString(e));
    if (ctx->profiling) {
        HIP_TRY(hipEventRecord(b, s));
        ctx->pending.push_back(PendingTiming{name, a, b});
    }
    return B3AMD_OK;
}

// Grows the context's per-batch scratch to n pairs.
int ensure_scratch(b3amd_context* ctx, uint32_t n);
// Enqueues the kernels of one batch on the context's streams, or on `streams` when given
// (no synchronisation); slots maps each pair to its cache/manifold slot (null: slot p).
int enqueue_batch(b3amd_context* ctx, uint32_t n, const uint32_t* pairs, const float* xfs, const float* sweeps,
                  ConvexCache* caches, Manifold* out, const BatchScratch& sc, const uint32_t* slots,
                  const BodyRefs* bodies = nullptr, const BatchStreams* streams = nullptr);

}  // namespace b3amd

#endif
