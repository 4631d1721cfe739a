// Device records and kernel launch interfaces shared by the kernels and the host API.
#ifndef B3AMD_KERNELS_H
#define B3AMD_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

#include "b3_math.h"
#include "device_hull.h"

namespace b3amd {

// b3amd_shape (include/bounce_amd.h).
struct ShapeRec {
    uint32_t type;
    uint32_t hull;
    float radius;
    float v[6];
    uint32_t hullUnit;  // first 16-byte unit of the hull's record in the pool (set by the library)
    uint32_t pad[2];
};
static_assert(sizeof(ShapeRec) == 48, "shape record layout");

// b3amd_manifold_point / b3amd_manifold.
struct ManifoldPoint {
    V3 localNormal1;
    V3 localPoint1;
    V3 localPoint2;
    uint32_t triangleKey;
    uint32_t key[4];
};
struct Manifold {
    uint32_t pointCount;
    uint32_t flags;
    uint32_t reserved[2];
    ManifoldPoint points[4];
};
static_assert(sizeof(ManifoldPoint) == 56, "manifold point layout");
static_assert(sizeof(Manifold) == 240, "manifold layout");

constexpr uint32_t kFlagPolyOverflow = 1u;

// The step after the narrowphase (b3amd_contact_init): the solver-owned fields of b3Manifold
// (manifold.h:66-93) and b3VelocityConstraintPoint / b3VelocityConstraintManifold
// (contact_solver.h:58-84) with the contact's mixed friction and restitution.
struct ManifoldImpulses {
    float normalImpulse[4];
    uint32_t persisting[4];
    float tangentImpulse[2];
    float motorImpulse;
    uint32_t reserved;
};
static_assert(sizeof(ManifoldImpulses) == 48, "impulse record layout");
struct VelocityPoint {
    float rA[3], rB[3], normal[3];
    float normalMass, normalImpulse, velocityBias;
};
struct VelocityManifold {
    VelocityPoint points[4];
    float rA[3], rB[3], normal[3], tangent1[3], tangent2[3];
    float tangentMass[4];
    float tangentImpulse[2];
    float motorImpulse, motorMass;
    uint32_t pointCount;
    float friction, restitution;
    uint32_t reserved[2];
};
static_assert(sizeof(VelocityManifold) == 304, "velocity constraint layout");
// A solver body: x (3), q (4), v (3), w (3), invMass, invI (9, columns), localCenter (3).
constexpr uint32_t kSolverBodyFloats = 26;
struct ContactInitArgs {
    uint32_t n, M;                     // contacts, manifold slots per contact
    const uint32_t* bodies;            // [n][2]
    const float* state;                // [B][kSolverBodyFloats]
    const float* material;             // [n][8] radiusA, radiusB, frictionA/B, restitutionA/B, pad
    const uint32_t* counts;            // [n] new manifolds
    const Manifold* manifolds;         // [n][M]
    const uint32_t* oldCounts;         // [n]
    const Manifold* oldManifolds;      // [n][M]
    const ManifoldImpulses* oldImpulses;  // [n][M]
    uint32_t warmStarting;
    ManifoldImpulses* impulses;        // [n][M] out
    VelocityManifold* constraints;     // [n][M] out
};
hipError_t launch_contact_init(const ContactInitArgs& a, hipStream_t s);

// b3SATFeaturePair (sat.h:74-80) inside b3ConvexCache (collide.h:38-42).
struct FeatureCache {
    uint32_t state, type, index1, index2;
};
struct ConvexCache {
    uint32_t simplex[5];  // b3SimplexCache, 20 bytes
    FeatureCache feature;
};
static_assert(sizeof(ConvexCache) == 36, "convex cache layout");

enum : uint32_t { kCacheSeparation = 0, kCacheOverlap = 1, kCacheEmpty = 2 };
enum : uint32_t { kFeatureEdge1 = 0, kFeatureFace1 = 1, kFeatureFace2 = 2 };

// Device equivalents of the reference's global statistics (collide_hulls.cpp:291-292,
// gjk.cpp:27): running totals over batches while counting is on.
struct Counters {
    unsigned long long convexCalls;      // b3_convexCalls: hull-hull collides
    unsigned long long convexCacheHits;  // b3_convexCacheHits: separation hits + rebuilt overlaps
    unsigned long long gjkCalls;         // b3_gjkCalls
    unsigned long long gjkIters;         // b3_gjkIters
    unsigned int gjkMaxIters;            // b3_gjkMaxIters
    unsigned int gjkCacheHits;           // b3_gjkCacheHits: simplex caches kept (sensor path)
};

// The totals are kept in kCounterSlots copies, block b adding into copy b % kCounterSlots, so
// that the per-wave atomics of a large batch do not all meet on one address; the host sums
// the copies (and takes the largest gjkMaxIters).
constexpr uint32_t kCounterSlots = 64;
__device__ __forceinline__ Counters* counter_slot(Counters* base) {
    return base + (blockIdx.x & (kCounterSlots - 1u));
}

struct PairInputs {
    uint32_t n;
    const uint32_t* pairs;   // [n][2] shape indices
    const ShapeRec* shapes;
    const float* xfs;        // [n][2][12]
    const float4* pool;      // hull records
    const uint32_t* offsets; // hull -> first 16-byte unit
    uint32_t shapeCount;
    uint32_t hullCount;
    const uint32_t* slots;   // [n] cache and manifold slot of each pair; null: slot p
    Counters* counters;      // statistics totals, or null when counting is off
    // Body-level inputs (a world step): when bodies is set, pair p joins bodies
    // bodies[2p], bodies[2p + 1] and its transforms (and sweeps) are read from the body
    // arrays instead of xfs (and the per-pair sweeps).
    const uint32_t* bodies = nullptr;    // [n][2] or null
    const float* bodyXfs = nullptr;      // [B][12]
    const float* bodySweeps = nullptr;   // [B][8] (orientation0, orientation) or null

    __device__ uint32_t slot_of(uint32_t p) const { return slots ? slots[p] : p; }
    // The 12 floats of pair p's transform of side 0 (shape A) or 1 (shape B).
    __device__ const float* xf_ptr(uint32_t p, uint32_t side) const {
        return bodies ? bodyXfs + 12 * (size_t)bodies[2 * p + side] : xfs + 24 * (size_t)p + 12 * side;
    }

    // 0: a pair kernels must not read (shape or hull index out of range, a type outside
    // sphere/capsule/hull, or typeA > typeB, which b3ContactManager::Create never makes,
    // contact_manager.cpp:238-242); 1: a hull-hull pair; 2: any other convex pair.
    __device__ uint32_t classify(uint32_t p) const {
        uint32_t a = pairs[2 * p], b = pairs[2 * p + 1];
        if (a >= shapeCount || b >= shapeCount) return 0u;
        const ShapeRec& A = shapes[a];
        const ShapeRec& B = shapes[b];
        if (A.type > 2u || B.type > 2u || A.type > B.type) return 0u;
        if ((A.type == 2u && A.hull >= hullCount) || (B.type == 2u && B.hull >= hullCount)) return 0u;
        return (A.type == 2u && B.type == 2u) ? 1u : 2u;
    }
};

constexpr uint32_t kFlagInvalidPair = 2u;
constexpr uint32_t kStageConvex = 5u;  // sphere/capsule pair: handled by the convex kernel

// Work split of the persistent SAT kernel over its list of L pairs: wave w of G takes list
// entries w, w + G, ... below staticEnd = L - max(L / 16, 2 chunk G) (an interleave gives
// every wave the same mix of hull sizes), then chunks of `chunk` entries (1 when L < chunk G)
// from a device counter until the list is done.  Only the tail is handed out dynamically:
// device-scope atomics on one address serialise across the XCDs.
struct SatSplit {
    uint32_t* counter;       // list entries handed out past the static part; zeroed before each batch
    uint32_t chunk;
};

// Per-batch scratch words: the stage word of each pair, the SAT list (pair indices) and the
// rebuild lists.  The rebuild lists are kRebuildLists sub-lists, one counter each, so that
// the per-wave appends of a large batch spread over many addresses: block b of the cache-read
// kernel (256 pairs) appends to sub-list b % kRebuildLists, whose capacity is
// rebuild_capacity(n).  The counter words (zeroed at the start of each batch): the SAT
// kernel's tail counter, the contact list's length, the SAT list's length, and from
// kCountRebuild the sub-lists' lengths.
constexpr uint32_t kRebuildLists = 64u;
constexpr uint32_t kReadPairsPerBlock = 256u;
__host__ __device__ inline uint32_t rebuild_capacity(uint32_t n) {
    const uint32_t blocks = (n + kReadPairsPerBlock - 1u) / kReadPairsPerBlock;
    return (blocks + kRebuildLists - 1u) / kRebuildLists * kReadPairsPerBlock;
}
// The wide passes' list lengths: pairs whose clip polygon outgrew the lane contact kernel's
// 32 vertices (kCountWide) or the lane cache-rebuild kernel's (kCountCacheWide).
enum : uint32_t {
    kCountSatTail = 0, kCountContacts = 1, kCountSat = 2, kCountWide = 3, kCountCacheWide = 4, kCountRebuild = 64,
    kCountWords = kCountRebuild + kRebuildLists
};
// Scratch words for a capacity of n pairs: stage, SAT list, rebuild lists.
inline uint64_t scratch_words(uint64_t n) {
    return 2u * n + (uint64_t)kRebuildLists * rebuild_capacity((uint32_t)n);
}

struct CacheArgs : PairInputs {
    uint32_t* counts;        // [kCountWords], zeroed before the cache kernels
    const float* sweeps;     // [n][4][4] or null (per-pair sweeps; body sweeps come with bodies)
    ConvexCache* caches;     // [n] or null
    uint32_t* stage;         // [n] out
    uint32_t* satList;       // [n] out: pairs that need the SAT
    uint32_t* rebuildList;   // [kRebuildLists][rebuild_capacity(n)] out: cached contacts to rebuild
    Manifold* out;           // [n]
    SatResult* sat;          // [n] out: stage kStageSkip for pairs that need no SAT

    // Pair p's sweep orientations (orientation0, orientation of body A, then of body B);
    // identities when no sweeps were given.
    __device__ void sweep(uint32_t p, Quat& q10, Quat& q1, Quat& q20, Quat& q2) const {
        const float4 one = make_float4(0.f, 0.f, 0.f, 1.f);
        float4 s0 = one, s1 = one, s2 = one, s3 = one;
        if (sweeps) {
            const float4* w = reinterpret_cast<const float4*>(sweeps + 16 * (size_t)p);
            s0 = w[0], s1 = w[1], s2 = w[2], s3 = w[3];
        } else if (bodies && bodySweeps) {
            const float4* wa = reinterpret_cast<const float4*>(bodySweeps + 8 * (size_t)bodies[2 * p]);
            const float4* wb = reinterpret_cast<const float4*>(bodySweeps + 8 * (size_t)bodies[2 * p + 1]);
            s0 = wa[0], s1 = wa[1], s2 = wb[0], s3 = wb[1];
        }
        q10 = Quat{s0.x, s0.y, s0.z, s0.w};
        q1 = Quat{s1.x, s1.y, s1.z, s1.w};
        q20 = Quat{s2.x, s2.y, s2.z, s2.w};
        q2 = Quat{s3.x, s3.y, s3.z, s3.w};
    }
};

// Stage words of the cache kernels: a hull pair with a cached feature to read, and one whose
// rebuilt clip polygon outgrows the group kernel (the lane-per-pair cache kernel takes it).
constexpr uint32_t kStageCached = 7u;
constexpr uint32_t kStageCacheFallback = 8u;
// A lane-form result whose clip polygon outgrew the lane's 32 vertices: nothing was written
// and the pair goes to a wide pass.
constexpr uint32_t kStageWide = 9u;

// What the cache and SAT kernels need of a hull pair, written into the pair's SatResult by
// the listing kernel before any query: the hull record units (f1Index, f2Index) and the
// radii (f1Sep, f2Sep) of shapes A and B, so later kernels read one record instead of the
// pair's shape indices and then their shape records; `stage` as given.
__device__ __forceinline__ float4 prefill_units(const PairInputs& a, uint32_t p) {
    const uint32_t sa = a.pairs[2 * p], sb = a.pairs[2 * p + 1];
    const ShapeRec& A = a.shapes[sa];
    const ShapeRec& B = a.shapes[sb];
    return make_float4(__uint_as_float(A.hullUnit), A.radius, __uint_as_float(B.hullUnit), B.radius);
}
__device__ __forceinline__ float4 sat_prefill(const PairInputs& a, uint32_t p, SatResult* r, uint32_t stage) {
    const float4 pre = prefill_units(a, p);
    *reinterpret_cast<float4*>(r) = pre;
    r->stage = stage;
    return pre;
}

struct SatArgs : PairInputs {
    const uint32_t* list;    // pairs to run, *listCount of them
    const uint32_t* listCount;
    SatResult* out;          // [n]
    SatSplit split;
};

// Launch shape of the persistent SAT kernel (chosen by the autotuner).
struct SatTuning {
    uint32_t wavesPerSimd;   // persistent waves per SIMD and register budget, 0 = 4
    uint32_t tailChunk;      // pairs per grab from the list's dynamic tail, 0 = 8
};

struct ContactArgs : PairInputs {
    SatResult* sat;          // [n]; the group kernel marks pairs it leaves with kStageFallback
    ConvexCache* caches;     // [n] or null
    Manifold* out;           // [n]
    uint32_t* list;          // [n] the overlapping pairs (reuses the cache kernel's stage array)
    uint32_t* listCount;     // zeroed by the cache kernel
    const uint32_t* satList;   // the pairs the SAT kernel ran, *satCount of them
    const uint32_t* satCount;
    uint32_t* wideList;        // [n] pairs for the wide contact pass (the SAT list's storage)
    uint32_t* wideCount;       // zeroed with the counter words
};

// A pair whose clip polygon outgrows the group kernel's capacity: the lane-per-pair
// contact kernel builds it instead.
constexpr uint32_t kStageFallback = 6u;

struct ConvexArgs : PairInputs {
    Manifold* out;           // [n]
};

// Sensor overlap tests (b3TestOverlap): overlap[p] = 1 overlapping, 0 apart, kOverlapInvalid for
// a pair the shape table cannot resolve; the simplex cache of caches[slot_of(p)] is read and
// written (null: fresh caches, not kept).
constexpr uint32_t kOverlapInvalid = 2u;
struct OverlapArgs : PairInputs {
    ConvexCache* caches;     // [slots] or null
    uint32_t* overlap;       // [n]
};

// A device-resident b3Mesh: vertices, triangles (v1, v2, v3, 0) and the static tree
// (8 words per node: lower xyz, upper xyz, child1, child2|triangle; static_tree.h:62-77).
struct MeshDev {
    const float4* verts;
    const uint4* tris;
    const uint32_t* nodes;
    uint32_t nodeCount;
};

// Triangles kept per mesh contact (the reference grows b3MeshContact::m_triangles without
// bound, mesh_contact.cpp:166-177); contacts overlapping more are flagged.
constexpr uint32_t kMeshTriangleCap = 64;
constexpr uint32_t kFlagTriangleOverflow = 4u;
constexpr uint32_t kFlagTreeStackOverflow = 8u;
// A fresh contact with more triangles than the list holds: the wide-contact pass queries
// again into a slab of kWideTris triangles and owns the contact's outputs (internal flag,
// cleared before any output).
constexpr uint32_t kFlagWideContact = 16u;
constexpr uint32_t kWideContactBlocks = 32;
constexpr uint32_t kWideTris = 2048;
constexpr uint32_t kWideObs = 4 * kWideTris;
constexpr uint64_t kWideSlabBytes = 16ull * 2 * kWideObs + 240ull * kWideTris + 4ull * kWideTris;

struct MeshArgs {
    Counters* counters;         // statistics totals, or null when counting is off
    uint32_t n;                 // mesh contacts
    const uint32_t* shapeA;     // [n] shape table index of the convex shape
    const ShapeRec* shapes;
    uint32_t shapeCount;
    uint32_t hullCount;
    const float* xfs;           // [n][2][12]: shape A body, mesh body
    const float4* pool;
    const uint32_t* offsets;
    MeshDev mesh;
    uint32_t cap;               // triangles per contact (<= kMeshTriangleCap)
    uint32_t* triCounts;        // [n] triangles the midphase found (may exceed cap)
    uint32_t* triLists;         // [n][cap] in DFS order
    uint32_t* contactFlags;     // [n]
    // Work items are grouped by the convex shape's collider so that a wave runs one of
    // them: segment 0 hulls, 1 capsules, 2 spheres, in that order in `temp`.
    uint32_t* contactOffset;    // [n] first work item of the contact within its segment
    uint32_t* totals;           // [8]: work items per segment, the clustering's count of
                                // contacts left to its second pass, the wide items' count
                                // (zeroed before the midphase)
    uint32_t* wideItems;        // [n * cap] work items for the wide pass (clip polygons beyond
                                // the lane's 32 vertices)
    uint32_t* wideContacts = nullptr;  // [n] fresh contacts beyond cap triangles (count in totals[5]),
                                       // null for persistent contacts
    void* wideSlabs = nullptr;         // [kWideContactBlocks] slabs of kWideSlabBytes
    uint2* pairContact;         // [3][n * cap] work items of a segment: (contact, triangle); once
                                // the narrowphase has run, the clustering's second-pass list
                                // (uint32 [n])
    Manifold* temp;             // [n * cap] per-triangle manifolds, segments back to back
    Manifold* out;              // [n][3] clustered manifolds
    uint32_t* manifoldCounts;   // [n]
    // Persistent contacts (b3MeshContact across steps); null for freshly created ones.
    ConvexCache* triCaches;     // [n][cap] b3TriangleCache::cache of each listed triangle
    const float* sweeps;        // [n][4][4] orientation0/orientation of body A, then the mesh body
    float* fat;                 // [n][6] m_aabbA in the mesh frame (lower xyz, upper xyz)
    const float* centers0;      // [n][3] body A's sweep worldCenter0 (the displacement's origin)
    uint32_t* overlap = nullptr;  // [n] sensor test result (b3MeshContact::TestOverlap)
};

// Mesh-contact state update: kMeshCreate is the b3MeshContact constructor followed by
// FindNewPairs (mesh_contact.cpp:30-55, 141-160); kMeshSync is SynchronizeShapes, MoveAABB
// and FindNewPairs (:62-160), which query the tree again only when the shape left its fat
// AABB and then reset the triangle caches.
enum : uint32_t { kMeshCreate = 0, kMeshSync = 1 };
hipError_t launch_mesh_sync(const MeshArgs& a, uint32_t mode, hipStream_t s);
// Work items of the listed triangles (as the fresh midphase builds them).
hipError_t launch_mesh_list(const MeshArgs& a, hipStream_t s);

hipError_t launch_mesh_midphase(const MeshArgs& a, hipStream_t s);
// The grid covers the n * cap bound: the work-item count stays on the device.
hipError_t launch_mesh_narrow(const MeshArgs& a, hipStream_t s);
hipError_t launch_mesh_cluster(const MeshArgs& a, hipStream_t s);
// The fresh contacts the midphase listed in wideContacts (more triangles than cap).
hipError_t launch_mesh_contact_wide(const MeshArgs& a, hipStream_t s);
// b3MeshContact::TestOverlap of every persistent contact (one lane per contact).
hipError_t launch_mesh_overlap(const MeshArgs& a, hipStream_t s);

// The shape pairs of one world step's contacts: contact i joins bodies contacts[2i],
// contacts[2i+1] (the kernels read their transforms and sweeps through PairInputs::bodies).
// A body index outside the world gives an invalid shape pair, which no kernel reads further.
struct BodyGatherArgs {
    uint32_t n;
    uint32_t bodyCount;
    const uint32_t* bodyShapes;  // [bodyCount]
    const uint32_t* contacts;    // [n][2]
    uint32_t* pairs;             // [n][2] out: shape indices
};
hipError_t launch_body_gather(const BodyGatherArgs& a, hipStream_t s);

hipError_t launch_convex(const ConvexArgs& a, hipStream_t s);
hipError_t launch_overlap(const OverlapArgs& a, hipStream_t s);
// Caches [0, n) to the empty state of a new contact (a fill kernel, safe under capture).
hipError_t launch_cache_reset(ConvexCache* caches, uint32_t n, hipStream_t s);
hipError_t launch_zero_words(uint32_t* words, uint32_t count, hipStream_t s);
// Zeroes the counter words, marks the cached pairs and lists the SAT-bound ones, reads the
// feature caches four lanes per cached pair (listing the contacts to rebuild), rebuilds them
// eight lanes per pair, and resolves the pairs whose rebuild outgrows the group on one lane
// each.
hipError_t launch_cache_hull(const CacheArgs& a, hipStream_t s);
// Persistent SAT over the pairs the cache kernel marked; any hull size within the record
// limits.  Fills a.split (the counter must be zero).
hipError_t launch_sat_persistent(SatArgs a, const SatTuning& t, hipStream_t s);
// Contact building after the SAT: a lane-per-pair pass writes the separated pairs and lists
// the overlapping ones; eight lanes per listed pair build the contact with the clip polygon
// spread across them; the lane-per-pair kernel takes the pairs marked kStageFallback.
hipError_t launch_contact_hull(const ContactArgs& a, hipStream_t s);

}  // namespace b3amd

#endif
