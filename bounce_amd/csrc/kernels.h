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
    uint32_t pad[3];
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
    unsigned int pad;
};

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

    __device__ uint32_t slot_of(uint32_t p) const { return slots ? slots[p] : p; }

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

// Work split of the persistent SAT kernel: wave w of G takes pairs w, w + G, ... below
// staticEnd (an interleave gives every wave the same mix of hull sizes), then chunks of
// `chunk` pairs from a device counter until the batch is done.  Only the tail is handed
// out dynamically: device-scope atomics on one address serialise across the XCDs.
struct SatSplit {
    uint32_t* counter;       // pairs handed out past staticEnd; zeroed before each batch
    uint32_t staticEnd;
    uint32_t chunk;
};

struct CacheArgs : PairInputs {
    uint32_t* satCounter;    // the SAT kernel's tail counter and the overlap-list count, zeroed here
    const float* sweeps;     // [n][4][4] or null
    ConvexCache* caches;     // [n] or null
    uint32_t* stage;         // [n] out
    Manifold* out;           // [n]
    SatResult* sat;          // [n] out: stage kStageSkip for pairs that need no SAT
};

struct SatArgs : PairInputs {
    const uint32_t* stage;   // [n] from the cache kernel
    SatResult* out;          // [n]
    SatSplit split;
};

// Launch shape of the persistent SAT kernel (chosen by the autotuner).
struct SatTuning {
    uint32_t wavesPerSimd;   // persistent waves per SIMD and register budget, 0 = 5
    uint32_t maxSlots;       // edge-query register slots per lane: 1, 2 or 3
};

struct ContactArgs : PairInputs {
    SatResult* sat;          // [n]; the group kernel marks pairs it leaves with kStageFallback
    ConvexCache* caches;     // [n] or null
    Manifold* out;           // [n]
    uint32_t* list;          // [n] the overlapping pairs (reuses the cache kernel's stage array)
    uint32_t* listCount;     // zeroed by the cache kernel
};

// A pair whose clip polygon outgrows the group kernel's capacity: the lane-per-pair
// contact kernel builds it instead.
constexpr uint32_t kStageFallback = 6u;

struct ConvexArgs : PairInputs {
    Manifold* out;           // [n]
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
    uint32_t* totals;           // [4] work items per segment (zeroed before the midphase)
    uint32_t* pairContact;      // [3][n * cap] work item of a segment -> contact
    Manifold* temp;             // [n * cap] per-triangle manifolds, segments back to back
    Manifold* out;              // [n][3] clustered manifolds
    uint32_t* manifoldCounts;   // [n]
    // Persistent contacts (b3MeshContact across steps); null for freshly created ones.
    ConvexCache* triCaches;     // [n][cap] b3TriangleCache::cache of each listed triangle
    const float* sweeps;        // [n][4][4] orientation0/orientation of body A, then the mesh body
    float* fat;                 // [n][6] m_aabbA in the mesh frame (lower xyz, upper xyz)
    const float* centers0;      // [n][3] body A's sweep worldCenter0 (the displacement's origin)
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

hipError_t launch_convex(const ConvexArgs& a, hipStream_t s);
hipError_t launch_zero_words(uint32_t* words, uint32_t count, hipStream_t s);
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
