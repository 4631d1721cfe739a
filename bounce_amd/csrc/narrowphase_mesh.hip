// Shape-vs-triangle-mesh narrowphase for gfx950 (b3MeshContact, mesh_contact.cpp).
//
//   mesh_midphase_kernel  one lane per mesh contact: the fat AABB of shape A in the mesh
//                         frame (mesh_contact.cpp:30-60) and the static-tree query in DFS
//                         order, child2 first (static_tree.h:100-142); triangle lists are
//                         compacted into (contact, triangle) work items, grouped by the
//                         convex shape's type so that each wave runs one collider.
//   mesh_narrow_kernel    one lane per (contact, triangle) work item: b3TriangleHull built
//                         in registers and b3CollideShapeAndShape (with the triangle's cache
//                         for persistent contacts), the triangle index stamped into every
//                         point key (:238-266).
//   mesh_cluster_kernel   sixteen lanes per mesh contact: b3ClusterSolver::Run over the
//                         triangle manifolds in midphase order (contact_cluster.cpp:202-590),
//                         observations in LDS, every one of a contact's up to 256 points.
#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "contact_build.h"
#include "convex_build.h"
#include "device_common.h"
#include "device_hull.h"
#include "kernels.h"

namespace b3amd {
namespace {

struct Aabb {
    V3 lo, hi;
};

__device__ __forceinline__ V3 vmin(V3 a, V3 b) {
    return V3{a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z};
}
__device__ __forceinline__ V3 vmax(V3 a, V3 b) {
    return V3{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z};
}

// b3Shape::ComputeAABB: sphere_shape.cpp:53-59, capsule_shape.cpp:130-137,
// hull_shape.cpp:162-166 (b3AABB3::Set over transformed vertices, then Extend(radius)).
__device__ Aabb shape_aabb(const ShapeRec& s, const Xf& xf, const float4* pool, const uint32_t* offsets) {
    Aabb a;
    if (s.type == 0u) {
        V3 c = mul(xf, V3{s.v[0], s.v[1], s.v[2]});
        V3 r{s.radius, s.radius, s.radius};
        a.lo = c - r;
        a.hi = c + r;
    } else if (s.type == 1u) {
        V3 c1 = mul(xf, V3{s.v[0], s.v[1], s.v[2]});
        V3 c2 = mul(xf, V3{s.v[3], s.v[4], s.v[5]});
        V3 r{s.radius, s.radius, s.radius};
        a.lo = vmin(c1, c2) - r;
        a.hi = vmax(c1, c2) + r;
    } else {
        const DevHull h = DevHull::at_unit(pool, s.hullUnit);
        a.lo = a.hi = mul(xf, h.vertex(0));
        for (uint32_t i = 1; i < h.V; ++i) {
            V3 v = mul(xf, h.vertex(i));
            a.lo = vmin(a.lo, v);
            a.hi = vmax(a.hi, v);
        }
        V3 r{s.radius, s.radius, s.radius};
        a.lo = a.lo - r;
        a.hi = a.hi + r;
    }
    return a;
}

// b3TestOverlap (aabb3.h:247-251) against one packed tree node.
__device__ __forceinline__ bool overlaps(const float4& lo, const float4& hi, const Aabb& b) {
    return (lo.x <= b.hi.x) && (lo.y <= b.hi.y) && (lo.z <= b.hi.z) && (lo.w >= b.lo.x) && (hi.x >= b.lo.y) &&
           (hi.y >= b.lo.z);
}

constexpr uint32_t kTreeStack = 128;

// Work-item segment of a convex shape type: hulls, capsules, spheres.
__device__ __forceinline__ uint32_t segment_of(uint32_t type) { return 2u - type; }

// Whether contact c's shape A is a convex shape the kernels may read.
__device__ __forceinline__ bool mesh_pair_valid(const MeshArgs& a, uint32_t si) {
    return si < a.shapeCount && a.shapes[si].type <= 2u && (a.shapes[si].type != 2u || a.shapes[si].hull < a.hullCount);
}

// b3StaticTree::QueryAABB (static_tree.h:100-142) reporting into contact c's triangle list
// (b3MeshContact::Report, mesh_contact.cpp:162-189): DFS, child2 first, every overlapping
// leaf in order.  Returns the number of triangles found (the list keeps the first cap).
// The DFS stack: its first kLdsStack entries in LDS (entry k of the lane at lds[256 k],
// consecutive lanes in consecutive banks), deeper ones (only an unbalanced tree reaches
// them) in private memory.
constexpr uint32_t kLdsStack = 32;
constexpr uint32_t kMidphaseBlock = 256;

__device__ uint32_t tree_query(const MeshArgs& a, uint32_t* list, uint32_t cap, const Aabb& box, uint32_t& flags,
                               uint32_t* lds, uint32_t ldsStride = kMidphaseBlock) {
    uint32_t found = 0;
    uint32_t deep[kTreeStack - kLdsStack];
    uint32_t top = 0;
    auto push = [&](uint32_t v) {
        if (top < kLdsStack) lds[ldsStride * top] = v;
        else deep[top - kLdsStack] = v;
        ++top;
    };
    if (a.mesh.nodeCount > 0) push(0u);
    while (top) {
        --top;
        const uint32_t ni = top < kLdsStack ? lds[ldsStride * top] : deep[top - kLdsStack];
        const float4* nd = reinterpret_cast<const float4*>(a.mesh.nodes + 8 * (size_t)ni);
        const float4 q0 = nd[0], q1 = nd[1];  // lower xyz, upper x | upper yz, child1, child2
        if (!overlaps(q0, q1, box)) continue;
        const uint32_t child1 = __float_as_uint(q1.z), child2 = __float_as_uint(q1.w);
        if (child1 == 0xFFFFFFFFu) {
            if (found < cap) list[found] = child2;
            ++found;
        } else if (top + 2 <= kTreeStack) {
            push(child1);
            push(child2);
        } else {
            flags |= kFlagTreeStackOverflow;
        }
    }
    if (found > cap) flags |= kFlagTriangleOverflow;
    return found;
}

// The fat AABB of contact c's shape A in the mesh frame (mesh_contact.cpp:30-60): the shape's
// AABB in the frame, extended by B3_AABB_EXTENSION (settings.h:54).
__device__ __forceinline__ Aabb fat_box(const MeshArgs& a, uint32_t c, const ShapeRec& s) {
    const Xf xfA = load_xf(a.xfs + 24 * (size_t)c);
    const Xf xfB = load_xf(a.xfs + 24 * (size_t)c + 12);
    Aabb box = shape_aabb(s, mulT(xfB, xfA), a.pool, a.offsets);
    const V3 ext{0.2f, 0.2f, 0.2f};
    box.lo = box.lo - ext;
    box.hi = box.hi + ext;
    return box;
}

// Appends contact c's m triangles as work items (contact, triangle) of its collider's
// segment and returns the first item's offset in the segment (0 without items).  Every lane
// of the wave must call it (one atomic per segment with a wave-uniform address, so each folds
// into a single wave-level add; a per-lane segment address would serialise every add).  The
// triangle indices come from the contact's list, which this lane has just written.
__device__ __forceinline__ uint32_t list_work(const MeshArgs& a, uint32_t c, uint32_t seg, uint32_t m) {
    uint32_t off = 0;
    for (uint32_t g = 0; g < 3u; ++g) {
        const uint32_t o = atomicAdd(a.totals + g, seg == g ? m : 0u);
        if (seg == g) off = o;
    }
    uint2* const items = a.pairContact + (size_t)(seg < 3u ? seg : 0u) * a.n * a.cap + off;
    const uint32_t* const list = a.triLists + (size_t)c * a.cap;
    for (uint32_t k = 0; k < m; ++k) items[k] = make_uint2(c, list[k]);
    return off;
}

// Freshly created contacts: the fat AABB of shape A in the mesh frame (mesh_contact.cpp:30-60),
// the tree query and the work items.
__global__ __launch_bounds__(kMidphaseBlock) void mesh_midphase_kernel(MeshArgs a) {
    __shared__ uint32_t stackLds[kLdsStack * kMidphaseBlock];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    uint32_t flags = 0;
    const uint32_t si = a.shapeA[c];
    uint32_t found = 0;
    if (mesh_pair_valid(a, si)) {
        found = tree_query(a, a.triLists + (size_t)c * a.cap, a.cap, fat_box(a, c, a.shapes[si]), flags,
                           stackLds + threadIdx.x);
    } else {
        flags |= kFlagInvalidPair;
    }
    uint32_t m = (flags & kFlagInvalidPair) ? 0u : (found < a.cap ? found : a.cap);
    if (a.wideContacts && (flags & kFlagTriangleOverflow) && !(flags & kFlagTreeStackOverflow)) {
        // More triangles than the list holds: mesh_contact_wide_kernel takes the contact whole.
        flags |= kFlagWideContact;
        m = 0;
        a.wideContacts[atomicAdd(a.totals + 5, 1u)] = c;
    }
    a.triCounts[c] = found;
    a.contactFlags[c] = flags;
    a.contactOffset[c] = list_work(a, c, m > 0 ? segment_of(a.shapes[si].type) : 3u, m);
}

// Persistent contacts: kMeshCreate (constructor + FindNewPairs) or kMeshSync
// (SynchronizeShapes + MoveAABB + FindNewPairs, mesh_contact.cpp:62-160).  A query resets
// the caches of the listed triangles to empty (b3MeshContact::Report).
__global__ __launch_bounds__(kMidphaseBlock) void mesh_sync_kernel(MeshArgs a, uint32_t mode) {
    __shared__ uint32_t stackLds[kLdsStack * kMidphaseBlock];
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const uint32_t si = a.shapeA[c];
    float* const fat = a.fat + 6 * (size_t)c;
    ConvexCache* const caches = a.triCaches + (size_t)c * a.cap;
    uint32_t* const countOut = a.triCounts + c;
    uint32_t* const flagsOut = a.contactFlags + c;
    const bool valid = mesh_pair_valid(a, si);
    bool query = valid;
    Aabb box{V3{0.0f, 0.0f, 0.0f}, V3{0.0f, 0.0f, 0.0f}};
    if (valid) {
        const ShapeRec s = a.shapes[si];
        const Xf xfA = load_xf(a.xfs + 24 * (size_t)c);
        const Xf xfB = load_xf(a.xfs + 24 * (size_t)c + 12);
        const Aabb aabb = shape_aabb(s, mulT(xfB, xfA), a.pool, a.offsets);
        const float ext = 0.2f;  // B3_AABB_EXTENSION (settings.h:54)
        box = Aabb{aabb.lo - V3{ext, ext, ext}, aabb.hi + V3{ext, ext, ext}};
        if (mode == kMeshSync) {
            const Aabb old{V3{fat[0], fat[1], fat[2]}, V3{fat[3], fat[4], fat[5]}};
            // b3AABB3::Contains (aabb3.h:156-167): both corners of the new box in the fat one.
            const bool inside = old.lo.x <= aabb.lo.x && aabb.lo.x <= old.hi.x && old.lo.y <= aabb.lo.y &&
                                aabb.lo.y <= old.hi.y && old.lo.z <= aabb.lo.z && aabb.lo.z <= old.hi.z &&
                                old.lo.x <= aabb.hi.x && aabb.hi.x <= old.hi.x && old.lo.y <= aabb.hi.y &&
                                aabb.hi.y <= old.hi.y && old.lo.z <= aabb.hi.z && aabb.hi.z <= old.hi.z;
            query = !inside;  // m_aabbMoved: a contained box keeps the list and its caches
            // Motion-predicted fat box (B3_AABB_MULTIPLIER, settings.h:59).
            const float* c0 = a.centers0 + 3 * (size_t)c;
            const V3 d = xfA.p - V3{c0[0], c0[1], c0[2]};
            const float kMul = 2.0f;
            if (d.x < 0.0f) box.lo.x += kMul * d.x;
            else box.hi.x += kMul * d.x;
            if (d.y < 0.0f) box.lo.y += kMul * d.y;
            else box.hi.y += kMul * d.y;
            if (d.z < 0.0f) box.lo.z += kMul * d.z;
            else box.hi.z += kMul * d.z;
        }
    }
    uint32_t flags = valid ? 0u : kFlagInvalidPair, found = 0;
    if (query) {
        fat[0] = box.lo.x, fat[1] = box.lo.y, fat[2] = box.lo.z;
        fat[3] = box.hi.x, fat[4] = box.hi.y, fat[5] = box.hi.z;
        found = tree_query(a, a.triLists + (size_t)c * a.cap, a.cap, box, flags, stackLds + threadIdx.x);
        const uint32_t m = found < a.cap ? found : a.cap;
        for (uint32_t k = 0; k < m; ++k) {
            ConvexCache e;
            for (int w = 0; w < 5; ++w) e.simplex[w] = 0u;  // b3SimplexCache with count 0
            e.feature.state = kCacheEmpty;
            e.feature.type = e.feature.index1 = e.feature.index2 = 0u;
            caches[k] = e;
        }
    }
    if (query || !valid) {
        *countOut = found;
        *flagsOut = flags;
    }
}

// Work items of persistent contacts from their current triangle lists.
__global__ __launch_bounds__(256) void mesh_list_kernel(MeshArgs a) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    const uint32_t flags = a.contactFlags[c], found = a.triCounts[c];
    const uint32_t m = (flags & kFlagInvalidPair) ? 0u : (found < a.cap ? found : a.cap);
    a.contactOffset[c] = list_work(a, c, m > 0 ? segment_of(a.shapes[a.shapeA[c]].type) : 3u, m);
}

// Statistics of the per-triangle collides a lane ran.
struct MeshTally {
    uint32_t hullCalls = 0, hits = 0, calls = 0, iters = 0, maxIters = 0;
};

// b3CollideShapeAndShape of contact c's convex shape s and triangle ti into m (mesh_contact.cpp:
// 238-266), hull-triangle face contacts built in the lane's workspace or the wide pass's; fc
// is the triangle's cache (persistent contacts) or null, q its four sweep orientations.
// Returns false, with nothing written and nothing counted, when a clip polygon outgrew the
// lane's workspace (the triangle's cache is as it was: collide_hulls_cached_seq stops before
// writing it).
template <bool Wide>
__device__ __forceinline__ bool collide_tri(const MeshArgs& a, uint32_t c, const ShapeRec& s, uint32_t ti,
                                            FeatureCache* fc, const Quat* q, WideFaceWs* wide, TriHull& tri,
                                            Contact& m, MeshTally& tally) {
    const uint4 t = a.mesh.tris[ti];
    const float4 A = a.mesh.verts[t.x], B = a.mesh.verts[t.y], C = a.mesh.verts[t.z];
    tri = TriHull::make(V3{A.x, A.y, A.z}, V3{B.x, B.y, B.z}, V3{C.x, C.y, C.z});
    const Xf xfA = load_xf(a.xfs + 24 * (size_t)c);
    const Xf xfB = load_xf(a.xfs + 24 * (size_t)c + 12);
    m.pointCount = 0;
    m.overflow = false;
    uint32_t hit = 0;
    if (s.type == 2u) {
        const DevHull h = DevHull::at_unit(a.pool, s.hullUnit);
        FaceWs lane;
        auto& ws = *select_ws<Wide>(&lane, wide);
        if (fc) collide_hulls_cached_seq(m, xfA, h, s.radius, xfB, tri, 0.0f, fc, q[0], q[1], q[2], q[3], hit, ws);
        else collide_cache_seq(m, xfA, h, s.radius, xfB, tri, 0.0f, nullptr, ws);
        if (m.overflow && !Wide) return false;
        tally.hullCalls += 1;
        tally.hits += hit;
    } else if (s.type == 0u) {
        collide_sphere_hull(m, xfA, s, xfB, tri, 0.0f);
    } else {
        collide_capsule_hull(m, xfA, s, xfB, tri, 0.0f);
    }
    tally.calls += m.gjkCalls;
    tally.iters += m.gjkIters;
    tally.maxIters = m.gjkIters > tally.maxIters ? m.gjkIters : tally.maxIters;
    return true;
}

// The manifold of a triangle collide with the triangle index in every point key.
__device__ __forceinline__ void write_tri_manifold(Manifold* out, const Contact& m, uint32_t ti) {
    write_manifold(out, m);
    for (uint32_t j = 0; j < m.pointCount; ++j) out->points[j].triangleKey = ti;
}

// Work item f (of t0 + t1 + t2, segments back to back) into temp[f]; false as collide_tri.
template <bool Wide>
__device__ __forceinline__ bool mesh_item(const MeshArgs& a, uint32_t f, uint32_t t0, uint32_t t1, WideFaceWs* wide,
                                          TriHull& tri, MeshTally& tally) {
    const uint32_t seg = f < t0 ? 0u : (f < t0 + t1 ? 1u : 2u);
    const uint32_t local = f - (seg == 0u ? 0u : (seg == 1u ? t0 : t0 + t1));
    const uint2 item = a.pairContact[(size_t)seg * a.n * a.cap + local];
    const uint32_t c = item.x, ti = item.y;
    const ShapeRec s = a.shapes[a.shapeA[c]];
    FeatureCache* fc = nullptr;
    Quat q[4];
    if (s.type == 2u && a.triCaches) {
        // b3MeshContact::Collide with the triangle's cache (mesh_contact.cpp:238-257); the
        // triangle hull belongs to the mesh body, whose sweep the rebuild reads.
        const uint32_t k = local - a.contactOffset[c];
        fc = &a.triCaches[(size_t)c * a.cap + k].feature;
        for (int w = 0; w < 4; ++w) {
            if (a.sweeps) {
                const float* qs = a.sweeps + 16 * (size_t)c + 4 * w;
                q[w] = Quat{qs[0], qs[1], qs[2], qs[3]};
            } else {
                q[w] = Quat{0.0f, 0.0f, 0.0f, 1.0f};
            }
        }
    }
    Contact m;
    if (!collide_tri<Wide>(a, c, s, ti, fc, q, wide, tri, m, tally)) return false;
    write_tri_manifold(a.temp + f, m, ti);
    return true;
}

// With counting on, hull-triangle collides go to b3_convexCalls and the sphere/capsule
// GJK calls to the b3_gjk* totals, as b3MeshContact::Collide's per-triangle
// b3CollideShapeAndShape calls count them in the reference (mesh_contact.cpp:217-275).
// Items whose clip polygon outgrows the lane's 32 vertices are listed (wideItems, count in
// totals[4]) for mesh_wide_kernel.
__global__ __launch_bounds__(256) void mesh_narrow_kernel(MeshArgs a) {
    // Each lane's triangle hull in LDS: the collide routines take it by reference.
    __shared__ TriHull triLds[256];
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t0 = a.totals[0], t1 = a.totals[1], t2 = a.totals[2];
    MeshTally tally;
    if (f < t0 + t1 + t2) {
        if (!mesh_item<false>(a, f, t0, t1, nullptr, triLds[threadIdx.x], tally))
            a.wideItems[atomicAdd(a.totals + 4, 1u)] = f;
    }
    if (a.counters) {  // all lanes of the block reach here
        Counters* const ct = counter_slot(a.counters);
        wave_count(&ct->convexCalls, tally.hullCalls);
        wave_count(&ct->convexCacheHits, tally.hits);
        wave_count(&ct->gjkCalls, tally.calls);
        wave_count(&ct->gjkIters, tally.iters);
        wave_max(&ct->gjkMaxIters, tally.maxIters);
    }
}

// The items mesh_narrow_kernel listed, one per block, its first lane colliding with the
// polygons in LDS (room for kWidePolyCap vertices).
constexpr uint32_t kMeshWideBlocks = 64;
__global__ __launch_bounds__(64) void mesh_wide_kernel(MeshArgs a) {
    __shared__ ClipV polyBuf[kWidePolyCap], scratchBuf[kWidePolyCap];
    __shared__ CPV keptBuf[kWidePolyCap];
    __shared__ TriHull tri;
    MeshTally tally;
    if (threadIdx.x == 0u) {
        const uint32_t t0 = a.totals[0], t1 = a.totals[1], count = a.totals[4];
        for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
            WideFaceWs ws{PolyView{0u, false, kWidePolyCap, polyBuf}, PolyView{0u, false, kWidePolyCap, scratchBuf},
                          CPolyView{0u, keptBuf}};
            mesh_item<true>(a, a.wideItems[e], t0, t1, &ws, tri, tally);  // beyond kWidePolyCap: flagged in the manifold
        }
    }
    if (a.counters) {  // all lanes of the block reach here (lanes other than 0 add nothing)
        Counters* const ct = counter_slot(a.counters);
        wave_count(&ct->convexCalls, tally.hullCalls);
        wave_count(&ct->convexCacheHits, tally.hits);
        wave_count(&ct->gjkCalls, tally.calls);
        wave_count(&ct->gjkIters, tally.iters);
        wave_max(&ct->gjkMaxIters, tally.maxIters);
    }
}

// ------------------------------------------------------------------ clustering

constexpr uint32_t kMaxObs = 4 * kMeshTriangleCap;

struct WPoint {
    V3 point, normal;
    float separation;
};

// b3WorldManifoldPoint::Initialize (manifold.cpp:57-69).
__device__ WPoint world_point(const ManifoldPoint& p, float rA, const Xf& xfA, float rB, const Xf& xfB) {
    WPoint w;
    V3 nA = mul(xfA.R, p.localNormal1);
    V3 cA = mul(xfA, p.localPoint1);
    V3 cB = mul(xfB, p.localPoint2);
    V3 pA = cA + rA * nA;
    V3 pB = cB - rB * nA;
    w.point = 0.5f * (pA + pB);
    w.normal = nA;
    w.separation = dot(cB - cA, nA) - rA - rB;
    return w;
}

__device__ __forceinline__ float dist_sq(V3 a, V3 b) {
    V3 d = a - b;
    return d.x * d.x + d.y * d.y + d.z * d.z;  // b3DistanceSquared (vec3.h:207-211)
}

// The clusters of b3ClusterSolver (contact_cluster.cpp:336-377), held by every lane of a
// contact's group with the same values.  InitializeClusters adds at most three.  The loops run
// over all three entries with a bound test, so that the centroids stay in registers.
struct Clusters {
    V3 c[3];
    uint32_t n;
    __device__ uint32_t find(V3 p) const {  // FindCluster (:361-377)
        uint32_t best = 0;
        float bestValue = kMaxFloat;
#pragma unroll
        for (uint32_t i = 0; i < 3u; ++i) {
            if (i >= n) break;
            float m = dist_sq(p, c[i]);
            if (m < bestValue) {
                bestValue = m;
                best = i;
            }
        }
        return best;
    }
    __device__ void add(V3 centroid) {  // AddCluster (:336-359)
        if (n == 0) {
            c[0] = centroid;
            n = 1;
            return;
        }
        const uint32_t b = find(centroid);
        const float kTol = 0.05f;
#pragma unroll
        for (uint32_t i = 0; i < 3u; ++i) {
            if (i != b) continue;
            if (dist_sq(centroid, c[i]) <= kTol * kTol) {
                c[i] = c[i] + centroid;
                normalize_in_place(c[i]);
                return;
            }
        }
#pragma unroll
        for (uint32_t i = 0; i < 3u; ++i)
            if (i == n) c[i] = centroid;
        ++n;
    }
};

// b3ClusterSolver::Run (contact_cluster.cpp:475-590) for one mesh contact by a group of
// kClusterGroup lanes.  The observations sit in LDS in manifold then point order, each as
// its b3WorldManifoldPoint (manifold.cpp:57-69: world normal, world point, separation,
// evaluated once) and its manifold point; observation o is lane o % 16's.  Everything that
// compares or selects runs across the lanes: the cluster assignment, the first maximum or
// minimum of a selection loop (larger value wins, equal values the smaller observation
// index, which is the sequential loop's strict comparison in index order), the member
// lists.  Every floating-point sum keeps the reference's order: a cluster's centroid sum is
// one lane's walk over the observations in index order, as the reference adds them.  The
// polygon of a cluster is its members in observation order, so a member's polygon index is
// never needed: a selection over the polygon is the same selection over the members.
//
// The contact's observations get Cap LDS slots.  The pass with Cap = kClusterLds (four
// contacts in 8 KB per wave) leaves a contact with more observations to the second pass
// (Cap = kMaxObs, every observation the triangle cap allows), which takes the listed
// contacts from scratch.
constexpr uint32_t kClusterGroup = 16;
constexpr uint32_t kClusterBlock = 64;  // one wave, four contacts
constexpr uint32_t kClusterLds = 64;
constexpr uint32_t kClusterWideBlocks = 256;  // grid of the second pass (1,024 contacts at a time)
constexpr uint32_t kNoIndex = 0xFFFFFFFFu;
static_assert(kClusterLds % 8u == 0u && kMaxObs % 8u == 0u, "the centroid sums read eight observations at a time");

// LDS written by one lane and read by another lane of the same wave: the fences order the
// accesses (no code motion across, outstanding LDS operations completed).
__device__ __forceinline__ void group_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// The group's bits of a wave ballot (bit g: lane g of the group).
__device__ __forceinline__ uint32_t group_ballot(bool p, uint32_t gbase) {
    return (uint32_t)(__ballot(p) >> gbase) & 0xFFFFu;
}

// First maximum across the group: each lane brings its own first maximum (v, o), o = kNoIndex
// when it has none (then v = -FLT_MAX, which no candidate of a strict `v > max` loop from
// -FLT_MAX can equal).
__device__ __forceinline__ void group_first_max(float& v, uint32_t& o) {
#pragma unroll
    for (int off = 1; off < (int)kClusterGroup; off <<= 1) {
        const float ov = __shfl_xor(v, off, (int)kClusterGroup);
        const uint32_t oo = (uint32_t)__shfl_xor((int)o, off, (int)kClusterGroup);
        if (ov > v || (ov == v && oo < o)) {
            v = ov;
            o = oo;
        }
    }
}

// First minimum across the group (v = FLT_MAX, o = kNoIndex for none).
__device__ __forceinline__ void group_first_min(float& v, uint32_t& o) {
#pragma unroll
    for (int off = 1; off < (int)kClusterGroup; off <<= 1) {
        const float ov = __shfl_xor(v, off, (int)kClusterGroup);
        const uint32_t oo = (uint32_t)__shfl_xor((int)o, off, (int)kClusterGroup);
        if (ov < v || (ov == v && oo < o)) {
            v = ov;
            o = oo;
        }
    }
}

__device__ __forceinline__ uint32_t group_min_u32(uint32_t x) {
#pragma unroll
    for (int off = 1; off < (int)kClusterGroup; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_xor((int)x, off, (int)kClusterGroup);
        x = y < x ? y : x;
    }
    return x;
}

__device__ __forceinline__ V3 xyz(float4 v) { return V3{v.x, v.y, v.z}; }
// An observation's fourth word: its manifold point (i << 2) | j in the low 16 bits, its
// cluster above.  Member lists pack up to four observation indices in 16-bit fields.
__device__ __forceinline__ uint32_t obs_point(float4 v) { return __float_as_uint(v.w) & 0xFFFFu; }
__device__ __forceinline__ uint32_t obs_cluster(float4 v) { return __float_as_uint(v.w) >> 16; }
__device__ __forceinline__ uint32_t field_at(uint64_t packed, uint32_t i) {
    return (uint32_t)(packed >> (16u * i)) & 0xFFFFu;
}

// b3ClosestPointOnSegment (geometry.h:174-193) and the squared distance to it, as
// InitializeClusters' third pick measures it (contact_cluster.cpp:303-330).
__device__ __forceinline__ float segment_dist_sq(V3 A, V3 B, V3 Pt) {
    float w[3];
    barycentric(w, A, B, Pt);
    V3 Q;
    if (w[1] <= 0.0f) Q = A;
    else if (w[0] <= 0.0f) Q = B;
    else {
        float sInv = 1.0f / w[2];
        float wA = sInv * w[0], wB = sInv * w[1];
        Q = wA * A + wB * B;
    }
    return dist_sq(Pt, Q);
}

// b3SortPolygon (contact_cluster.cpp:33-56) on up to four members (observation indices packed
// in the 16-bit fields of q), positions in ob; the pivot B is read once per i, before any swap.
__device__ __forceinline__ uint64_t sort_members(uint64_t q, uint32_t count, const float4* ob, V3 N) {
    if (count == 0u) return q;
    const V3 A = xyz(ob[field_at(q, 0)]);
    for (uint32_t i = 1; i < count; ++i) {
        const V3 B = xyz(ob[field_at(q, i)]);
        for (uint32_t j = i + 1; j < count; ++j) {
            const V3 C = xyz(ob[field_at(q, j)]);
            if (!is_ccw(A, B, C, N)) {
                const uint64_t bi = field_at(q, i), bj = field_at(q, j);
                q = (q & ~(0xFFFFull << (16u * i)) & ~(0xFFFFull << (16u * j))) | (bj << (16u * i)) |
                    (bi << (16u * j));
            }
        }
    }
    return q;
}

// Clusters contact ci's m triangle manifolds `in` (midphase order) with Cap observation slots
// (obN: world normal and the packed manifold point and cluster; obP: world point, later
// moved along the cluster normal, and the separation), the outputs carrying `flags`.  With
// more than Cap observations the group lists the contact in `defer` and writes nothing;
// returns whether it did.
template <uint32_t Cap>
__device__ __forceinline__ bool cluster_manifolds(const MeshArgs& a, uint32_t ci, const Manifold* in, uint32_t m,
                                                  uint32_t flags, float4* obN, float4* obP, uint32_t g,
                                                  uint32_t gbase, uint32_t* defer, uint32_t* deferCount) {
    Manifold* const out = a.out + 3 * (size_t)ci;
    uint32_t numOut = 0;
    uint64_t qs0 = 0, qs1 = 0, qs2 = 0;  // each output's manifold points (16-bit fields)
    uint32_t qn0 = 0, qn1 = 0, qn2 = 0;  // and their counts
    if (!(flags & kFlagInvalidPair) && m > 0) {
        const ShapeRec s = a.shapes[a.shapeA[ci]];
        const Xf xfA = load_xf(a.xfs + 24 * (size_t)ci);
        const Xf xfB = load_xf(a.xfs + 24 * (size_t)ci + 12);
        const float rA = s.radius, rB = 0.0f;

        // Observations (:480-495), manifold by manifold.  Lane g takes manifolds g, g + 16, ...
        // and loads all four point records with the count (one round trip); an exclusive scan
        // of the counts places each manifold's points.
        uint32_t nObs = 0;
        for (uint32_t r = 0; r * kClusterGroup < m; ++r) {
            const uint32_t i = r * kClusterGroup + g;
            uint32_t pc = 0;
            ManifoldPoint pts[4];
            if (i < m) {
                pc = in[i].pointCount;
#pragma unroll
                for (uint32_t j = 0; j < 4u; ++j) pts[j] = in[i].points[j];
            }
            pc = pc < 4u ? pc : 4u;
            uint32_t incl = pc;
#pragma unroll
            for (uint32_t d = 1; d < kClusterGroup; d <<= 1) {
                const uint32_t t = (uint32_t)__shfl_up((int)incl, d, (int)kClusterGroup);
                if (g >= d) incl += t;
            }
            const uint32_t total = (uint32_t)__shfl((int)incl, (int)kClusterGroup - 1, (int)kClusterGroup);
            if (nObs + total > Cap) {  // only with Cap < kMaxObs: nObs <= 4 m <= kMaxObs
                if (g == 0u && defer) defer[atomicAdd(deferCount, 1u)] = ci;
                return true;
            }
            const uint32_t off = nObs + incl - pc;
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j)
                if (j < pc) {
                    const WPoint w = world_point(pts[j], rA, xfA, rB, xfB);
                    obN[off + j] = make_float4(w.normal.x, w.normal.y, w.normal.z, __uint_as_float((i << 2) | j));
                    obP[off + j] = make_float4(w.point.x, w.point.y, w.point.z, w.separation);
                }
            nObs += total;
        }
        group_lds_sync();

        // InitializeClusters (:212-334).
        Clusters C;
        C.n = 0;
        if (nObs <= 3u) {
            for (uint32_t o = 0; o < nObs; ++o) C.add(xyz(obN[o]));
        } else {
            C.add(xyz(obN[0]));
            uint32_t c1;
            {
                const V3 A = C.c[0];
                float v = -kMaxFloat;
                uint32_t idx = kNoIndex;
                for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                    if (o == 0u) continue;
                    const float dd = dist_sq(A, xyz(obN[o]));
                    if (dd > v) {
                        v = dd;
                        idx = o;
                    }
                }
                group_first_max(v, idx);
                c1 = idx == kNoIndex ? 0u : idx;
                C.add(xyz(obN[c1]));
            }
            float v = -kMaxFloat;
            uint32_t idx = kNoIndex;
            if (C.n == 1u) {
                const V3 A = C.c[0];
                for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                    if (o == 0u || o == c1) continue;
                    const float dd = dist_sq(A, xyz(obN[o]));
                    if (dd > v) {
                        v = dd;
                        idx = o;
                    }
                }
            } else {
                const V3 A = C.c[0], B = C.c[1];
                for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                    if (o == 0u || o == c1) continue;
                    const float dd = segment_dist_sq(A, B, xyz(obN[o]));
                    if (dd > v) {
                        v = dd;
                        idx = o;
                    }
                }
            }
            group_first_max(v, idx);
            C.add(xyz(obN[idx == kNoIndex ? 0u : idx]));
        }

        // Ten Lloyd iterations (:404-444).  The assignment of observation o is written into
        // its LDS word; lane k < n then sums cluster k's observations in index order (eight
        // LDS reads in flight, the adds in order).  The sums and counts of the last
        // iteration are the polygon normals' (:512-530 add the same observations in the same
        // order from zero).  An iteration that assigns every observation as the one before
        // is a fixed point: the same members give the same centroids (an empty cluster keeps
        // its centroid either way), so every later iteration repeats it, and the loop stops
        // there with the state the tenth iteration would leave.
        V3 sum{0.0f, 0.0f, 0.0f};
        uint32_t cnt = 0;
        for (uint32_t iter = 0; iter < 10u; ++iter) {
            bool changed = iter == 0u;
            for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                const float4 v = obN[o];
                const uint32_t c = C.find(xyz(v));
                changed = changed || c != obs_cluster(v);
                reinterpret_cast<uint32_t*>(obN + o)[3] = obs_point(v) | (c << 16);
            }
            if (group_ballot(changed, gbase) == 0u) break;
            group_lds_sync();
            sum = V3{0.0f, 0.0f, 0.0f};
            cnt = 0;
            if (g < C.n) {
                for (uint32_t o = 0; o < nObs; o += 8u) {  // o + 7 < Cap: nObs <= Cap, o % 8 == 0
                    float4 v[8];
#pragma unroll
                    for (uint32_t i = 0; i < 8u; ++i) v[i] = obN[o + i];
#pragma unroll
                    for (uint32_t i = 0; i < 8u; ++i)
                        if (o + i < nObs && obs_cluster(v[i]) == g) {
                            sum = sum + xyz(v[i]);
                            ++cnt;
                        }
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < 3u; ++k) {
                const float sx = __shfl(sum.x, (int)k, (int)kClusterGroup);
                const float sy = __shfl(sum.y, (int)k, (int)kClusterGroup);
                const float sz = __shfl(sum.z, (int)k, (int)kClusterGroup);
                const uint32_t ck = (uint32_t)__shfl((int)cnt, (int)k, (int)kClusterGroup);
                if (k < C.n && ck > 0u) {
                    const float sInv = 1.0f / (float)ck;  // b3Vec3::operator/= (vec3.h:70-76)
                    C.c[k] = V3{sx * sInv, sy * sInv, sz * sInv};
                }
            }
            group_lds_sync();
        }
        group_lds_sync();

        // Per non-empty cluster, in cluster order (the renumbering of :449-472 keeps the
        // order): the members' world points moved along the normal, the deepest member and
        // b3ReducePolygon (:58-200).
        for (uint32_t k = 0; k < C.n && numOut < 3u; ++k) {
            const uint32_t ck = (uint32_t)__shfl((int)cnt, (int)k, (int)kClusterGroup);
            V3 normal{__shfl(sum.x, (int)k, (int)kClusterGroup), __shfl(sum.y, (int)k, (int)kClusterGroup),
                      __shfl(sum.z, (int)k, (int)kClusterGroup)};
            if (ck == 0u) continue;
            normalize_in_place(normal);
            float sepMin = kMaxFloat;
            uint32_t minO = kNoIndex, first = kNoIndex;
            for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                if (obs_cluster(obN[o]) != k) continue;
                const float4 w = obP[o];
                if (w.w < sepMin) {
                    sepMin = w.w;
                    minO = o;
                }
                first = first < o ? first : o;
                const V3 pos = xyz(w) + w.w * normal;
                obP[o] = make_float4(pos.x, pos.y, pos.z, w.w);
            }
            group_first_min(sepMin, minO);
            first = group_min_u32(first);
            group_lds_sync();
            // b3ReducePolygon: with four points or fewer all of them, sorted; else the
            // deepest, then the farthest from it, the largest triangle, the most negative
            // signed area, sorted.  Members are tested by their cluster; chosen members are
            // skipped.  A selection with no candidate keeps index 0: the first member.
            uint64_t q = 0;
            uint32_t qn = 0;
            if (ck <= 4u) {
                for (uint32_t r = 0; r * kClusterGroup < nObs && qn < ck; ++r) {
                    const uint32_t o = r * kClusterGroup + g;
                    uint32_t bits = group_ballot(o < nObs && obs_cluster(obN[o]) == k, gbase);
                    while (bits != 0u && qn < 4u) {
                        q |= (uint64_t)(r * kClusterGroup + (uint32_t)__builtin_ctz(bits)) << (16u * qn);
                        ++qn;
                        bits &= bits - 1u;
                    }
                }
                q = sort_members(q, qn, obP, normal);
            } else {
                const uint32_t q0 = minO == kNoIndex ? first : minO;
                q = q0;
                qn = 1;
                do {
                    const V3 A = xyz(obP[q0]);
                    float v = -kMaxFloat;
                    uint32_t idx = kNoIndex;
                    for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                        if (obs_cluster(obN[o]) != k || o == q0) continue;
                        const V3 d = xyz(obP[o]) - A;
                        const float dd = dot(d, d);
                        if (dd > v) {
                            v = dd;
                            idx = o;
                        }
                    }
                    group_first_max(v, idx);
                    if (v < kEpsilon * kEpsilon) break;
                    const uint32_t q1 = idx;
                    q |= (uint64_t)q1 << 16;
                    qn = 2;
                    const V3 B = xyz(obP[q1]);
                    v = -kMaxFloat;
                    idx = kNoIndex;
                    for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                        if (obs_cluster(obN[o]) != k || o == q0 || o == q1) continue;
                        const V3 n = cross(B - A, xyz(obP[o]) - A);
                        const float a2 = fabsf(dot(n, normal));
                        if (a2 > v) {
                            v = a2;
                            idx = o;
                        }
                    }
                    group_first_max(v, idx);
                    const float kAreaTol = 0.01f;
                    if (v < 2.0f * kAreaTol) break;
                    const uint32_t q2 = idx;
                    if (!is_ccw(A, B, xyz(obP[q2]), normal)) q = (uint64_t)q1 | ((uint64_t)q0 << 16);
                    q |= (uint64_t)q2 << 32;
                    qn = 3;
                    const V3 A2 = xyz(obP[field_at(q, 0)]), B2 = xyz(obP[field_at(q, 1)]);
                    v = kMaxFloat;
                    idx = kNoIndex;
                    for (uint32_t o = g; o < nObs; o += kClusterGroup) {
                        if (obs_cluster(obN[o]) != k || o == q0 || o == q1 || o == q2) continue;
                        const V3 n = cross(B2 - A2, xyz(obP[o]) - A2);
                        const float sa2 = dot(n, normal);
                        if (sa2 < v) {
                            v = sa2;
                            idx = o;
                        }
                    }
                    group_first_min(v, idx);
                    if (fabsf(v) < 2.0f * kAreaTol) break;
                    q |= (uint64_t)(idx == kNoIndex ? first : idx) << 48;
                    qn = 4;
                    q = sort_members(q, 4u, obP, normal);
                } while (false);
            }
            // The members' manifold points (i << 2) | j.
            uint64_t pts = 0;
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j)
                if (j < qn) pts |= (uint64_t)obs_point(obN[field_at(q, j)]) << (16u * j);
            if (numOut == 0u) qs0 = pts, qn0 = qn;
            else if (numOut == 1u) qs1 = pts, qn1 = qn;
            else qs2 = pts, qn2 = qn;
            ++numOut;
            group_lds_sync();
        }
    }
    // The output manifolds (:587-589): lane 4 t + j copies point j of output t (its original
    // manifold point, or zeros), lane 12 + t writes output t's header; outputs past numOut
    // are empty manifolds with the contact's flags, 15 16-byte words each.
    static_assert(sizeof(Manifold) == 15 * 16, "manifold of 15 16-byte words");
    if (g < 12u) {
        const uint32_t t = g >> 2, j = g & 3u;
        if (t < numOut) {
            const uint64_t qsT = t == 0u ? qs0 : (t == 1u ? qs1 : qs2);
            const uint32_t qnT = t == 0u ? qn0 : (t == 1u ? qn1 : qn2);
            ManifoldPoint pt;
            if (j < qnT) {
                const uint32_t pi = field_at(qsT, j);
                pt = in[pi >> 2].points[pi & 3u];
            } else {
                pt.localNormal1 = V3{0.0f, 0.0f, 0.0f};
                pt.localPoint1 = V3{0.0f, 0.0f, 0.0f};
                pt.localPoint2 = V3{0.0f, 0.0f, 0.0f};
                pt.triangleKey = 0;
                pt.key[0] = pt.key[1] = pt.key[2] = pt.key[3] = 0;
            }
            out[t].points[j] = pt;
        }
    } else if (g < 15u) {
        const uint32_t t = g - 12u;
        if (t < numOut) {
            const uint32_t qnT = t == 0u ? qn0 : (t == 1u ? qn1 : qn2);
            *reinterpret_cast<uint4*>(out + t) = make_uint4(qnT, flags, 0u, 0u);
        }
    }
    for (uint32_t t = numOut; t < 3u; ++t)
        if (g < 15u)
            reinterpret_cast<uint4*>(out + t)[g] = g == 0u ? make_uint4(0u, flags, 0u, 0u) : make_uint4(0u, 0u, 0u, 0u);
    if (g == 0u) a.manifoldCounts[ci] = numOut;
    return false;
}

// Contact ci from the narrowphase's work items (its manifolds sit at its segment's offset in
// temp); a contact the wide-contact pass owns is left to it.
template <uint32_t Cap>
__device__ __forceinline__ void cluster_contact(const MeshArgs& a, uint32_t ci, float4* obN, float4* obP, uint32_t g,
                                                uint32_t gbase, uint32_t* defer, uint32_t* deferCount) {
    const uint32_t flags = a.contactFlags[ci];
    if (flags & kFlagWideContact) return;
    const uint32_t found = a.triCounts[ci];
    const uint32_t m = found < a.cap ? found : a.cap;
    const Manifold* in = nullptr;
    if (!(flags & kFlagInvalidPair) && m > 0) {
        const uint32_t seg = segment_of(a.shapes[a.shapeA[ci]].type);
        const uint32_t base = seg == 0u ? 0u : (seg == 1u ? a.totals[0] : a.totals[0] + a.totals[1]);
        in = a.temp + base + a.contactOffset[ci];
    }
    cluster_manifolds<Cap>(a, ci, in, m, flags, obN, obP, g, gbase, defer, deferCount);
}

// Every contact, four per wave, kClusterLds observation slots each; contacts with more
// observations are listed (pairContact, free once the narrowphase has run; count in
// totals[3]) for mesh_cluster_wide_kernel.
__global__ __launch_bounds__(kClusterBlock) void mesh_cluster_kernel(MeshArgs a) {
    __shared__ float4 obsN[kClusterBlock / kClusterGroup][kClusterLds];
    __shared__ float4 obsP[kClusterBlock / kClusterGroup][kClusterLds];
    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane & (kClusterGroup - 1u);
    const uint32_t gbase = lane & ~(kClusterGroup - 1u);
    const uint32_t ci = blockIdx.x * (kClusterBlock / kClusterGroup) + lane / kClusterGroup;
    if (ci >= a.n) return;  // whole groups leave; nothing below waits on other groups
    cluster_contact<kClusterLds>(a, ci, obsN[lane / kClusterGroup], obsP[lane / kClusterGroup], g, gbase,
                                 reinterpret_cast<uint32_t*>(a.pairContact), a.totals + 3);
}

// The listed contacts with every observation (kMaxObs slots each): group q of the grid takes
// entries q, q + groups, ...
__global__ __launch_bounds__(kClusterBlock) void mesh_cluster_wide_kernel(MeshArgs a) {
    __shared__ float4 obsN[kClusterBlock / kClusterGroup][kMaxObs];
    __shared__ float4 obsP[kClusterBlock / kClusterGroup][kMaxObs];
    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane & (kClusterGroup - 1u);
    const uint32_t gbase = lane & ~(kClusterGroup - 1u);
    const uint32_t groups = gridDim.x * (kClusterBlock / kClusterGroup);
    const uint32_t count = a.totals[3];
    for (uint32_t e = blockIdx.x * (kClusterBlock / kClusterGroup) + lane / kClusterGroup; e < count; e += groups) {
        cluster_contact<kMaxObs>(a, reinterpret_cast<const uint32_t*>(a.pairContact)[e], obsN[lane / kClusterGroup], obsP[lane / kClusterGroup], g,
                                 gbase, nullptr, nullptr);
        group_lds_sync();
    }
}

// Contacts whose tree query found more triangles than the lists hold (the reference's
// b3MeshContact grows m_triangles without bound, mesh_contact.cpp:166-189), one per block
// from the midphase's list: the query again into the block's slab (up to kWideTris
// triangles), the per-triangle collides across the block's lanes (a clip polygon beyond the
// lane's 32 vertices is redone by the first lane with the polygons in LDS), then the
// clustering by the first 16 lanes with the observations in the slab.
__global__ __launch_bounds__(64) void mesh_contact_wide_kernel(MeshArgs a) {
    __shared__ ClipV polyBuf[kWidePolyCap], scratchBuf[kWidePolyCap];
    __shared__ CPV keptBuf[kWidePolyCap];
    __shared__ TriHull triLds[64];
    __shared__ uint32_t stack[kLdsStack];
    __shared__ uint32_t redo[64];
    __shared__ uint32_t sh[2];  // triangles found, items to redo
    const uint32_t lane = threadIdx.x;
    char* const slab = static_cast<char*>(a.wideSlabs) + kWideSlabBytes * blockIdx.x;
    float4* const obN = reinterpret_cast<float4*>(slab);
    float4* const obP = obN + kWideObs;
    Manifold* const mans = reinterpret_cast<Manifold*>(obP + kWideObs);
    uint32_t* const list = reinterpret_cast<uint32_t*>(mans + kWideTris);
    MeshTally tally;
    const uint32_t count = a.totals[5];
    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const uint32_t c = a.wideContacts[e];
        const ShapeRec s = a.shapes[a.shapeA[c]];
        if (lane == 0u) {
            uint32_t flags = 0;
            sh[0] = tree_query(a, list, kWideTris, fat_box(a, c, s), flags, stack, 1u);
            sh[1] = 0u;
        }
        __syncthreads();
        const uint32_t found = sh[0];
        const uint32_t mt = found < kWideTris ? found : kWideTris;
        for (uint32_t k0 = 0; k0 < mt; k0 += 64u) {
            const uint32_t k = k0 + lane;
            if (k < mt) {
                Contact m;
                if (collide_tri<false>(a, c, s, list[k], nullptr, nullptr, nullptr, triLds[lane], m, tally))
                    write_tri_manifold(mans + k, m, list[k]);
                else
                    redo[atomicAdd(&sh[1], 1u)] = k;
            }
            __syncthreads();
            if (lane == 0u) {
                for (uint32_t r = 0; r < sh[1]; ++r) {
                    WideFaceWs ws{PolyView{0u, false, kWidePolyCap, polyBuf},
                                  PolyView{0u, false, kWidePolyCap, scratchBuf}, CPolyView{0u, keptBuf}};
                    Contact m;
                    const uint32_t kr = redo[r];
                    collide_tri<true>(a, c, s, list[kr], nullptr, nullptr, &ws, triLds[0], m, tally);
                    write_tri_manifold(mans + kr, m, list[kr]);
                }
                sh[1] = 0u;
            }
            __syncthreads();
        }
        if (lane < kClusterGroup) {
            // Every triangle within the slab: the contact's result is the reference's, unflagged.
            const uint32_t keep = found <= kWideTris ? ~(kFlagWideContact | kFlagTriangleOverflow) : ~kFlagWideContact;
            cluster_manifolds<kWideObs>(a, c, mans, mt, a.contactFlags[c] & keep, obN, obP, lane, 0u, nullptr,
                                        nullptr);
        }
        __syncthreads();
    }
    if (a.counters) {  // all lanes of the block reach here
        Counters* const ct = counter_slot(a.counters);
        wave_count(&ct->convexCalls, tally.hullCalls);
        wave_count(&ct->convexCacheHits, tally.hits);
        wave_count(&ct->gjkCalls, tally.calls);
        wave_count(&ct->gjkIters, tally.iters);
        wave_max(&ct->gjkMaxIters, tally.maxIters);
    }
}

}  // namespace

hipError_t launch_mesh_midphase(const MeshArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    mesh_midphase_kernel<<<(a.n + kMidphaseBlock - 1) / kMidphaseBlock, kMidphaseBlock, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_mesh_sync(const MeshArgs& a, uint32_t mode, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    mesh_sync_kernel<<<(a.n + kMidphaseBlock - 1) / kMidphaseBlock, kMidphaseBlock, 0, s>>>(a, mode);
    return hipGetLastError();
}

hipError_t launch_mesh_list(const MeshArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    mesh_list_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_mesh_narrow(const MeshArgs& a, hipStream_t s) {
    const uint64_t bound = (uint64_t)a.n * a.cap;
    if (bound == 0) return hipSuccess;
    mesh_narrow_kernel<<<(uint32_t)((bound + 255) / 256), 256, 0, s>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    mesh_wide_kernel<<<kMeshWideBlocks, 64, 0, s>>>(a);
    return hipGetLastError();
}

// b3MeshContact::TestOverlap (mesh_contact.cpp:191-215), the sensor path of a mesh contact: one
// lane per contact walks its triangle list in order and runs b3TestOverlap (collide.cpp:84-95)
// of shape A against each triangle (proxy: the three vertices, radius B3_HULL_RADIUS = 0),
// reading and writing that triangle's simplex cache, and stops at the first overlap, so the
// caches of later triangles keep their contents.  Lists cut at the capacity (flagged
// B3AMD_FLAG_TRIANGLE_OVERFLOW by the midphase) are tested up to the capacity.
__global__ __launch_bounds__(256) void mesh_overlap_kernel(MeshArgs a) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t calls = 0, iters = 0, maxIters = 0, hits = 0;
    if (c < a.n) {
        uint32_t* const flag = a.overlap + c;
        uint32_t result = 0;
        const ShapeRec s = a.shapes[a.shapeA[c]];
        const Xf xfA = load_xf(a.xfs + 24 * (size_t)c);
        const Xf xfB = load_xf(a.xfs + 24 * (size_t)c + 12);
        const Proxy pA = s.type == 2u ? hull_proxy(DevHull::at_unit(a.pool, s.hullUnit)) : round_proxy(s);
        const uint32_t count = a.triCounts[c] < a.cap ? a.triCounts[c] : a.cap;
        for (uint32_t k = 0; k < count && result == 0u; ++k) {
            const uint32_t ti = a.triLists[(size_t)c * a.cap + k];
            const uint4 t = a.mesh.tris[ti];
            const float4 A = a.mesh.verts[t.x], B = a.mesh.verts[t.y], C = a.mesh.verts[t.z];
            Proxy pB;
            pB.pts[0] = V3{A.x, A.y, A.z};
            pB.pts[1] = V3{B.x, B.y, B.z};
            pB.pts[2] = V3{C.x, C.y, C.z};
            pB.count = 3;
            pB.hullVerts = nullptr;
            uint32_t* const cw = a.triCaches[(size_t)c * a.cap + k].simplex;
            uint32_t w[5];
            for (int q = 0; q < 5; ++q) w[q] = cw[q];
            bool hit = false;
            const GJKOut g = gjk_cached(xfA, pA, xfB, pB, w, hit);
            for (int q = 0; q < 5; ++q) cw[q] = w[q];
            calls += 1;
            iters += g.iterations;
            maxIters = g.iterations > maxIters ? g.iterations : maxIters;
            hits += hit ? 1u : 0u;
            if (overlap_of(g, s.radius, 0.0f)) result = 1u;
        }
        *flag = result;
    }
    if (a.counters) {  // all lanes of the block reach here
        Counters* const ct = counter_slot(a.counters);
        wave_count(&ct->gjkCalls, calls);
        wave_count(&ct->gjkIters, iters);
        wave_max(&ct->gjkMaxIters, maxIters);
        wave_count(&ct->gjkCacheHits, hits);
    }
}

hipError_t launch_mesh_overlap(const MeshArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    mesh_overlap_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_mesh_contact_wide(const MeshArgs& a, hipStream_t s) {
    if (a.n == 0 || !a.wideContacts) return hipSuccess;
    mesh_contact_wide_kernel<<<kWideContactBlocks, 64, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_mesh_cluster(const MeshArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    constexpr uint32_t perBlock = kClusterBlock / kClusterGroup;
    mesh_cluster_kernel<<<(a.n + perBlock - 1) / perBlock, kClusterBlock, 0, s>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    mesh_cluster_wide_kernel<<<kClusterWideBlocks, kClusterBlock, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
