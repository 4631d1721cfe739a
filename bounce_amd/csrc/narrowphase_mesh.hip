// Shape-vs-triangle-mesh narrowphase for gfx950 (b3MeshContact, mesh_contact.cpp).
//
//   mesh_midphase_kernel  one lane per mesh contact: the fat AABB of shape A in the mesh
//                         frame (mesh_contact.cpp:30-60) and the static-tree query in DFS
//                         order, child2 first (static_tree.h:100-142); triangle lists are
//                         compacted into (contact, triangle) work items, grouped by the
//                         convex shape's type so that each wave runs one collider.
//   mesh_narrow_kernel    one lane per (contact, triangle): b3TriangleHull built in
//                         registers and b3CollideShapeAndShape with an empty cache, the
//                         triangle index stamped into every point key (:238-266).
//   mesh_cluster_kernel   one lane per mesh contact: b3ClusterSolver::Run over the
//                         triangle manifolds in midphase order (contact_cluster.cpp:202-590).
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

__device__ uint32_t tree_query(const MeshArgs& a, uint32_t c, const Aabb& box, uint32_t& flags, uint32_t* lds) {
    uint32_t found = 0;
    uint32_t deep[kTreeStack - kLdsStack];
    uint32_t top = 0;
    auto push = [&](uint32_t v) {
        if (top < kLdsStack) lds[kMidphaseBlock * top] = v;
        else deep[top - kLdsStack] = v;
        ++top;
    };
    if (a.mesh.nodeCount > 0) push(0u);
    uint32_t* list = a.triLists + (size_t)c * a.cap;
    while (top) {
        --top;
        const uint32_t ni = top < kLdsStack ? lds[kMidphaseBlock * top] : deep[top - kLdsStack];
        const float4* nd = reinterpret_cast<const float4*>(a.mesh.nodes + 8 * (size_t)ni);
        const float4 q0 = nd[0], q1 = nd[1];  // lower xyz, upper x | upper yz, child1, child2
        if (!overlaps(q0, q1, box)) continue;
        const uint32_t child1 = __float_as_uint(q1.z), child2 = __float_as_uint(q1.w);
        if (child1 == 0xFFFFFFFFu) {
            if (found < a.cap) list[found] = child2;
            ++found;
        } else if (top + 2 <= kTreeStack) {
            push(child1);
            push(child2);
        } else {
            flags |= kFlagTreeStackOverflow;
        }
    }
    if (found > a.cap) flags |= kFlagTriangleOverflow;
    return found;
}

// Appends contact c's m triangles as work items of its collider's segment and returns the
// first item's offset in the segment (0 without items).  Every lane of the wave must call it
// (one atomic per segment with a wave-uniform address, so each folds into a single
// wave-level add; a per-lane segment address would serialise every add).
__device__ __forceinline__ uint32_t list_work(const MeshArgs& a, uint32_t c, uint32_t seg, uint32_t m) {
    uint32_t off = 0;
    for (uint32_t g = 0; g < 3u; ++g) {
        const uint32_t o = atomicAdd(a.totals + g, seg == g ? m : 0u);
        if (seg == g) off = o;
    }
    uint32_t* const items = a.pairContact + (size_t)(seg < 3u ? seg : 0u) * a.n * a.cap + off;
    for (uint32_t k = 0; k < m; ++k) items[k] = c;
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
        const ShapeRec s = a.shapes[si];
        const Xf xfA = load_xf(a.xfs + 24 * (size_t)c);
        const Xf xfB = load_xf(a.xfs + 24 * (size_t)c + 12);
        Aabb box = shape_aabb(s, mulT(xfB, xfA), a.pool, a.offsets);
        const V3 ext{0.2f, 0.2f, 0.2f};  // B3_AABB_EXTENSION (settings.h:54)
        box.lo = box.lo - ext;
        box.hi = box.hi + ext;
        found = tree_query(a, c, box, flags, stackLds + threadIdx.x);
    } else {
        flags |= kFlagInvalidPair;
    }
    const uint32_t m = (flags & kFlagInvalidPair) ? 0u : (found < a.cap ? found : a.cap);
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
        found = tree_query(a, c, box, flags, stackLds + threadIdx.x);
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

// With counting on, hull-triangle collides go to b3_convexCalls and the sphere/capsule
// GJK calls to the b3_gjk* totals, as b3MeshContact::Collide's per-triangle
// b3CollideShapeAndShape calls count them in the reference (mesh_contact.cpp:217-275).
__global__ __launch_bounds__(256) void mesh_narrow_kernel(MeshArgs a) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t0 = a.totals[0], t1 = a.totals[1], t2 = a.totals[2];
    uint32_t hullCalls = 0, hits = 0, calls = 0, iters = 0;
    if (f < t0 + t1 + t2) {
        const uint32_t seg = f < t0 ? 0u : (f < t0 + t1 ? 1u : 2u);
        const uint32_t local = f - (seg == 0u ? 0u : (seg == 1u ? t0 : t0 + t1));
        const uint32_t c = a.pairContact[(size_t)seg * a.n * a.cap + local];
        const uint32_t k = local - a.contactOffset[c];
        const uint32_t ti = a.triLists[(size_t)c * a.cap + k];
        Manifold* const out = a.temp + f;
        const uint4 t = a.mesh.tris[ti];
        const float4 A = a.mesh.verts[t.x], B = a.mesh.verts[t.y], C = a.mesh.verts[t.z];
        const TriHull tri = TriHull::make(V3{A.x, A.y, A.z}, V3{B.x, B.y, B.z}, V3{C.x, C.y, C.z});
        const ShapeRec s = a.shapes[a.shapeA[c]];
        const Xf xfA = load_xf(a.xfs + 24 * (size_t)c);
        const Xf xfB = load_xf(a.xfs + 24 * (size_t)c + 12);
        Contact m;
        m.pointCount = 0;
        m.overflow = false;
        if (s.type == 2u) {
            const DevHull h = DevHull::at_unit(a.pool, s.hullUnit);
            if (a.triCaches) {
                // b3MeshContact::Collide with the triangle's cache (mesh_contact.cpp:238-257);
                // the triangle hull belongs to the mesh body, whose sweep the rebuild reads.
                FeatureCache* const fc = &a.triCaches[(size_t)c * a.cap + k].feature;
                Quat q[4];
                for (int w = 0; w < 4; ++w) {
                    if (a.sweeps) {
                        const float* qs = a.sweeps + 16 * (size_t)c + 4 * w;
                        q[w] = Quat{qs[0], qs[1], qs[2], qs[3]};
                    } else {
                        q[w] = Quat{0.0f, 0.0f, 0.0f, 1.0f};
                    }
                }
                collide_hulls_cached_seq(m, xfA, h, s.radius, xfB, tri, 0.0f, fc, q[0], q[1], q[2], q[3], hits);
            } else {
                collide_cache_seq(m, xfA, h, s.radius, xfB, tri, 0.0f, nullptr);
            }
            hullCalls = 1;
        } else if (s.type == 0u) {
            collide_sphere_hull(m, xfA, s, xfB, tri, 0.0f);
        } else {
            collide_capsule_hull(m, xfA, s, xfB, tri, 0.0f);
        }
        write_manifold(out, m);
        for (uint32_t j = 0; j < m.pointCount; ++j) out->points[j].triangleKey = ti;
        calls = m.gjkCalls;
        iters = m.gjkIters;
    }
    if (a.counters) {  // all lanes of the block reach here
        Counters* const c = counter_slot(a.counters);
        wave_count(&c->convexCalls, hullCalls);
        wave_count(&c->convexCacheHits, hits);
        wave_count(&c->gjkCalls, calls);
        wave_count(&c->gjkIters, iters);
        wave_max(&c->gjkMaxIters, iters);
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

struct Clusters {
    V3 c[4];  // InitializeClusters adds at most three
    uint32_t n;
    __device__ uint32_t find(V3 p) const {  // contact_cluster.cpp:361-377
        uint32_t best = 0;
        float bestValue = kMaxFloat;
        for (uint32_t i = 0; i < n; ++i) {
            float m = dist_sq(p, c[i]);
            if (m < bestValue) {
                bestValue = m;
                best = i;
            }
        }
        return best;
    }
    __device__ void add(V3 centroid) {  // contact_cluster.cpp:336-359
        if (n == 0) {
            c[n++] = centroid;
            return;
        }
        uint32_t b = find(centroid);
        const float kTol = 0.05f;
        if (dist_sq(centroid, c[b]) <= kTol * kTol) {
            c[b] = c[b] + centroid;
            normalize_in_place(c[b]);
            return;
        }
        c[n++] = centroid;
    }
};

__global__ __launch_bounds__(64) void mesh_cluster_kernel(MeshArgs a) {
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= a.n) return;
    Manifold* out = a.out + 3 * (size_t)ci;
    Contact none;
    none.pointCount = 0;
    none.overflow = false;
    const uint32_t flags = a.contactFlags[ci];
    const uint32_t found = a.triCounts[ci];
    const uint32_t m = found < a.cap ? found : a.cap;
    uint32_t numOut = 0;
    if (!(flags & kFlagInvalidPair) && m > 0) {
        const ShapeRec s = a.shapes[a.shapeA[ci]];
        const uint32_t seg = segment_of(s.type);
        const uint32_t base = seg == 0u ? 0u : (seg == 1u ? a.totals[0] : a.totals[0] + a.totals[1]);
        const Manifold* in = a.temp + base + a.contactOffset[ci];
        const Xf xfA = load_xf(a.xfs + 24 * (size_t)ci);
        const Xf xfB = load_xf(a.xfs + 24 * (size_t)ci + 12);
        const float rA = s.radius, rB = 0.0f;
        // Observations: world normals of every point, in manifold order (:480-495).
        V3 obs[kMaxObs];
        uint32_t om[kMaxObs], cl[kMaxObs];
        uint32_t nObs = 0;
        for (uint32_t i = 0; i < m; ++i)
            for (uint32_t j = 0; j < in[i].pointCount; ++j) {
                obs[nObs] = world_point(in[i].points[j], rA, xfA, rB, xfB).normal;
                om[nObs] = (i << 2) | j;
                ++nObs;
            }
        Clusters C;
        C.n = 0;
        // InitializeClusters (:212-334).
        if (nObs <= 3) {
            for (uint32_t i = 0; i < nObs; ++i) C.add(obs[i]);
        } else {
            uint64_t chosenLo = 0, chosenHi[3] = {0, 0, 0};  // bit set over nObs <= 256
            auto chosen = [&](uint32_t i) -> bool {
                return i < 64 ? ((chosenLo >> i) & 1ull) : ((chosenHi[(i - 64) >> 6] >> ((i - 64) & 63)) & 1ull);
            };
            auto choose = [&](uint32_t i) {
                if (i < 64) chosenLo |= 1ull << i;
                else chosenHi[(i - 64) >> 6] |= 1ull << ((i - 64) & 63);
            };
            C.add(obs[0]);
            choose(0);
            {
                V3 A = C.c[0];
                uint32_t index = 0;
                float mx = -kMaxFloat;
                for (uint32_t i = 0; i < nObs; ++i) {
                    if (chosen(i)) continue;
                    float dd = dist_sq(A, obs[i]);
                    if (dd > mx) {
                        mx = dd;
                        index = i;
                    }
                }
                C.add(obs[index]);
                choose(index);
            }
            if (C.n == 1) {
                V3 A = C.c[0];
                uint32_t index = 0;
                float mx = -kMaxFloat;
                for (uint32_t i = 0; i < nObs; ++i) {
                    if (chosen(i)) continue;
                    float dd = dist_sq(A, obs[i]);
                    if (dd > mx) {
                        mx = dd;
                        index = i;
                    }
                }
                C.add(obs[index]);
                choose(index);
            } else {
                V3 A = C.c[0], B = C.c[1];
                uint32_t index = 0;
                float mx = -kMaxFloat;
                for (uint32_t i = 0; i < nObs; ++i) {
                    if (chosen(i)) continue;
                    V3 Pt = obs[i];
                    // b3ClosestPointOnSegment (geometry.h:174-193)
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
                    float dd = dist_sq(Pt, Q);
                    if (dd > mx) {
                        mx = dd;
                        index = i;
                    }
                }
                C.add(obs[index]);
                choose(index);
            }
        }
        // Ten Lloyd iterations (:404-444).  The reference assigns every observation, then sums
        // each cluster's observations in observation order; one pass that assigns observation
        // j and adds it to its cluster's sum adds the same terms in the same order to each
        // sum, so the centroids are the same, and the observations are read once per
        // iteration instead of once per cluster.
        for (uint32_t iter = 0; iter < 10; ++iter) {
            V3 sum[4] = {{0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f}};
            uint32_t count[4] = {0u, 0u, 0u, 0u};
            for (uint32_t j = 0; j < nObs; ++j) {
                const V3 o = obs[j];
                const uint32_t c = C.find(o);
                if (iter == 9u) cl[j] = c;
#pragma unroll
                for (uint32_t k = 0; k < 4u; ++k)
                    if (c == k) {
                        sum[k] = sum[k] + o;
                        ++count[k];
                    }
            }
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i)
                if (i < C.n && count[i] > 0) {
                    float sInv = 1.0f / (float)count[i];  // b3Vec3::operator/= (vec3.h:70-76)
                    C.c[i] = V3{sum[i].x * sInv, sum[i].y * sInv, sum[i].z * sInv};
                }
        }
        // Drop empty clusters, renumbering the rest in order (:449-472).
        uint32_t used = 0;
        for (uint32_t i = 0; i < C.n; ++i) {
            uint32_t cnt = 0;
            for (uint32_t j = 0; j < nObs; ++j) cnt += cl[j] == i;
            if (cnt > 0) {
                for (uint32_t j = 0; j < nObs; ++j)
                    if (cl[j] == i) cl[j] = used;
                C.c[used++] = C.c[i];
            }
        }
        C.n = used;
        // Per cluster: deepest point, offset polygon, reduction, original points (:497-587).
        for (uint32_t i = 0; i < C.n && numOut < 3; ++i) {
            CPoly poly, quad;
            poly.count = 0;
            V3 normal{0.0f, 0.0f, 0.0f};
            for (uint32_t j = 0; j < nObs; ++j) {
                if (cl[j] != i) continue;
                const ManifoldPoint& mp = in[om[j] >> 2].points[om[j] & 3u];
                if (poly.count < (uint32_t)kPolyCap) {
                    poly.v[poly.count].position = world_point(mp, rA, xfA, rB, xfB).point;
                    poly.v[poly.count].clipIndex = j;
                    ++poly.count;
                }
                normal = normal + obs[j];
            }
            if (poly.count == 0) continue;
            normalize_in_place(normal);
            uint32_t minIndex = 0;
            float minSep = kMaxFloat;
            for (uint32_t j = 0; j < poly.count; ++j) {
                const uint32_t o = poly.v[j].clipIndex;
                WPoint w = world_point(in[om[o] >> 2].points[om[o] & 3u], rA, xfA, rB, xfB);
                if (w.separation < minSep) {
                    minIndex = j;
                    minSep = w.separation;
                }
                poly.v[j].position = poly.v[j].position + w.separation * normal;
            }
            reduce_polygon(quad, poly, normal, minIndex);
            Manifold mo;
            mo.pointCount = quad.count;
            mo.flags = flags;
            mo.reserved[0] = mo.reserved[1] = 0;
            for (uint32_t j = 0; j < 4; ++j) {
                if (j < quad.count) {
                    const uint32_t o = quad.v[j].clipIndex;
                    mo.points[j] = in[om[o] >> 2].points[om[o] & 3u];
                } else {
                    mo.points[j].localNormal1 = V3{0.0f, 0.0f, 0.0f};
                    mo.points[j].localPoint1 = V3{0.0f, 0.0f, 0.0f};
                    mo.points[j].localPoint2 = V3{0.0f, 0.0f, 0.0f};
                    mo.points[j].triangleKey = 0;
                    mo.points[j].key[0] = mo.points[j].key[1] = mo.points[j].key[2] = mo.points[j].key[3] = 0;
                }
            }
            out[numOut++] = mo;
        }
    }
    for (uint32_t i = numOut; i < 3; ++i) {
        write_manifold(out + i, none);
        out[i].flags = flags;
    }
    a.manifoldCounts[ci] = numOut;
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

hipError_t launch_mesh_cluster(const MeshArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    mesh_cluster_kernel<<<(a.n + 63) / 64, 64, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
