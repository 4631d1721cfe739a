// Sphere and capsule pairs for gfx950: one lane per pair, a block's pairs grouped by collider.
//
//   sphere-sphere    collide_spheres.cpp:24-56           closed form
//   sphere-capsule   collide_sphere_capsule.cpp:24-133    closed form
//   capsule-capsule  collide_capsules.cpp:176-259         segment closest points, parallel clip
//   sphere-hull      collide_sphere_hull.cpp:26-87        GJK, SAT face query when deep
//   capsule-hull     collide_capsule_hull.cpp:142-222     GJK, face clip or SAT face/edge query
//
// GJK is b3GJK with an empty simplex cache (gjk.cpp:579-737, 845-851): Voronoi-region
// simplex solver, at most 20 support iterations, duplicate-support termination.  The hull
// proxy reads its vertices straight from the device hull record.  Arithmetic keeps the
// reference's FP32 operation order (b3_math.h).
#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "convex_build.h"
#include "device_common.h"
#include "device_hull.h"
#include "kernels.h"

namespace b3amd {
namespace {

// The collider of a sphere/capsule pair (typeA <= typeB): 0 sphere-hull, 1 capsule-hull,
// 2 sphere-sphere, 3 sphere-capsule, 4 capsule-capsule.
__device__ __forceinline__ uint32_t convex_kind(uint32_t t1, uint32_t t2) {
    if (t2 == 2u) return t1 == 0u ? 0u : 1u;
    return t1 == 0u ? (t2 == 0u ? 2u : 3u) : 4u;
}

// s_CollideMatrix (collide.cpp:172-177) for every valid pair that is not hull-hull.  A block
// takes 256 consecutive pairs and orders them by collider in LDS before colliding, so that a
// wave runs one collider (two at a boundary) instead of all five; each pair writes its own
// manifold, so the order inside the block changes nothing.  With counting on, GJK calls and
// support iterations go to the b3_gjk* totals (gjk.cpp:583,673,697).
__global__ __launch_bounds__(256) void convex_kernel(ConvexArgs a) {
    __shared__ uint32_t order[256];
    __shared__ uint32_t kindCount[8];
    const uint32_t p0 = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t kind = 5u;
    if (p0 < a.n && a.classify(p0) == 2u)
        kind = convex_kind(a.shapes[a.pairs[2 * p0]].type, a.shapes[a.pairs[2 * p0 + 1]].type);
    if (threadIdx.x < 8u) kindCount[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t rank = kind < 5u ? atomicAdd(&kindCount[kind], 1u) : 0u;
    __syncthreads();
    uint32_t base = 0, total = 0;
#pragma unroll
    for (uint32_t k = 0; k < 5u; ++k) {
        if (k == kind) base = total;
        total += kindCount[k];
    }
    if (kind < 5u) order[base + rank] = p0;
    __syncthreads();
    uint32_t calls = 0, iters = 0;
    if (threadIdx.x < total) {
        const uint32_t p = order[threadIdx.x];
        Manifold* const out = a.out + a.slot_of(p);
        const ShapeRec s1 = a.shapes[a.pairs[2 * p]];
        const ShapeRec s2 = a.shapes[a.pairs[2 * p + 1]];
        const Xf xf1 = load_xf(a.xf_ptr(p, 0));
        const Xf xf2 = load_xf(a.xf_ptr(p, 1));
        Contact m;
        m.pointCount = 0;
        m.overflow = false;
        if (s2.type == 2u) {
            const DevHull h2 = DevHull::at_unit(a.pool, s2.hullUnit);
            if (s1.type == 0u) collide_sphere_hull(m, xf1, s1, xf2, h2, s2.radius);
            else collide_capsule_hull(m, xf1, s1, xf2, h2, s2.radius);
        } else if (s1.type == 0u && s2.type == 0u) {
            collide_spheres(m, xf1, s1, xf2, s2);
        } else if (s1.type == 0u) {
            collide_sphere_capsule(m, xf1, s1, xf2, s2);
        } else {
            collide_capsules(m, xf1, s1, xf2, s2);
        }
        write_manifold(out, m);
        calls = m.gjkCalls;
        iters = m.gjkIters;
    }
    if (a.counters) {  // all lanes of the block reach here
        Counters* const c = counter_slot(a.counters);
        wave_count(&c->gjkCalls, calls);
        wave_count(&c->gjkIters, iters);
        wave_max(&c->gjkMaxIters, iters);
    }
}

// b3TestOverlap (collide.cpp:84-95) for every pair of convex shapes, one lane per pair: the
// sensor path of b3ConvexContact::TestOverlap (convex_contact.cpp:39-48).  The shape proxies
// are b3ShapeGJKProxy::Set's (collide.cpp:30-82): a sphere's centre, a capsule's two centres
// or a hull's vertices, with the shape radius; GJK reads and writes the pair's simplex cache.
// The destinations are computed before any divergent branch and stored once at the end.
__global__ __launch_bounds__(256) void overlap_kernel(OverlapArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t calls = 0, iters = 0, hits = 0;
    if (p < a.n) {
        uint32_t* const flag = a.overlap + p;
        uint32_t* const cw = a.caches ? a.caches[a.slot_of(p)].simplex : nullptr;
        uint32_t result = kOverlapInvalid;
        uint32_t w[5] = {0u, 0u, 0u, 0u, 0u};
        const uint32_t cls = a.classify(p);
        if (cls != 0u) {
            const ShapeRec s1 = a.shapes[a.pairs[2 * p]];
            const ShapeRec s2 = a.shapes[a.pairs[2 * p + 1]];
            const Xf xf1 = load_xf(a.xf_ptr(p, 0));
            const Xf xf2 = load_xf(a.xf_ptr(p, 1));
            const Proxy p1 = s1.type == 2u ? hull_proxy(DevHull::at_unit(a.pool, s1.hullUnit)) : round_proxy(s1);
            const Proxy p2 = s2.type == 2u ? hull_proxy(DevHull::at_unit(a.pool, s2.hullUnit)) : round_proxy(s2);
            if (cw)
                for (int t = 0; t < 5; ++t) w[t] = cw[t];
            bool hit = false;
            const GJKOut g = gjk_cached(xf1, p1, xf2, p2, cw ? w : nullptr, hit);
            result = overlap_of(g, s1.radius, s2.radius) ? 1u : 0u;
            calls = 1;
            iters = g.iterations;
            hits = hit ? 1u : 0u;
        }
        *flag = result;
        if (cw && cls != 0u)
            for (int t = 0; t < 5; ++t) cw[t] = w[t];
    }
    if (a.counters) {  // all lanes of the block reach here
        Counters* const c = counter_slot(a.counters);
        wave_count(&c->gjkCalls, calls);
        wave_count(&c->gjkIters, iters);
        wave_max(&c->gjkMaxIters, iters);
        wave_count(&c->gjkCacheHits, hits);
    }
}

// The state a new contact's cache starts in (b3ConvexContact's constructor,
// convex_contact.cpp:35-36: simplex count 0, feature cache empty); the fields the reference
// leaves unset are zero.
__global__ __launch_bounds__(256) void cache_reset_kernel(ConvexCache* c, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ConvexCache e;
    for (int t = 0; t < 5; ++t) e.simplex[t] = 0u;
    e.feature.state = kCacheEmpty;
    e.feature.type = 0u;
    e.feature.index1 = 0u;
    e.feature.index2 = 0u;
    c[i] = e;
}

}  // namespace

hipError_t launch_cache_reset(ConvexCache* caches, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    cache_reset_kernel<<<(n + 255) / 256, 256, 0, s>>>(caches, n);
    return hipGetLastError();
}

hipError_t launch_overlap(const OverlapArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    overlap_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_convex(const ConvexArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    convex_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
