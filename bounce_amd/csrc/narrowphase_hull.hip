// Hull-hull contact narrowphase for gfx950.
//
// Kernels per batch (all on the context stream), around the persistent SAT kernel of
// sat_persistent.hip:
//   cache_hull_kernel    one lane per pair: b3FeatureCache read and the cached rebuilds
//                        (collide_hulls_cache.cpp:279-344); marks pairs that need a SAT.
//   contact_hull_kernel  one lane per pair: the decision tree of b3CollideCache
//                        (collide_hulls_cache.cpp:155-275), edge/face contact, clipping,
//                        reduction, keys, manifold and feature-cache write.
// All arithmetic keeps the reference's FP32 operation order (b3_math.h); the build uses
// -ffp-contract=off and IEEE division/square root.
#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "device_hull.h"
#include "kernels.h"
#include "device_common.h"
#include "contact_build.h"

namespace b3amd {

__device__ __forceinline__ void set_feature(FeatureCache* c, uint32_t state, uint32_t type, uint32_t i1, uint32_t i2) {
    if (!c) return;
    c->state = state;
    c->type = type;
    c->index1 = i1;
    c->index2 = i2;
}

// b3CollideCache after the queries (collide_hulls_cache.cpp:155-275).
__global__ __launch_bounds__(256) void contact_hull_kernel(ContactArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const SatResult r = a.sat[p];
    if (r.stage == kStageSkip) return;  // the cache kernel wrote this pair
    const uint32_t slot = a.slot_of(p);
    Manifold* const out = a.out + slot;
    FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
    Contact m;
    m.pointCount = 0;
    m.overflow = false;
    if (r.stage != kStageOverlap) {
        if (r.stage == kStageSepFace1) set_feature(fc, kCacheSeparation, kFeatureFace1, r.f1Index, r.f1Index);
        else if (r.stage == kStageSepFace2) set_feature(fc, kCacheSeparation, kFeatureFace2, r.f2Index, r.f2Index);
        else set_feature(fc, kCacheSeparation, kFeatureEdge1, r.eIndex1, r.eIndex2);
        write_manifold(out, m);
        return;
    }
    const uint32_t sa = a.pairs[2 * p], sb = a.pairs[2 * p + 1];
    const float r1 = a.shapes[sa].radius, r2 = a.shapes[sb].radius;
    const Xf xf1 = load_xf(a.xfs + 24 * (size_t)p);
    const Xf xf2 = load_xf(a.xfs + 24 * (size_t)p + 12);
    const DevHull h1 = DevHull::at(a.pool, a.offsets, a.shapes[sa].hull);
    const DevHull h2 = DevHull::at(a.pool, a.offsets, a.shapes[sb].hull);

    const float kTol = 0.1f * kLinearSlop;
    const float fmax12 = r.f1Sep > r.f2Sep ? r.f1Sep : r.f2Sep;  // b3Max
    if (r.eSep > fmax12 + kTol) {
        build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2);
            write_manifold(out, m);
            return;
        }
    } else if (r.f1Sep + kTol > r.f2Sep) {
        build_face_contact(m, xf1, r.f1Index, h1, r1, xf2, h2, r2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index);
            write_manifold(out, m);
            return;
        }
    } else {
        build_face_contact(m, xf2, r.f2Index, h2, r2, xf1, h1, r1, true);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index);
            write_manifold(out, m);
            return;
        }
    }
    // Heuristic failed: fallback without the tolerance.
    if (r.eSep > fmax12) {
        build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2);
            write_manifold(out, m);
            return;
        }
    } else if (r.f1Sep > r.f2Sep) {
        build_face_contact(m, xf1, r.f1Index, h1, r1, xf2, h2, r2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index);
            write_manifold(out, m);
            return;
        }
    } else {
        build_face_contact(m, xf2, r.f2Index, h2, r2, xf1, h1, r1, true);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index);
            write_manifold(out, m);
            return;
        }
    }
    // Last resort: a contact between the edge lines.
    build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
    set_feature(fc, kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2);
    write_manifold(out, m);
}

// ------------------------------------------------------------------ feature-cache kernel

// b3CollideHulls with a feature cache (collide_hulls_cache.cpp:279-344) up to the flush.
// Writes the manifold through `out` and the cache through `fc` when the pair is resolved
// here; returns the pair's stage (kStageOverlap when it needs the full SAT).
__device__ __forceinline__ uint32_t cache_pair(const CacheArgs& a, uint32_t p, uint32_t cls, Manifold* out,
                                               FeatureCache* fc) {
    if (cls != 1u) {
        if (cls == 0u) {
            Contact m;
            m.pointCount = 0;
            m.overflow = false;
            write_manifold(out, m);
            out->flags = kFlagInvalidPair;
        }
        return cls == 0u ? kStageSkip : kStageConvex;
    }
    if (!fc || fc->state == kCacheEmpty) {
        // ReadState on an empty cache evaluates face 0 of hull 1 and the result is
        // discarded by the flush (sat.cpp:182-185): go straight to the SAT.
        if (fc) fc->state = kCacheEmpty;
        return kStageOverlap;
    }
    const uint32_t sa = a.pairs[2 * p], sb = a.pairs[2 * p + 1];
    const float r1 = a.shapes[sa].radius, r2 = a.shapes[sb].radius;
    const float totalRadius = r1 + r2;
    const Xf xf1 = load_xf(a.xfs + 24 * (size_t)p);
    const Xf xf2 = load_xf(a.xfs + 24 * (size_t)p + 12);
    const DevHull h1 = DevHull::at(a.pool, a.offsets, a.shapes[sa].hull);
    const DevHull h2 = DevHull::at(a.pool, a.offsets, a.shapes[sb].hull);
    const uint32_t state0 = fc->state, type = fc->type, i1 = fc->index1, i2 = fc->index2;
    // A cached feature outside its hull cannot come from this pair: treat it as a miss.
    const bool inRange = (type == kFeatureEdge1 && i1 + 1 < h1.HE && i2 + 1 < h2.HE && !(i1 & 1u) && !(i2 & 1u)) ||
                         (type == kFeatureFace1 && i1 < h1.F) || (type == kFeatureFace2 && i1 < h2.F);
    if (!inRange || state0 > kCacheEmpty) {
        fc->state = kCacheEmpty;
        return kStageOverlap;
    }
    uint32_t state1;
    switch (type) {
        case kFeatureEdge1: state1 = read_edge(i1, i2, xf1, h1, xf2, h2, totalRadius); break;
        case kFeatureFace1: state1 = read_face(i1, xf1, h1, xf2, h2, totalRadius); break;
        case kFeatureFace2: state1 = read_face(i1, xf2, h2, xf1, h1, totalRadius); break;
        default: state1 = kCacheEmpty; break;
    }
    Contact m;
    m.pointCount = 0;
    m.overflow = false;
    if (state0 == kCacheSeparation && state1 == kCacheSeparation) {
        write_manifold(out, m);
        return kStageSkip;
    }
    if (state0 == kCacheOverlap && state1 == kCacheOverlap) {
        Quat q10{0, 0, 0, 1}, q1{0, 0, 0, 1}, q20{0, 0, 0, 1}, q2{0, 0, 0, 1};
        if (a.sweeps) {
            const float* s = a.sweeps + 16 * (size_t)p;
            q10 = Quat{s[0], s[1], s[2], s[3]};
            q1 = Quat{s[4], s[5], s[6], s[7]};
            q20 = Quat{s[8], s[9], s[10], s[11]};
            q2 = Quat{s[12], s[13], s[14], s[15]};
        }
        switch (type) {
            case kFeatureEdge1: build_edge_contact(m, xf1, i1, h1, xf2, i2, h2, true); break;
            case kFeatureFace1: rebuild_face_contact(m, xf1, i1, h1, r1, xf2, h2, r2, false, q10, q1, q20, q2); break;
            case kFeatureFace2: rebuild_face_contact(m, xf2, i1, h2, r2, xf1, h1, r1, true, q20, q2, q10, q1); break;
            default: break;
        }
        if (m.pointCount > 0) {
            write_manifold(out, m);
            return kStageSkip;
        }
    }
    fc->state = kCacheEmpty;
    return kStageOverlap;
}

// Every destination is addressed once, before the divergent branches, and the stage words
// are stored unconditionally at the end, so no store address is a merge of per-path
// values.
// With counting on, b3_convexCalls counts the hull pairs and b3_convexCacheHits the ones
// the cache resolved (collide_hulls.cpp:304, collide_hulls_cache.cpp:300,334).
__global__ __launch_bounds__(256) void cache_hull_kernel(CacheArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) *a.satCounter = 0u;  // the SAT kernel runs after this one on the stream
    uint32_t call = 0, hit = 0;
    if (p < a.n) {
        const uint32_t slot = a.slot_of(p);
        Manifold* const out = a.out + slot;
        FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
        uint32_t* const stage = a.stage + p;
        SatResult* const sat = a.sat + p;
        const uint32_t cls = a.classify(p);
        const uint32_t st = cache_pair(a, p, cls, out, fc);
        *stage = st;
        sat->stage = st == kStageOverlap ? kStageOverlap : kStageSkip;
        call = cls == 1u ? 1u : 0u;
        hit = (cls == 1u && st == kStageSkip) ? 1u : 0u;
    }
    if (a.counters) {  // all lanes of the block reach here
        wave_count(&a.counters->convexCalls, call);
        wave_count(&a.counters->convexCacheHits, hit);
    }
}

// ------------------------------------------------------------------ launchers

hipError_t launch_cache_hull(const CacheArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    cache_hull_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_contact_hull(const ContactArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    contact_hull_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
