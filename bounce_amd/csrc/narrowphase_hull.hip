// Hull-hull contact narrowphase for gfx950.
//
// Three kernels per batch (all on the context stream):
//   cache_hull_kernel    one lane per pair: b3FeatureCache read and the cached rebuilds
//                        (collide_hulls_cache.cpp:279-344); marks pairs that need a SAT.
//   sat_hull_kernel      one 64-lane wave per pair: face queries in both directions and
//                        the edge query (sat.cpp:33-173), staged through LDS.
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

// ------------------------------------------------------------------ SAT kernel

// Ring buffer of Gauss-map hits per wave: at most 64 pending before an iteration plus
// 64 lanes x 6 register slots added by it.
constexpr uint32_t kQueueCap = 512;

// LDS of one wave.  The face phase (vertices, transformed hull-1 vertices and normals)
// and the edge-pair queue of the edge phase share storage.
template <int MAXV, int MAXF, int MAXE>
struct SatLds {
    // Per unique edge, 3 float4: {A.xyz, B.x}, {B.yz, nE.xy}, {nE.z, P.xyz} with
    // hull 1 (hull-2 frame): A = U1, B = V1, nE = -E1, P = P1;
    // hull 2 (own frame):    A = -U2, B = -V2, nE = -E2, P = P2.
    float4 rec1[MAXE * 3];
    float4 rec2[MAXE * 3];
    union {
        struct {
            float4 v1[MAXV];   // hull 1 vertices (hull-1 frame)
            float4 v2[MAXV];   // hull 2 vertices (hull-2 frame)
            float4 tv1[MAXV];  // hull 1 vertices in the hull-2 frame
            float4 tn1[MAXF];  // hull 1 face normals in the hull-2 frame
        } f;
        uint32_t queue[kQueueCap];  // Gauss-map hits awaiting projection: i << 16 | j
    } u;
};

// One face query (sat.cpp:33-58): planes of hull A, transformed by xf, against the
// vertices of hull B staged in LDS.  Each face's support search is split over `split`
// adjacent lanes (contiguous vertex ranges); combining the partial maxima with the lower
// index on ties reproduces the sequential first-maximum of GetSupportVertex
// (hull.inl:31-45).  Returns (separation, face index) reduced over the wave.
template <bool WRITE_TN>
__device__ __forceinline__ void face_query(const Xf& xf, const DevHull& hA, const float4* vB, uint32_t VB,
                                           float4* tn, uint32_t lane, float& bestOut, uint32_t& keyOut) {
    const uint32_t F = hA.F;
    uint32_t split = 1, bestCost = 0xFFFFFFFFu;
    for (uint32_t s = 1; s <= 8; s <<= 1) {
        uint32_t cost = ((F * s + 63) / 64) * ((VB + s - 1) / s);
        if (cost < bestCost) {
            bestCost = cost;
            split = s;
        }
    }
    const uint32_t chunk = (VB + split - 1) / split;
    const uint32_t iters = (F * split + 63) / 64;
    float best = -kMaxFloat;
    uint32_t key = 0;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t t = it * 64 + lane;
        const uint32_t face = t / split, part = t % split;
        const bool active = face < F;
        Plane pl = mul(xf, hA.plane(active ? face : 0u));
        V3 d = -pl.n;
        uint32_t lo = part * chunk, hi = lo + chunk;
        if (hi > VB) hi = VB;
        bool have = active && lo < hi;
        uint32_t mi = lo;
        float mp = have ? dot(d, ld3(vB[lo])) : 0.0f;
        for (uint32_t k = lo + 1; k < hi; ++k) {
            float q = dot(d, ld3(vB[k]));
            if (q > mp) {
                mp = q;
                mi = k;
            }
        }
        for (uint32_t off = split >> 1; off >= 1; off >>= 1) {
            float op = __shfl_xor(mp, off, 64);
            uint32_t oi = __shfl_xor(mi, off, 64);
            bool oh = __shfl_xor((int)have, off, 64) != 0;
            if (oh && (!have || op > mp || (op == mp && oi < mi))) {
                mp = op;
                mi = oi;
                have = true;
            }
        }
        if (active && part == 0) {
            if (WRITE_TN) tn[face] = make_float4(pl.n.x, pl.n.y, pl.n.z, 0.0f);
            float s = distance(ld3(vB[mi]), pl);
            if (s > best) {
                best = s;
                key = face;
            }
        }
    }
    wave_argmax(best, key);
    bestOut = best;
    keyOut = key;
}

__device__ __forceinline__ uint32_t lanemask_lt_count(uint64_t mask, uint32_t lane) {
    return __popcll(mask & ((1ull << lane) - 1ull));
}

// Projects queued Gauss-map hits (sat.cpp:73-108) and folds them into (best, key) with the
// lower (i, j) winning ties, as the sequential loop's strict '>' does.
__device__ __forceinline__ void drain(const float4* rec1, const float4* rec2, const uint32_t* queue, uint32_t head,
                                      uint32_t count, uint32_t lane, V3 C1, float& best, uint32_t& key) {
    if (lane < count) {
        uint32_t e = queue[(head + lane) & (kQueueCap - 1)];
        uint32_t i = e >> 16, j = e & 0xFFFFu;
        float4 a1 = rec1[3 * i + 1], b1 = rec1[3 * i + 2];
        float4 a2 = rec2[3 * j + 1], b2 = rec2[3 * j + 2];
        V3 E1 = -V3{a1.z, a1.w, b1.x};
        V3 P1 = V3{b1.y, b1.z, b1.w};
        V3 E2 = -V3{a2.z, a2.w, b2.x};
        V3 P2 = V3{b2.y, b2.z, b2.w};
        float s = project_edges(P1, E1, length(E1), P1 - C1, P2, E2, length(E2));
        if (s > best || (s == best && e < key)) {
            best = s;
            key = e;
        }
    }
}

// Edge query loop: the register set R (ER edges) is spread over lanes in groups of S
// lanes (K slots each), the broadcast set B (EB edges) is read from LDS, G = 64/S
// broadcast edges per iteration.  REG1: hull 1 edges are the register set.
template <bool REG1, int KMAX>
__device__ __forceinline__ void edge_loop(const float4* recR, const float4* recB, uint32_t ER, uint32_t EB,
                                          uint32_t S, uint32_t K, uint32_t lane, uint32_t* queue,
                                          const float4* rec1, const float4* rec2, V3 C1, float& best,
                                          uint32_t& key) {
    const uint32_t G = 64 / S, s = lane % S, g = lane / S;
    V3 rA[KMAX], rB[KMAX], rE[KMAX];
    bool rv[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) {
        uint32_t r = s + S * k;
        rv[k] = (uint32_t)k < K && r < ER;
        uint32_t rr = rv[k] ? r : 0u;
        float4 q0 = recR[3 * rr], q1 = recR[3 * rr + 1], q2 = recR[3 * rr + 2];
        rA[k] = V3{q0.x, q0.y, q0.z};
        rB[k] = V3{q0.w, q1.x, q1.y};
        rE[k] = V3{q1.z, q1.w, q2.x};
    }
    uint32_t head = 0, tail = 0;
    const uint32_t iters = (EB + G - 1) / G;
    for (uint32_t it = 0; it < iters; ++it) {
        const uint32_t b = it * G + g;
        const bool bv = b < EB;
        const uint32_t bb = bv ? b : 0u;
        const float4 q0 = recB[3 * bb], q1 = recB[3 * bb + 1], q2 = recB[3 * bb + 2];
        const V3 bA = V3{q0.x, q0.y, q0.z}, bB = V3{q0.w, q1.x, q1.y}, bE = V3{q1.z, q1.w, q2.x};
#pragma unroll
        for (int k = 0; k < KMAX; ++k) {
            if ((uint32_t)k >= K) break;
            // b3IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2) (sat.cpp:60-71)
            const V3& U1 = REG1 ? rA[k] : bA;
            const V3& V1 = REG1 ? rB[k] : bB;
            const V3& nE1 = REG1 ? rE[k] : bE;
            const V3& nU2 = REG1 ? bA : rA[k];
            const V3& nV2 = REG1 ? bB : rB[k];
            const V3& nE2 = REG1 ? bE : rE[k];
            float ADC = dot(U1, nE2);
            float BDC = dot(V1, nE2);
            float CBA = dot(nU2, nE1);
            float DBA = dot(nV2, nE1);
            bool pass = bv && rv[k] && (CBA * DBA < 0.0f) && (ADC * BDC < 0.0f) && (CBA * BDC > 0.0f);
            uint64_t mask = __ballot(pass);
            if (mask) {
                if (pass) {
                    uint32_t r = s + S * k;
                    uint32_t i = REG1 ? r : b, j = REG1 ? b : r;
                    queue[(tail + lanemask_lt_count(mask, lane)) & (kQueueCap - 1)] = (i << 16) | j;
                }
                tail += __popcll(mask);
            }
        }
        if (tail - head >= 64) {
            __syncthreads();
            while (tail - head >= 64) {
                drain(rec1, rec2, queue, head, 64, lane, C1, best, key);
                head += 64;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    if (tail > head) drain(rec1, rec2, queue, head, tail - head, lane, C1, best, key);
}

template <int MAXV, int MAXF, int MAXE>
__global__ __launch_bounds__(64) void sat_hull_kernel(SatArgs a) {
    __shared__ SatLds<MAXV, MAXF, MAXE> lds;
    constexpr int KMAX = 6;
    const uint32_t lane = threadIdx.x;
    const uint32_t p = blockIdx.x;
    if (p >= a.n) return;
    if (a.stage[p] != kStageOverlap) {  // resolved by the cache, or not a hull pair
        if (lane == 0) a.out[p].stage = kStageSkip;
        return;
    }

    const uint32_t sa = a.pairs[2 * p], sb = a.pairs[2 * p + 1];
    const uint32_t ha = a.shapes[sa].hull, hb = a.shapes[sb].hull;
    const float totalRadius = a.shapes[sa].radius + a.shapes[sb].radius;
    const Xf xf1 = load_xf(a.xfs + 24 * (size_t)p);
    const Xf xf2 = load_xf(a.xfs + 24 * (size_t)p + 12);
    const DevHull h1 = DevHull::at(a.pool, a.offsets, ha);
    const DevHull h2 = DevHull::at(a.pool, a.offsets, hb);

    for (uint32_t k = lane; k < h1.V; k += 64) lds.u.f.v1[k] = h1.base[2 + k];
    for (uint32_t k = lane; k < h2.V; k += 64) lds.u.f.v2[k] = h2.base[2 + k];
    __syncthreads();

    SatResult r;
    r.f2Index = 0;
    r.f2Sep = -kMaxFloat;
    r.eIndex1 = 0;
    r.eIndex2 = 0;
    r.eSep = -kMaxFloat;
    // ---- face query, hull 1 planes against hull 2 vertices (sat.cpp:33-58)
    const Xf xf = mulT(xf2, xf1);
    face_query<true>(xf, h1, lds.u.f.v2, h2.V, lds.u.f.tn1, lane, r.f1Sep, r.f1Index);
    if (r.f1Sep > totalRadius) {
        r.stage = kStageSepFace1;
        if (lane == 0) a.out[p] = r;
        return;
    }
    // ---- face query, hull 2 planes against hull 1 vertices
    face_query<false>(mulT(xf1, xf2), h2, lds.u.f.v1, h1.V, nullptr, lane, r.f2Sep, r.f2Index);
    if (r.f2Sep > totalRadius) {
        r.stage = kStageSepFace2;
        if (lane == 0) a.out[p] = r;
        return;
    }
    // ---- edge query (sat.cpp:110-173)
    const V3 C1 = mul(xf, h1.centroid());
    for (uint32_t k = lane; k < h1.V; k += 64) {
        V3 t = mul(xf, ld3(lds.u.f.v1[k]));
        lds.u.f.tv1[k] = make_float4(t.x, t.y, t.z, 0.0f);
    }
    __syncthreads();
    const uint32_t E1n = h1.HE >> 1, E2n = h2.HE >> 1;
    for (uint32_t i = lane; i < E1n; i += 64) {
        uint32_t u = h1.uedge(i);
        V3 P1 = ld3(lds.u.f.tv1[u & 0xFFu]);
        V3 Q1 = ld3(lds.u.f.tv1[(u >> 8) & 0xFFu]);
        V3 nE1 = -(Q1 - P1);
        V3 U1 = ld3(lds.u.f.tn1[(u >> 16) & 0xFFu]);
        V3 V1 = ld3(lds.u.f.tn1[u >> 24]);
        lds.rec1[3 * i + 0] = make_float4(U1.x, U1.y, U1.z, V1.x);
        lds.rec1[3 * i + 1] = make_float4(V1.y, V1.z, nE1.x, nE1.y);
        lds.rec1[3 * i + 2] = make_float4(nE1.z, P1.x, P1.y, P1.z);
    }
    for (uint32_t j = lane; j < E2n; j += 64) {
        uint32_t u = h2.uedge(j);
        V3 P2 = ld3(lds.u.f.v2[u & 0xFFu]);
        V3 Q2 = ld3(lds.u.f.v2[(u >> 8) & 0xFFu]);
        V3 nE2 = -(Q2 - P2);
        V3 nU2 = -h2.plane((u >> 16) & 0xFFu).n;
        V3 nV2 = -h2.plane(u >> 24).n;
        lds.rec2[3 * j + 0] = make_float4(nU2.x, nU2.y, nU2.z, nV2.x);
        lds.rec2[3 * j + 1] = make_float4(nV2.y, nV2.z, nE2.x, nE2.y);
        lds.rec2[3 * j + 2] = make_float4(nE2.z, P2.x, P2.y, P2.z);
    }
    __syncthreads();
    // Register set and lane grouping with the fewest wave iterations.
    uint32_t bestCost = 0xFFFFFFFFu, S = 64, K = 1;
    bool reg1 = false;
    for (int side = 0; side < 2; ++side) {
        uint32_t ER = side ? E1n : E2n, EB = side ? E2n : E1n;
        for (uint32_t s = 64; s >= 16; s >>= 1) {
            uint32_t k = (ER + s - 1) / s;
            if (k > (uint32_t)KMAX) continue;
            uint32_t cost = ((EB + 64 / s - 1) / (64 / s)) * k;
            if (cost < bestCost) {
                bestCost = cost;
                S = s;
                K = k;
                reg1 = side;
            }
        }
    }
    float best = -kMaxFloat;
    uint32_t key = 0;
    if (reg1)
        edge_loop<true, KMAX>(lds.rec1, lds.rec2, E1n, E2n, S, K, lane, lds.u.queue, lds.rec1, lds.rec2, C1, best, key);
    else
        edge_loop<false, KMAX>(lds.rec2, lds.rec1, E2n, E1n, S, K, lane, lds.u.queue, lds.rec1, lds.rec2, C1, best, key);
    wave_argmax(best, key);
    r.eIndex1 = 2 * (key >> 16);
    r.eIndex2 = 2 * (key & 0xFFFFu);
    r.eSep = best;
    r.stage = r.eSep > totalRadius ? kStageSepEdge : kStageOverlap;
    if (lane == 0) a.out[p] = r;
}

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
    FeatureCache* fc = a.caches ? &a.caches[p].feature : nullptr;
    Contact m;
    m.pointCount = 0;
    m.overflow = false;
    if (r.stage != kStageOverlap) {
        if (r.stage == kStageSepFace1) set_feature(fc, kCacheSeparation, kFeatureFace1, r.f1Index, r.f1Index);
        else if (r.stage == kStageSepFace2) set_feature(fc, kCacheSeparation, kFeatureFace2, r.f2Index, r.f2Index);
        else set_feature(fc, kCacheSeparation, kFeatureEdge1, r.eIndex1, r.eIndex2);
        write_manifold(a.out + p, m);
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
            write_manifold(a.out + p, m);
            return;
        }
    } else if (r.f1Sep + kTol > r.f2Sep) {
        build_face_contact(m, xf1, r.f1Index, h1, r1, xf2, h2, r2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index);
            write_manifold(a.out + p, m);
            return;
        }
    } else {
        build_face_contact(m, xf2, r.f2Index, h2, r2, xf1, h1, r1, true);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index);
            write_manifold(a.out + p, m);
            return;
        }
    }
    // Heuristic failed: fallback without the tolerance.
    if (r.eSep > fmax12) {
        build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2);
            write_manifold(a.out + p, m);
            return;
        }
    } else if (r.f1Sep > r.f2Sep) {
        build_face_contact(m, xf1, r.f1Index, h1, r1, xf2, h2, r2, false);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index);
            write_manifold(a.out + p, m);
            return;
        }
    } else {
        build_face_contact(m, xf2, r.f2Index, h2, r2, xf1, h1, r1, true);
        if (m.pointCount > 0) {
            set_feature(fc, kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index);
            write_manifold(a.out + p, m);
            return;
        }
    }
    // Last resort: a contact between the edge lines.
    build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
    set_feature(fc, kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2);
    write_manifold(a.out + p, m);
}

// ------------------------------------------------------------------ feature-cache kernel

// b3CollideHulls with a feature cache (collide_hulls_cache.cpp:279-344) up to the flush.
// Returns true when the pair needs the full SAT.
__device__ __forceinline__ bool cache_pair(const CacheArgs& a, uint32_t p) {
    const uint32_t cls = a.classify(p);
    if (cls != 1u) {
        if (cls == 0u) {
            Contact m;
            m.pointCount = 0;
            m.overflow = false;
            write_manifold(a.out + p, m);
            a.out[p].flags = kFlagInvalidPair;
        }
        a.stage[p] = cls == 0u ? kStageSkip : kStageConvex;
        return false;
    }
    FeatureCache* fc = a.caches ? &a.caches[p].feature : nullptr;
    if (!fc || fc->state == kCacheEmpty) {
        // ReadState on an empty cache evaluates face 0 of hull 1 and the result is
        // discarded by the flush (sat.cpp:182-185): go straight to the SAT.
        if (fc) fc->state = kCacheEmpty;
        a.stage[p] = kStageOverlap;
        return true;
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
        a.stage[p] = kStageOverlap;
        return true;
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
        write_manifold(a.out + p, m);
        a.stage[p] = kStageSkip;
        return false;
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
            write_manifold(a.out + p, m);
            a.stage[p] = kStageSkip;
            return false;
        }
    }
    fc->state = kCacheEmpty;
    a.stage[p] = kStageOverlap;
    return true;
}

__global__ __launch_bounds__(256) void cache_hull_kernel(CacheArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    if (!cache_pair(a, p)) a.sat[p].stage = kStageSkip;
}

// ------------------------------------------------------------------ launchers

hipError_t launch_cache_hull(const CacheArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    cache_hull_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_sat_hull(const SatArgs& a, int sizeClass, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    if (sizeClass == 0) sat_hull_kernel<64, 128, 192><<<a.n, 64, 0, s>>>(a);
    else sat_hull_kernel<255, 255, 512><<<a.n, 64, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_contact_hull(const ContactArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    contact_hull_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
