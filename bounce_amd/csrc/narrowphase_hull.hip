// Hull-hull contact narrowphase for gfx950.
//
// Kernels per batch (all on the context stream), around the persistent SAT kernel of
// sat_persistent.hip:
//   cache_hull_kernel    one lane per pair: b3FeatureCache read and the cached rebuilds
//                        (collide_hulls_cache.cpp:279-344); marks pairs that need a SAT.
//   contact_split_kernel one lane per pair: the separated exits (empty manifold, feature
//                        cache) and a list of the overlapping pairs.
//   contact_group_kernel eight lanes per listed pair: the decision tree of b3CollideCache
//                        (collide_hulls_cache.cpp:155-275), edge/face contact, clipping
//                        with the polygon spread across the lanes, reduction, keys,
//                        manifold and feature-cache write.
//   contact_hull_kernel  one lane per listed pair, the same for the pairs whose clip
//                        polygon outgrows the group kernel (kStageFallback).
//   contact_wide_kernel  the pairs whose polygon outgrows the lane's 32 vertices, with the
//                        polygons in LDS (the reference's b3StackArray grows on the heap).
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

// b3CollideCache after the queries (collide_hulls_cache.cpp:155-275) for pair p with the
// query results r, face contacts built in ws.  Leaves the manifold in m and the feature the
// cache records in f (state, type, index1, index2); returns false, with nothing decided, when
// a clip polygon outgrew ws.
struct FeatureOut {
    uint32_t state, type, i1, i2;
};

template <class WS>
__device__ bool contact_hull_pair(const ContactArgs& a, uint32_t p, const SatResult& r, WS& ws, Contact& m,
                                  FeatureOut& f) {
    m.pointCount = 0;
    m.overflow = false;
    const uint32_t sa = a.pairs[2 * p], sb = a.pairs[2 * p + 1];
    const float r1 = a.shapes[sa].radius, r2 = a.shapes[sb].radius;
    const Xf xf1 = load_xf(a.xf_ptr(p, 0));
    const Xf xf2 = load_xf(a.xf_ptr(p, 1));
    const DevHull h1 = DevHull::at_unit(a.pool, a.shapes[sa].hullUnit);
    const DevHull h2 = DevHull::at_unit(a.pool, a.shapes[sb].hullUnit);
    const float kTol = 0.1f * kLinearSlop;
    const float fmax12 = r.f1Sep > r.f2Sep ? r.f1Sep : r.f2Sep;  // b3Max
    f = FeatureOut{kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2};
    if (r.eSep > fmax12 + kTol) {
        build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
        if (m.pointCount > 0) return true;
    } else if (r.f1Sep + kTol > r.f2Sep) {
        build_face_contact(m, xf1, r.f1Index, h1, r1, xf2, h2, r2, false, ws);
        if (m.overflow) return false;
        if (m.pointCount > 0) {
            f = FeatureOut{kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index};
            return true;
        }
    } else {
        build_face_contact(m, xf2, r.f2Index, h2, r2, xf1, h1, r1, true, ws);
        if (m.overflow) return false;
        if (m.pointCount > 0) {
            f = FeatureOut{kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index};
            return true;
        }
    }
    // Heuristic failed: fallback without the tolerance.
    if (r.eSep > fmax12) {
        build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
        if (m.pointCount > 0) return true;
    } else if (r.f1Sep > r.f2Sep) {
        build_face_contact(m, xf1, r.f1Index, h1, r1, xf2, h2, r2, false, ws);
        if (m.overflow) return false;
        if (m.pointCount > 0) {
            f = FeatureOut{kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index};
            return true;
        }
    } else {
        build_face_contact(m, xf2, r.f2Index, h2, r2, xf1, h1, r1, true, ws);
        if (m.overflow) return false;
        if (m.pointCount > 0) {
            f = FeatureOut{kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index};
            return true;
        }
    }
    // Last resort: a contact between the edge lines.
    build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
    return true;
}

// One lane per pair for the pairs the group kernel marked kStageFallback (polygons beyond its
// capacity), with the reference's 32-vertex inline polygon; a pair whose polygon outgrows
// that is listed for contact_wide_kernel (in the SAT list's storage, which contact_split_kernel
// has finished reading) and writes nothing here.
__global__ __launch_bounds__(256) void contact_hull_kernel(ContactArgs a) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *a.listCount) return;
    const uint32_t p = a.list[k];
    const SatResult r = a.sat[p];
    if (r.stage != kStageFallback) return;  // written by the group kernel
    const uint32_t slot = a.slot_of(p);
    FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
    FaceWs ws;
    Contact m;
    FeatureOut f;
    if (!contact_hull_pair(a, p, r, ws, m, f)) {
        a.wideList[atomicAdd(a.wideCount, 1u)] = p;
        return;
    }
    set_feature(fc, f.state, f.type, f.i1, f.i2);
    write_manifold(a.out + slot, m);
}

// The pairs contact_hull_kernel listed, one per block, its first lane building the contact
// with the polygons in LDS (room for kWidePolyCap vertices).
constexpr uint32_t kWideBlocks = 64;
__global__ __launch_bounds__(64) void contact_wide_kernel(ContactArgs a) {
    __shared__ ClipV polyBuf[kWidePolyCap], scratchBuf[kWidePolyCap];
    __shared__ CPV keptBuf[kWidePolyCap];
    if (threadIdx.x != 0u) return;
    const uint32_t count = *a.wideCount;
    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const uint32_t p = a.wideList[e];
        const uint32_t slot = a.slot_of(p);
        FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
        WideFaceWs ws{PolyView{0u, false, kWidePolyCap, polyBuf}, PolyView{0u, false, kWidePolyCap, scratchBuf},
                      CPolyView{0u, keptBuf}};
        Contact m;
        FeatureOut f;
        contact_hull_pair(a, p, a.sat[p], ws, m, f);  // beyond kWidePolyCap: flagged in the manifold
        set_feature(fc, f.state, f.type, f.i1, f.i2);
        write_manifold(a.out + slot, m);
    }
}

// ------------------------------------------------------------------ group contact kernel
//
// Eight lanes per pair.  Every lane of a group walks the same decision tree; in a face
// contact (collide_hulls.cpp:93-204) lane g holds clip-polygon vertex g.  One clip plane
// (clip.cpp:99-149): each lane takes the previous vertex from its neighbour, emits 0, 1 or
// 2 vertices exactly as the sequential loop does for that vertex, and the outputs are
// compacted in vertex order with two ballots through a small per-group LDS buffer, so the
// polygon is the sequential one.  Reductions over the group keep the first maximum
// (strict comparisons in index order).  A polygon longer than kGroupPoly vertices sends
// the pair to the lane-per-pair kernel (kStageFallback).

constexpr uint32_t kGroup = 8;
constexpr uint32_t kGroupPoly = 8;   // polygon capacity of the group path
constexpr uint32_t kGroupBuf = 16;   // a clip step emits at most two vertices per input

struct GroupLds {
    ClipV poly[kGroupBuf];  // the compacted output of the last clip step
    CPV kept[kGroupPoly];   // vertices within totalRadius, projected (b3ClusterPolygonVertex)
    CPV quad[4];            // the reduced polygon
};

__device__ __forceinline__ void group_sync() {
    // The group lives in one wave: order its LDS writes before the reads of other lanes.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The group's bits of a wave ballot.
__device__ __forceinline__ uint32_t group_bits(bool pred, uint32_t gbase) {
    return (uint32_t)((__ballot(pred) >> gbase) & 0xFFull);
}

__device__ __forceinline__ ClipV shfl_clipv(const ClipV& v, uint32_t src) {
    ClipV o;
    o.position = V3{__shfl(v.position.x, (int)src, (int)kGroup), __shfl(v.position.y, (int)src, (int)kGroup),
                    __shfl(v.position.z, (int)src, (int)kGroup)};
    o.inEdge1 = __shfl(v.inEdge1, (int)src, (int)kGroup);
    o.outEdge1 = __shfl(v.outEdge1, (int)src, (int)kGroup);
    o.inEdge2 = __shfl(v.inEdge2, (int)src, (int)kGroup);
    o.outEdge2 = __shfl(v.outEdge2, (int)src, (int)kGroup);
    return o;
}

// sort_polygon (contact_cluster.cpp:33-56) on LDS, one lane.
__device__ inline void sort_quad(CPV* q, uint32_t count, V3 N) {
    V3 A = q[0].position;
    for (uint32_t i = 1; i < count; ++i) {
        V3 B = q[i].position;
        for (uint32_t j = i + 1; j < count; ++j) {
            V3 C = q[j].position;
            if (!is_ccw(A, B, C, N)) {
                CPV t = q[i];
                q[i] = q[j];
                q[j] = t;
            }
        }
    }
}

// reduce_polygon (contact_cluster.cpp:58-200) from LDS to LDS, one lane; returns the count.
__device__ inline uint32_t reduce_quad(const CPV* in, uint32_t n, V3 N, uint32_t initial, CPV* out) {
    if (n <= 4) {
        for (uint32_t i = 0; i < n; ++i) out[i] = in[i];
        sort_quad(out, n, N);
        return n;
    }
    uint32_t count = 0, chosen = 0;
    out[count++] = in[initial];
    chosen |= 1u << initial;
    {
        V3 A = out[0].position;
        uint32_t index = 0;
        float mx = -kMaxFloat;
        for (uint32_t i = 0; i < n; ++i) {
            if (chosen & (1u << i)) continue;
            V3 d = in[i].position - A;
            float dd = dot(d, d);
            if (dd > mx) {
                mx = dd;
                index = i;
            }
        }
        if (mx < kEpsilon * kEpsilon) return count;
        out[count++] = in[index];
        chosen |= 1u << index;
    }
    {
        V3 A = out[0].position, B = out[1].position;
        uint32_t index = 0;
        float mx = -kMaxFloat;
        for (uint32_t i = 0; i < n; ++i) {
            if (chosen & (1u << i)) continue;
            V3 C = in[i].position;
            V3 nn = cross(B - A, C - A);
            float a2 = fabsf(dot(nn, N));
            if (a2 > mx) {
                mx = a2;
                index = i;
            }
        }
        const float kAreaTol = 0.01f;
        if (mx < 2.0f * kAreaTol) return count;
        V3 C = in[index].position;
        if (!is_ccw(A, B, C, N)) {
            CPV t = out[0];
            out[0] = out[1];
            out[1] = t;
        }
        out[count++] = in[index];
        chosen |= 1u << index;
    }
    {
        V3 A = out[0].position, B = out[1].position;
        uint32_t index = 0;
        float mn = kMaxFloat;
        for (uint32_t i = 0; i < n; ++i) {
            if (chosen & (1u << i)) continue;
            V3 D = in[i].position;
            V3 nn = cross(B - A, D - A);
            float sa2 = dot(nn, N);
            if (sa2 < mn) {
                mn = sa2;
                index = i;
            }
        }
        const float kAreaTol = 0.01f;
        if (fabsf(mn) < 2.0f * kAreaTol) return count;
        out[count++] = in[index];
        chosen |= 1u << index;
    }
    sort_quad(out, count, N);
    return count;
}

// b3BuildFaceContact (collide_hulls.cpp:93-204) by the group; returns the point count, or
// -1 if a polygon outgrew the group's capacity.  With points it writes the manifold
// (lanes 0-3 one point each, lane 0 the header).
__device__ __forceinline__ int group_face_contact(GroupLds& L, uint32_t g, uint32_t gbase, Manifold* out, const Xf& xf1,
                                  uint32_t index1, const DevHull& h1, float r1, const Xf& xf2, const DevHull& h2,
                                  float r2, bool flip) {
    const float totalRadius = r1 + r2;
    const Plane plane1 = mul(xf1, h1.plane(index1));
    const V3 normal1 = mulT(xf2.R, plane1.n);
    // Support face of hull 2 along -normal1 (hull.inl:47-61): strided first maximum.
    const V3 dn = -normal1;
    float bv = -__builtin_inff();
    uint32_t bi = 0xFFFFFFFFu;
    for (uint32_t i = g; i < h2.F; i += kGroup) {
        const float q = dot(dn, h2.plane(i).n);
        if (q > bv) {
            bv = q;
            bi = i;
        }
    }
#pragma unroll
    for (uint32_t off = 1; off < kGroup; off <<= 1) {
        const float ov = __shfl_xor(bv, (int)off, (int)kGroup);
        const uint32_t oi = __shfl_xor(bi, (int)off, (int)kGroup);
        if (ov > bv || (ov == bv && oi < bi)) {
            bv = ov;
            bi = oi;
        }
    }
    const uint32_t index2 = bi;
    // Incident polygon (clip.cpp:33-57): lane g takes the face's g-th half-edge.
    const uint32_t begin2 = h2.face_edge(index2);
    uint32_t e = begin2, deg = 0, mine = begin2;
    do {
        if (deg == g) mine = e;
        ++deg;
        e = h2.next(e);
    } while (e != begin2 && deg <= kGroupPoly);
    if (e != begin2) return -1;
    uint32_t count = deg;
    ClipV v = clipv(mul(xf2, h2.vertex(h2.origin(mine))), h2.next(mine ^ 1u), kNullEdge, mine, kNullEdge);
    // Clip against the side planes of the reference face (clip.cpp:243-280), in the face's
    // edge order.  Lane j builds the side plane of the face's j-th edge up front (the same
    // operations, all planes at once instead of one dependent chain of record loads per
    // clip step); step k takes plane k from lane k.
    const uint32_t begin1 = h1.face_edge(index1);
    uint32_t w1 = begin1, deg1 = 0, myEdge = begin1;
    do {
        if (deg1 == g) myEdge = w1;
        ++deg1;
        w1 = h1.next(w1);
    } while (w1 != begin1 && deg1 <= kGroup);
    if (w1 != begin1) return -1;  // a reference face of more than kGroup edges
    Plane side = edge_side_plane(h1, myEdge);
    side.d += totalRadius;
    const Plane wmine = mul(xf1, side);
    for (uint32_t step = 0; step < deg1; ++step) {
        const uint32_t e1 = (uint32_t)__shfl((int)myEdge, (int)step, (int)kGroup);
        const Plane wp{V3{__shfl(wmine.n.x, (int)step, (int)kGroup), __shfl(wmine.n.y, (int)step, (int)kGroup),
                          __shfl(wmine.n.z, (int)step, (int)kGroup)},
                       __shfl(wmine.d, (int)step, (int)kGroup)};
        const bool valid = g < count;
        const float d2 = valid ? distance(v.position, wp) : 0.0f;
        const uint32_t src = g == 0 ? count - 1u : g - 1u;
        const V3 p1 = V3{__shfl(v.position.x, (int)src, (int)kGroup), __shfl(v.position.y, (int)src, (int)kGroup),
                         __shfl(v.position.z, (int)src, (int)kGroup)};
        const float d1 = __shfl(d2, (int)src, (int)kGroup);
        const uint32_t in1 = __shfl(v.inEdge1, (int)src, (int)kGroup);
        uint32_t nout = 0;
        ClipV o0 = v, o1 = v;
        if (valid) {
            if (d1 <= 0.0f && d2 <= 0.0f) {
                nout = 1;
            } else if (d1 <= 0.0f && d2 > 0.0f) {
                const float f = d1 / (d1 - d2);
                o0 = clipv(p1 + f * (v.position - p1), in1, kNullEdge, kNullEdge, e1);
                nout = 1;
            } else if (d1 > 0.0f && d2 <= 0.0f) {
                const float f = d1 / (d1 - d2);
                o0 = clipv(p1 + f * (v.position - p1), in1, e1, kNullEdge, kNullEdge);
                nout = 2;
            }
        }
        const uint32_t b1 = group_bits(nout >= 1u, gbase), b2 = group_bits(nout == 2u, gbase);
        const uint32_t below = (1u << g) - 1u;
        const uint32_t at = (uint32_t)(__builtin_popcount(b1 & below) + __builtin_popcount(b2 & below));
        count = (uint32_t)(__builtin_popcount(b1) + __builtin_popcount(b2));
        if (nout >= 1u) L.poly[at] = o0;
        if (nout == 2u) L.poly[at + 1u] = o1;
        group_sync();
        if (count > kGroupPoly) return -1;
        if (g < count) v = L.poly[g];
        group_sync();
        if (count == 0) break;
    }
    if (count == 0) return 0;
    // Points within totalRadius, projected onto the reference plane; the deepest first.
    const bool valid = g < count;
    const float sep = valid ? distance(v.position, plane1) : 0.0f;
    const bool keep = valid && sep <= totalRadius;
    const uint32_t kb = group_bits(keep, gbase);
    const uint32_t nk = (uint32_t)__builtin_popcount(kb);
    if (nk == 0) return 0;
    const uint32_t ki = (uint32_t)__builtin_popcount(kb & ((1u << g) - 1u));
    if (keep) {
        L.kept[ki].position = closest_on_plane(v.position, plane1);
        L.kept[ki].clipIndex = g;
    }
    float ms = keep ? sep : __builtin_inff();
    uint32_t mi = keep ? ki : 0xFFFFFFFFu;
#pragma unroll
    for (uint32_t off = 1; off < kGroup; off <<= 1) {
        const float os = __shfl_xor(ms, (int)off, (int)kGroup);
        const uint32_t oi = __shfl_xor(mi, (int)off, (int)kGroup);
        if (os < ms || (os == ms && oi < mi)) {
            ms = os;
            mi = oi;
        }
    }
    group_sync();
    const V3 normal = plane1.n;
    const V3 s_normal = flip ? -normal : normal;
    uint32_t nq = 0;
    if (g == 0) nq = reduce_quad(L.kept, nk, s_normal, mi, L.quad);
    nq = __shfl(nq, 0, (int)kGroup);
    group_sync();
    // Manifold: lane i < 4 writes point i (collide_hulls.cpp:170-199).
    const uint32_t ci = g < nq ? L.quad[g].clipIndex : 0u;
    const ClipV v2 = shfl_clipv(v, ci);
    if (g < 4u) {
        ManifoldPoint q;
        if (g < nq) {
            const V3 v1 = closest_on_plane(v2.position, plane1);
            if (flip) {
                q.localNormal1 = mulT(xf2.R, s_normal);
                q.localPoint1 = mulT(xf2, v2.position);
                q.localPoint2 = mulT(xf1, v1);
                q.key[0] = v2.inEdge2;
                q.key[1] = v2.outEdge2;
                q.key[2] = v2.inEdge1;
                q.key[3] = v2.outEdge1;
            } else {
                q.localNormal1 = mulT(xf1.R, normal);
                q.localPoint1 = mulT(xf1, v1);
                q.localPoint2 = mulT(xf2, v2.position);
                q.key[0] = v2.inEdge1;
                q.key[1] = v2.outEdge1;
                q.key[2] = v2.inEdge2;
                q.key[3] = v2.outEdge2;
            }
            q.triangleKey = kNullTriangle;
        } else {
            q.localNormal1 = V3{0.0f, 0.0f, 0.0f};
            q.localPoint1 = V3{0.0f, 0.0f, 0.0f};
            q.localPoint2 = V3{0.0f, 0.0f, 0.0f};
            q.triangleKey = 0;
            q.key[0] = q.key[1] = q.key[2] = q.key[3] = 0;
        }
        if (nq > 0) out->points[g] = q;
    }
    if (g == 0 && nq > 0) {
        out->pointCount = nq;
        out->flags = 0u;
        out->reserved[0] = 0u;
        out->reserved[1] = 0u;
    }
    return (int)nq;
}

// Field-wise selections: a reference chosen at run time between two locals would keep both
// in private memory; selecting the values keeps them in registers.
__device__ __forceinline__ V3 pick(bool c, V3 a, V3 b) { return V3{c ? b.x : a.x, c ? b.y : a.y, c ? b.z : a.z}; }
__device__ __forceinline__ Xf pick(bool c, const Xf& a, const Xf& b) {
    Xf r;
    r.R.c0 = pick(c, a.R.c0, b.R.c0);
    r.R.c1 = pick(c, a.R.c1, b.R.c1);
    r.R.c2 = pick(c, a.R.c2, b.R.c2);
    r.p = pick(c, a.p, b.p);
    return r;
}
__device__ __forceinline__ DevHull pick(bool c, const DevHull& a, const DevHull& b) {
    DevHull r;
    r.base = c ? b.base : a.base;
    r.V = c ? b.V : a.V;
    r.F = c ? b.F : a.F;
    r.HE = c ? b.HE : a.HE;
    return r;
}

// The separated exits of b3CollideCache (collide_hulls_cache.cpp:155-275) for every pair the
// SAT ran on, one lane per pair; the overlapping pairs are appended to a list (one atomic
// per block) so that the group kernel's waves hold only pairs that build a contact.
__global__ __launch_bounds__(256) void contact_split_kernel(ContactArgs a) {
    __shared__ uint32_t waveCount[4], waveBase[4];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t count = *a.satCount;
    if (blockIdx.x * blockDim.x >= count) return;  // the whole block is past the list
    const uint32_t wave = threadIdx.x >> 6;
    bool overlap = false;
    uint32_t p = 0;
    if (i < count) {
        p = a.satList[i];
        const uint32_t slot = a.slot_of(p);
        Manifold* const out = a.out + slot;
        FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
        const SatResult r = a.sat[p];
        overlap = r.stage == kStageOverlap;
        if (r.stage != kStageSkip && !overlap) {
            if (r.stage == kStageSepFace1) set_feature(fc, kCacheSeparation, kFeatureFace1, r.f1Index, r.f1Index);
            else if (r.stage == kStageSepFace2) set_feature(fc, kCacheSeparation, kFeatureFace2, r.f2Index, r.f2Index);
            else set_feature(fc, kCacheSeparation, kFeatureEdge1, r.eIndex1, r.eIndex2);
            Contact none;
            none.pointCount = 0;
            none.overflow = false;
            write_manifold(out, none);
        }
    }
    const uint64_t m = __ballot(overlap);
    if ((threadIdx.x & 63u) == 0) waveCount[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = waveCount[0] + waveCount[1] + waveCount[2] + waveCount[3];
        uint32_t base = total ? atomicAdd(a.listCount, total) : 0u;
        for (uint32_t w = 0; w < 4u; ++w) {
            waveBase[w] = base;
            base += waveCount[w];
        }
    }
    __syncthreads();
    if (overlap) a.list[waveBase[wave] + mbcnt(m)] = p;
}

// b3CollideCache after the queries (collide_hulls_cache.cpp:155-275) for the listed pairs,
// eight lanes per pair.  Edge contacts are built by every lane and written by lane 0.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void contact_group_kernel(ContactArgs a) {
    __shared__ GroupLds lds[256 / kGroup];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t k = t / kGroup, g = threadIdx.x & (kGroup - 1u);
    const uint32_t gbase = threadIdx.x & 63u & ~(kGroup - 1u);
    GroupLds& L = lds[threadIdx.x / kGroup];
    if (k >= *a.listCount) return;  // whole groups
    const uint32_t p = a.list[k];
    const SatResult r = a.sat[p];
    const uint32_t slot = a.slot_of(p);
    Manifold* const out = a.out + slot;
    FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
    SatResult* const mark = a.sat + p;
    const bool lead = g == 0;
    const uint32_t sa = a.pairs[2 * p], sb = a.pairs[2 * p + 1];
    const float r1 = a.shapes[sa].radius, r2 = a.shapes[sb].radius;
    const Xf xf1 = load_xf(a.xf_ptr(p, 0));
    const Xf xf2 = load_xf(a.xf_ptr(p, 1));
    const DevHull h1 = DevHull::at_unit(a.pool, a.shapes[sa].hullUnit);
    const DevHull h2 = DevHull::at_unit(a.pool, a.shapes[sb].hullUnit);

    const float kTol = 0.1f * kLinearSlop;
    const float fmax12 = r.f1Sep > r.f2Sep ? r.f1Sep : r.f2Sep;  // b3Max
    // Which builder each step of the tree uses: 0 edge, 1 face of hull 1, 2 face of hull 2.
    const uint32_t first = r.eSep > fmax12 + kTol ? 0u : (r.f1Sep + kTol > r.f2Sep ? 1u : 2u);
    const uint32_t second = r.eSep > fmax12 ? 0u : (r.f1Sep > r.f2Sep ? 1u : 2u);
    for (uint32_t step = 0; step < 3u; ++step) {
        const uint32_t kind = step == 0u ? first : (step == 1u ? second : 0u);
        int n = 0;
        if (kind == 0u) {
            Contact m;
            m.pointCount = 0;
            m.overflow = false;
            build_edge_contact(m, xf1, r.eIndex1, h1, xf2, r.eIndex2, h2, false);
            n = (int)m.pointCount;
            if ((n > 0 || step == 2u) && lead) write_manifold(out, m);
        } else {
            // One call site for both reference faces: hull 1's face, or hull 2's flipped.
            const bool flip = kind == 2u;
            n = group_face_contact(L, g, gbase, out, flip ? xf2 : xf1, flip ? r.f2Index : r.f1Index, flip ? h2 : h1,
                                   flip ? r2 : r1, flip ? xf1 : xf2, flip ? h1 : h2, flip ? r1 : r2, flip);
        }
        if (n < 0) {  // beyond the group's polygon capacity
            if (lead) mark->stage = kStageFallback;
            return;
        }
        if (n > 0 || step == 2u) {
            if (lead) {
                if (kind == 0u) set_feature(fc, kCacheOverlap, kFeatureEdge1, r.eIndex1, r.eIndex2);
                else if (kind == 1u) set_feature(fc, kCacheOverlap, kFeatureFace1, r.f1Index, r.f1Index);
                else set_feature(fc, kCacheOverlap, kFeatureFace2, r.f2Index, r.f2Index);
            }
            return;
        }
    }
}

// ------------------------------------------------------------------ feature-cache kernel

// b3CollideHulls with a feature cache (collide_hulls_cache.cpp:279-344) up to the flush, face
// rebuilds in ws.  Writes the manifold through `out` and the cache through `fc` when the pair
// is resolved here; returns the pair's stage (kStageOverlap when it needs the full SAT,
// kStageWide, with nothing written, when a rebuild's polygon outgrew ws).
template <class WS>
__device__ __forceinline__ uint32_t cache_pair(const CacheArgs& a, uint32_t p, uint32_t cls, Manifold* out,
                                               FeatureCache* fc, WS& ws) {
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
    const Xf xf1 = load_xf(a.xf_ptr(p, 0));
    const Xf xf2 = load_xf(a.xf_ptr(p, 1));
    const DevHull h1 = DevHull::at_unit(a.pool, a.shapes[sa].hullUnit);
    const DevHull h2 = DevHull::at_unit(a.pool, a.shapes[sb].hullUnit);
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
        Quat q10, q1, q20, q2;
        a.sweep(p, q10, q1, q20, q2);
        switch (type) {
            case kFeatureEdge1: build_edge_contact(m, xf1, i1, h1, xf2, i2, h2, true); break;
            case kFeatureFace1: rebuild_face_contact(m, xf1, i1, h1, r1, xf2, h2, r2, false, q10, q1, q20, q2, ws); break;
            case kFeatureFace2: rebuild_face_contact(m, xf2, i1, h2, r2, xf1, h1, r1, true, q20, q2, q10, q1, ws); break;
            default: break;
        }
        if (m.overflow) return kStageWide;  // nothing written: the wide pass runs the pair again
        if (m.pointCount > 0) {
            write_manifold(out, m);
            return kStageSkip;
        }
    }
    fc->state = kCacheEmpty;
    return kStageOverlap;
}

// Appends the flagged pairs of a 256-lane block to a list with one atomic per block (waves
// in block order, lanes in lane order).  Every lane of the block calls it; sh holds 8 words.
__device__ __forceinline__ void block_append(bool flag, uint32_t p, uint32_t* list, uint32_t* count, uint32_t* sh) {
    const uint32_t wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    if ((threadIdx.x & 63u) == 0) sh[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t total = sh[0] + sh[1] + sh[2] + sh[3];
        uint32_t base = total ? atomicAdd(count, total) : 0u;
        for (uint32_t w = 0; w < 4u; ++w) {
            sh[4 + w] = base;
            base += sh[w];
        }
    }
    __syncthreads();
    if (flag) list[sh[4 + wave] + mbcnt(m)] = p;
    __syncthreads();
}

// Appends the flagged pairs of one wave to a list with one atomic (lane 0 of the wave must
// be active; every lane of the wave calls it).
__device__ __forceinline__ void wave_append(bool flag, uint32_t p, uint32_t* list, uint32_t* count) {
    const uint64_t m = __ballot(flag);
    if (!m) return;
    uint32_t base = 0;
    if ((threadIdx.x & 63u) == 0) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = (uint32_t)__shfl((int)base, 0, 64);
    if (flag) list[base + mbcnt(m)] = p;
}

// The start of b3CollideHulls with a feature cache (collide_hulls_cache.cpp:279-297) for
// pair p, by its lane: invalid pairs get an empty manifold and the invalid flag,
// sphere/capsule pairs go to the convex kernel, hull pairs with a cached feature are marked
// kStageCached (the return value) and the others (no cache, an empty one, or a state outside
// the enum) go to the SAT list (toSat) with their cache flushed.  ReadState on an empty
// cache evaluates face 0 of hull 1 and the result is discarded by the flush
// (sat.cpp:182-185), so those go straight to the SAT.  `call` is 1 for a hull pair
// (b3_convexCalls, collide_hulls.cpp:304); `pre` receives a hull pair's record units and
// radii, stored in its SAT result slot for a pair bound for the SAT.
__device__ __forceinline__ bool list_pair(const CacheArgs& a, uint32_t p, bool& toSat, uint32_t& call, float4& pre) {
    const uint32_t slot = a.slot_of(p);
    Manifold* const out = a.out + slot;
    FeatureCache* const fc = a.caches ? &a.caches[slot].feature : nullptr;
    SatResult* const sat = a.sat + p;
    const uint32_t cls = a.classify(p);
    bool cached = false;
    if (cls == 0u) {
        Contact m;
        m.pointCount = 0;
        m.overflow = false;
        write_manifold(out, m);
        out->flags = kFlagInvalidPair;
    } else if (cls == 1u) {
        call = 1u;
        cached = fc && fc->state < kCacheEmpty;
        toSat = !cached;
        if (fc && toSat) fc->state = kCacheEmpty;
    }
    a.stage[p] = cls == 2u ? kStageConvex : (toSat ? kStageOverlap : (cached ? kStageCached : kStageSkip));
    if (toSat) pre = sat_prefill(a, p, sat, kStageOverlap);
    else if (cached) pre = prefill_units(a, p);  // stored only if the read sends it on
    return cached;
}

// The batch without feature caches, one lane per pair (list_pair); every hull pair goes to
// the SAT list.
__global__ __launch_bounds__(256) void cache_list_kernel(CacheArgs a) {
    __shared__ uint32_t sh[8];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool toSat = false;
    uint32_t call = 0;
    float4 pre;
    if (p < a.n) list_pair(a, p, toSat, call, pre);
    block_append(toSat, p, a.satList, a.counts + kCountSat, sh);
    if (a.counters) wave_count(&counter_slot(a.counters)->convexCalls, call);
}

// ReadFace (sat.cpp:209-222) by a group of G lanes: the support search over hull 2's
// vertices strided across the lanes.  Only the maximum of dot(-n, v) is needed: the
// reference's separation dot(n, v*) - d equals -max - d exactly, since negating n negates
// every product and sum exactly, whichever of tied maxima the reference picks.
template <uint32_t G>
__device__ __forceinline__ uint32_t group_read_face(uint32_t g, uint32_t index1, const Xf& xf1, const DevHull& h1,
                                                    const Xf& xf2, const DevHull& h2, float totalRadius) {
    const Xf xf = mulT(xf2, xf1);
    const Plane pl = mul(xf, h1.plane(index1));
    const V3 d = -pl.n;
    float mp = -__builtin_inff();
    for (uint32_t k = g; k < h2.V; k += G) mp = fmaxf(mp, dot(d, h2.vertex(k)));
#pragma unroll
    for (uint32_t off = 1; off < G; off <<= 1) mp = fmaxf(mp, __shfl_xor(mp, (int)off, (int)G));
    const float sep = -mp - pl.d;
    return sep > totalRadius ? kCacheSeparation : kCacheOverlap;
}

// Whether ReadFace's state is kCacheSeparation by the bounding sphere of hull 2 alone
// (centroid c, radius R >= |v - c| for every vertex): max dot(-n, v) <= dot(-n, c) + |n| R
// for the same plane (n, offset) as group_read_face computes.  The FP32 search result lies
// within (9 u) |n| (|c| + R) of the exact maximum and the FP32 bound within as much of its
// exact value, so when the bound clears totalRadius by more than 1e-5 (|n| (|c| + R) +
// |offset|), far above those rounding terms, the search's separation is above totalRadius
// too.  Non-finite inputs fail the comparison (and take the search).
__device__ __forceinline__ bool face_bound_separated(uint32_t index1, const Xf& xf1, const DevHull& h1, const Xf& xf2,
                                                     const DevHull& h2, float totalRadius) {
    const Xf xf = mulT(xf2, xf1);
    const Plane pl = mul(xf, h1.plane(index1));
    const V3 d = -pl.n;
    const float4 cs = h2.base[1];  // centroid, bounding radius
    const V3 c{cs.x, cs.y, cs.z};
    const float nd = sqrtf(dot(d, d)), cn = sqrtf(dot(c, c));
    const float bound = dot(d, c) + nd * cs.w;
    const float slack = 1e-5f * (nd * (cn + cs.w) + fabsf(pl.d));
    return (-bound - pl.d) - slack > totalRadius;
}

// An empty manifold (all 240 bytes zero, as write_manifold writes pointCount 0) stored by a
// group of G lanes, one 16-byte word per lane and store.
template <uint32_t G>
__device__ __forceinline__ void group_write_empty(Manifold* out, uint32_t g) {
    static_assert(sizeof(Manifold) == 15 * 16, "manifold of 15 16-byte words");
    float4* const w = reinterpret_cast<float4*>(out);
    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (uint32_t k = 0; k < 15u; k += G)
        if (k + g < 15u) w[k + g] = z;
}

enum : uint32_t { kCachedHit = 0, kCachedMiss = 1, kCachedFallback = 2, kCachedRebuild = 3 };

// The inputs of one cached pair.
struct CachedPair {
    Xf xf1, xf2;
    DevHull h1, h2;
    float r1, r2;
    uint32_t state0, type, i1, i2;
    __device__ static CachedPair load(const CacheArgs& a, uint32_t p, const FeatureCache* fc) {
        // Record units and radii, as list_pair stored them (sat_prefill).
        return make(a, p, fc, *reinterpret_cast<const float4*>(a.sat + p));
    }
    __device__ static CachedPair make(const CacheArgs& a, uint32_t p, const FeatureCache* fc, float4 pre) {
        CachedPair c;
        c.r1 = pre.y;
        c.r2 = pre.w;
        c.xf1 = load_xf(a.xf_ptr(p, 0));
        c.xf2 = load_xf(a.xf_ptr(p, 1));
        c.h1 = DevHull::at_unit(a.pool, __float_as_uint(pre.x));
        c.h2 = DevHull::at_unit(a.pool, __float_as_uint(pre.z));
        c.state0 = fc->state;
        c.type = fc->type;
        c.i1 = fc->index1;
        c.i2 = fc->index2;
        return c;
    }
    // A cached feature outside its hull cannot come from this pair.
    __device__ bool in_range() const {
        return (type == kFeatureEdge1 && i1 + 1 < h1.HE && i2 + 1 < h2.HE && !(i1 & 1u) && !(i2 & 1u)) ||
               (type == kFeatureFace1 && i1 < h1.F) || (type == kFeatureFace2 && i1 < h2.F);
    }
};

// b3CollideHulls with a feature cache (collide_hulls_cache.cpp:279-344) up to the rebuilds,
// given ReadState's new state (sat.cpp:177-278): the separation hit, or the decision to
// rebuild: an overlap hit needs the edge rebuild, or the face rebuild when the relative
// rotation allows it (collide_hulls_cache.cpp:117-126; hull 2's face swaps the bodies,
// :325-329); anything else is a miss.
__device__ __forceinline__ uint32_t cached_decide(const CacheArgs& a, uint32_t p, const CachedPair& c, uint32_t state1) {
    if (c.state0 == kCacheSeparation && state1 == kCacheSeparation) return kCachedHit;
    if (c.state0 != kCacheOverlap || state1 != kCacheOverlap) return kCachedMiss;
    if (c.type == kFeatureEdge1) return kCachedRebuild;
    Quat q10, q1, q20, q2;
    a.sweep(p, q10, q1, q20, q2);
    const bool flip = c.type == kFeatureFace2;
    const bool allowed = flip ? rebuild_face_allowed(q20, q2, q10, q1) : rebuild_face_allowed(q10, q1, q20, q2);
    return allowed ? kCachedRebuild : kCachedMiss;
}

// The cached rebuilds (collide_hulls_cache.cpp:35-153) of one pair by its group of eight:
// the edge contact on every lane, or the face contact with the clip polygon across the group.
// Writes the manifold on a hit.
__device__ __forceinline__ uint32_t cached_rebuild(const CacheArgs& a, GroupLds& L, uint32_t g, uint32_t gbase,
                                                   uint32_t p, Manifold* out, const FeatureCache* fc) {
    const CachedPair c = CachedPair::load(a, p, fc);
    if (c.type == kFeatureEdge1) {
        Contact m;
        m.pointCount = 0;
        m.overflow = false;
        build_edge_contact(m, c.xf1, c.i1, c.h1, c.xf2, c.i2, c.h2, true);
        if (m.pointCount == 0) return kCachedMiss;
        if (g == 0) write_manifold(out, m);
        return kCachedHit;
    }
    const bool flip = c.type == kFeatureFace2;
    const Xf xa = pick(flip, c.xf1, c.xf2), xb = pick(flip, c.xf2, c.xf1);
    const DevHull ha = pick(flip, c.h1, c.h2), hb = pick(flip, c.h2, c.h1);
    const int n = group_face_contact(L, g, gbase, out, xa, c.i1, ha, flip ? c.r2 : c.r1, xb, hb, flip ? c.r1 : c.r2,
                                     flip);
    if (n < 0) return kCachedFallback;
    return n > 0 ? kCachedHit : kCachedMiss;
}

constexpr uint32_t kReadGroup = 4;
constexpr uint32_t kCachedNone = 4;
static_assert(kReadPairsPerBlock == 256u, "cache-read block: one lane per pair");

// The outcome of one cached pair's read (r: kCached*, or kCachedNone for no pair): a miss
// flushes the cache and appends the pair to the SAT list, a rebuild appends it to the block's
// rebuild sub-list; both store the record units and radii `pre` in the pair's SAT result
// slot, where the SAT and rebuild kernels read them.  Every lane of the block calls it
// together.
__device__ __forceinline__ void cached_finish(const CacheArgs& a, uint32_t r, uint32_t p, FeatureCache* fc,
                                              float4 pre) {
    if (r == kCachedMiss) {
        fc->state = kCacheEmpty;
        *reinterpret_cast<float4*>(a.sat + p) = pre;
        a.sat[p].stage = kStageOverlap;
    } else if (r == kCachedRebuild) {
        *reinterpret_cast<float4*>(a.sat + p) = pre;
    }
    wave_append(r == kCachedMiss, p, a.satList, a.counts + kCountSat);
    const uint32_t k = blockIdx.x % kRebuildLists;
    wave_append(r == kCachedRebuild, p, a.rebuildList + (size_t)k * rebuild_capacity(a.n), a.counts + kCountRebuild + k);
    if (a.counters) wave_count(&counter_slot(a.counters)->convexCacheHits, r == kCachedHit ? 1u : 0u);
}

// The batch with feature caches: list_pair, then for the pairs marked kStageCached the
// feature-cache read (ReadState, sat.cpp:177-278), one lane per pair over the block's 256
// pairs.  An edge feature is read by its lane; a face feature whose state the bounding
// sphere of the other hull settles (face_bound_separated) needs no vertex; the remaining
// face features are queued in LDS and read by groups of four lanes (group_read_face).
// Separation hits get their empty manifold here, written by the whole block; overlap hits
// go to the rebuild lists; a miss flushes the cache and appends the pair to the SAT list.
// With counting on, b3_convexCacheHits counts the separation hits
// (collide_hulls_cache.cpp:300).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void cache_read_kernel(CacheArgs a) {
    __shared__ uint32_t sh[8], readQ[256], hitQ[256], qn[2];
    __shared__ float4 preQ[256];  // the queued pairs' record units and radii
    if (threadIdx.x < 2u) qn[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t p = blockIdx.x * 256u + threadIdx.x;
    bool toSat = false;
    uint32_t call = 0, r = kCachedNone;
    FeatureCache* fc = nullptr;
    float4 pre;
    if (p < a.n && list_pair(a, p, toSat, call, pre)) {
        const uint32_t slot = a.slot_of(p);
        fc = &a.caches[slot].feature;
        const CachedPair c = CachedPair::make(a, p, fc, pre);
        const float totalRadius = c.r1 + c.r2;
        if (!c.in_range()) {
            r = kCachedMiss;
        } else if (c.type == kFeatureEdge1) {
            r = cached_decide(a, p, c, read_edge(c.i1, c.i2, c.xf1, c.h1, c.xf2, c.h2, totalRadius));
        } else {
            const bool two = c.type == kFeatureFace2;
            if (face_bound_separated(c.i1, pick(two, c.xf1, c.xf2), pick(two, c.h1, c.h2), pick(two, c.xf2, c.xf1),
                                     pick(two, c.h2, c.h1), totalRadius))
                r = cached_decide(a, p, c, kCacheSeparation);
            else {
                const uint32_t q = atomicAdd(&qn[0], 1u);
                readQ[q] = p;
                preQ[q] = pre;
            }
        }
        if (r == kCachedHit) hitQ[atomicAdd(&qn[1], 1u)] = slot;
    }
    block_append(toSat, p, a.satList, a.counts + kCountSat, sh);
    if (a.counters) wave_count(&counter_slot(a.counters)->convexCalls, call);
    cached_finish(a, r, p, fc, pre);
    __syncthreads();
    const uint32_t m = qn[0], nh = qn[1];
    // The empty manifolds of the hits, 15 16-byte words each, across the block.
    static_assert(sizeof(Manifold) == 15 * 16, "manifold of 15 16-byte words");
    for (uint32_t w = threadIdx.x; w < 15u * nh; w += 256u)
        reinterpret_cast<float4*>(a.out + hitQ[w / 15u])[w % 15u] = make_float4(0.f, 0.f, 0.f, 0.f);
    // The queued face reads, 64 pairs per round.
    for (uint32_t base = 0; base < m; base += 256u / kReadGroup) {
        const uint32_t q = base + threadIdx.x / kReadGroup, g = threadIdx.x & (kReadGroup - 1u);
        uint32_t r2 = kCachedNone, p2 = 0;
        FeatureCache* fc2 = nullptr;
        float4 pre2 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (q < m) {
            p2 = readQ[q];
            pre2 = preQ[q];
            const uint32_t slot = a.slot_of(p2);
            fc2 = &a.caches[slot].feature;
            const CachedPair c = CachedPair::make(a, p2, fc2, pre2);
            const bool two = c.type == kFeatureFace2;
            const uint32_t state1 = group_read_face<kReadGroup>(g, c.i1, pick(two, c.xf1, c.xf2), pick(two, c.h1, c.h2),
                                                                pick(two, c.xf2, c.xf1), pick(two, c.h2, c.h1),
                                                                c.r1 + c.r2);
            r2 = cached_decide(a, p2, c, state1);
            if (r2 == kCachedHit) group_write_empty<kReadGroup>(a.out + slot, g);
            if (g != 0) r2 = kCachedNone;
        }
        cached_finish(a, r2, p2, fc2, pre2);
    }
}

// The rebuild lists, eight lanes per pair.  Group q of the grid's G groups takes entry
// i = q / kRebuildLists of sub-list q % kRebuildLists, then q + G, ..., until i passes the
// longest sub-list.  A miss flushes the cache and appends the pair to the SAT list; a face
// rebuild whose polygon outgrows the group marks the pair for the lane-per-pair kernel.
// With counting on, b3_convexCacheHits counts the overlap hits
// (collide_hulls_cache.cpp:334).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void cache_group_kernel(CacheArgs a) {
    __shared__ GroupLds lds[256 / kGroup];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = threadIdx.x & (kGroup - 1u);
    const uint32_t gbase = threadIdx.x & 63u & ~(kGroup - 1u);
    GroupLds& L = lds[threadIdx.x / kGroup];
    const uint32_t groups = gridDim.x * (blockDim.x / kGroup);
    const uint32_t cap = rebuild_capacity(a.n);
    uint32_t longest = 0;
    for (uint32_t k = g; k < kRebuildLists; k += kGroup) longest = max(longest, a.counts[kCountRebuild + k]);
#pragma unroll
    for (uint32_t off = 1; off < kGroup; off <<= 1) longest = max(longest, (uint32_t)__shfl_xor((int)longest, (int)off, (int)kGroup));
    uint32_t hits = 0;
    for (uint32_t q = t / kGroup;; q += groups) {
        const uint32_t i = q / kRebuildLists, k = q % kRebuildLists;
        if (i >= longest) break;
        if (i >= a.counts[kCountRebuild + k]) continue;
        const uint32_t p = a.rebuildList[(size_t)k * cap + i];
        const uint32_t slot = a.slot_of(p);
        Manifold* const out = a.out + slot;
        FeatureCache* const fc = &a.caches[slot].feature;
        uint32_t* const stage = a.stage + p;
        const uint32_t r = cached_rebuild(a, L, g, gbase, p, out, fc);
        if (g == 0) {
            hits += r == kCachedHit ? 1u : 0u;
            if (r == kCachedMiss) {
                fc->state = kCacheEmpty;
                a.sat[p].stage = kStageOverlap;
                a.satList[atomicAdd(a.counts + kCountSat, 1u)] = p;
            }
            *stage = r == kCachedFallback ? kStageCacheFallback : kStageSkip;
        }
    }
    if (a.counters) wave_count(&counter_slot(a.counters)->convexCacheHits, hits);
}

// The pairs the group kernel marked kStageCacheFallback, one lane each, with the
// reference's 32-vertex clip polygon; a pair whose polygon outgrows that is listed for
// cache_wide_kernel (in the rebuild lists' storage, which cache_group_kernel has finished
// reading) and writes nothing here.
__global__ __launch_bounds__(256) void cache_fallback_kernel(CacheArgs a) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool miss = false, hit = false;
    if (p < a.n && a.stage[p] == kStageCacheFallback) {
        const uint32_t slot = a.slot_of(p);
        Manifold* const out = a.out + slot;
        FeatureCache* const fc = &a.caches[slot].feature;
        SatResult* const sat = a.sat + p;
        FaceWs ws;
        const uint32_t st = cache_pair(a, p, 1u, out, fc, ws);
        if (st == kStageWide) {
            a.rebuildList[atomicAdd(a.counts + kCountCacheWide, 1u)] = p;
        } else {
            hit = st == kStageSkip;
            miss = !hit;
            if (miss) sat->stage = kStageOverlap;
        }
    }
    wave_append(miss, p, a.satList, a.counts + kCountSat);
    if (a.counters) wave_count(&counter_slot(a.counters)->convexCacheHits, hit ? 1u : 0u);
}

// The pairs cache_fallback_kernel listed, one per block, its first lane rebuilding with the
// polygons in LDS (room for kWidePolyCap vertices); a miss goes to the SAT list.
__global__ __launch_bounds__(64) void cache_wide_kernel(CacheArgs a) {
    __shared__ ClipV polyBuf[kWidePolyCap], scratchBuf[kWidePolyCap];
    __shared__ CPV keptBuf[kWidePolyCap];
    if (threadIdx.x != 0u) return;
    const uint32_t count = a.counts[kCountCacheWide];
    uint32_t hits = 0;
    for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
        const uint32_t p = a.rebuildList[e];
        const uint32_t slot = a.slot_of(p);
        WideFaceWs ws{PolyView{0u, false, kWidePolyCap, polyBuf}, PolyView{0u, false, kWidePolyCap, scratchBuf},
                      CPolyView{0u, keptBuf}};
        // Beyond kWidePolyCap the rebuild keeps its flagged polygon and the pair takes the SAT.
        uint32_t st = cache_pair(a, p, 1u, a.out + slot, &a.caches[slot].feature, ws);
        if (st == kStageWide) {
            a.caches[slot].feature.state = kCacheEmpty;
            st = kStageOverlap;
        }
        if (st == kStageSkip) {
            ++hits;
        } else {
            a.sat[p].stage = kStageOverlap;
            a.satList[atomicAdd(a.counts + kCountSat, 1u)] = p;
        }
    }
    if (a.counters && hits) atomicAdd(&counter_slot(a.counters)->convexCacheHits, hits);
}

// ------------------------------------------------------------------ launchers

hipError_t launch_cache_hull(const CacheArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    hipError_t e = launch_zero_words(a.counts, kCountWords, s);
    if (e != hipSuccess) return e;
    if (!a.caches) {  // nothing is cached
        cache_list_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
        return hipGetLastError();
    }
    // The grids cover the batch; the stage words select the pairs.
    cache_read_kernel<<<(a.n + kReadPairsPerBlock - 1) / kReadPairsPerBlock, kReadPairsPerBlock, 0, s>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // A grid of at most 2048 blocks (65,536 groups) walks the rebuild lists.
    uint64_t blocks = ((uint64_t)a.n * kGroup + 255) / 256;
    if (blocks > 2048u) blocks = 2048u;
    cache_group_kernel<<<(uint32_t)blocks, 256, 0, s>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    cache_fallback_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    cache_wide_kernel<<<kWideBlocks, 64, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_contact_hull(const ContactArgs& a, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    contact_split_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // Grids cover the n-pair bound: the list length stays on the device.
    const uint64_t threads = (uint64_t)a.n * kGroup;
    contact_group_kernel<<<(uint32_t)((threads + 255) / 256), 256, 0, s>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    contact_hull_kernel<<<(a.n + 255) / 256, 256, 0, s>>>(a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    contact_wide_kernel<<<kWideBlocks, 64, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
