// Lane-per-pair contact building, templated over the hull accessor (a device hull record
// or a triangle hull built on the fly): clipping (clip.cpp), polygon reduction
// (contact_cluster.cpp:22-200), edge/face contacts and the cached rebuilds
// (collide_hulls.cpp:26-204, collide_hulls_cache.cpp:35-153) and the feature-cache reads
// (sat.cpp:177-278).  A hull accessor provides V, F, HE, vertex(i), plane(i), centroid(),
// origin(h), face(h), next(h) and face_edge(f), with the twin of half-edge h at h ^ 1.
#ifndef B3AMD_CONTACT_BUILD_H
#define B3AMD_CONTACT_BUILD_H

#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "device_common.h"
#include "device_hull.h"
#include "kernels.h"

namespace b3amd {

// sat.cpp:73-108 with the per-edge lengths and P1 - C1 precomputed (same operations).
__device__ inline float project_edges(V3 P1, V3 E1, float L1, V3 D1, V3 P2, V3 E2, float L2) {
    if (L1 < kLinearSlop) return -kMaxFloat;
    if (L2 < kLinearSlop) return -kMaxFloat;
    const float kTol = 0.005f;
    V3 E1xE2 = cross(E1, E2);
    float L = length(E1xE2);
    if (L < kTol * L1 * L2) return -kMaxFloat;
    V3 N = (1.0f / L) * E1xE2;
    if (dot(N, D1) < 0.0f) N = -N;
    return dot(N, P2 - P1);
}

// ------------------------------------------------------------------ contact building

// Clip polygons.  The lane paths hold b3StackArray's inline capacity (clip.cpp); a polygon
// that outgrows it sets `overflow`, and the pair goes to the wide pass, whose polygons live
// in LDS with room for kWidePolyCap vertices: a convex polygon clipped by a convex face
// gains at most one vertex per plane, so two faces of at most 255 vertices (the hull limit)
// stay within 510.
constexpr uint32_t kPolyCap = 32;
constexpr uint32_t kWidePolyCap = 512;

template <uint32_t Cap>
struct PolyArr {  // in the lane's own storage
    static constexpr uint32_t cap = Cap;
    uint32_t count;
    bool overflow;
    ClipV v[Cap];
    __device__ void push(const ClipV& x) {
        if (count < Cap) v[count++] = x;
        else overflow = true;
    }
};
struct PolyView {  // over caller-provided storage (the wide pass's LDS)
    uint32_t count;
    bool overflow;
    uint32_t cap;
    ClipV* v;
    __device__ void push(const ClipV& x) {
        if (count < cap) v[count++] = x;
        else overflow = true;
    }
};
using Poly = PolyArr<kPolyCap>;

struct CPV {
    V3 position;
    uint32_t clipIndex;
};
template <uint32_t Cap>
struct CPolyArr {
    uint32_t count;
    CPV v[Cap];
};
struct CPolyView {
    uint32_t count;
    CPV* v;
};
using CPoly = CPolyArr<kPolyCap>;

// The workspace of b3BuildFaceContact: the clip polygon, its scratch and the points kept
// within the radius.  FaceWs is the lane's own; WideFaceWs points into LDS.
struct FaceWs {
    Poly poly, scratch;
    CPoly kept;
};
struct WideFaceWs {
    PolyView poly, scratch;
    CPolyView kept;
};
// The lane's own workspace, or the wide pass's.
template <bool Wide>
__device__ __forceinline__ auto select_ws(FaceWs* lane, WideFaceWs* wide) {
    if constexpr (Wide) return wide;
    else return lane;
}

// clip.cpp:33-57
template <class H, class P>
__device__ inline void build_polygon(P& out, const Xf& xf, uint32_t index, const H& h) {
    out.count = 0;
    uint32_t begin = h.face_edge(index);
    uint32_t e = begin;
    uint32_t guard = 0;
    do {
        uint32_t twin = e ^ 1u;
        uint32_t edgeId = twin ^ 1u;  // twin->twin
        uint32_t prevId = h.next(twin);
        out.push(clipv(mul(xf, h.vertex(h.origin(e))), prevId, kNullEdge, edgeId, kNullEdge));
        e = h.next(e);
    } while (e != begin && ++guard < 256);
}

// clip.cpp:99-149
template <class P>
__device__ inline void clip_to_plane(P& out, const P& in, const Plane& pl, uint32_t id) {
    out.count = 0;
    ClipV v1 = in.v[in.count - 1];
    float d1 = distance(v1.position, pl);
    for (uint32_t i = 0; i < in.count; ++i) {
        ClipV v2 = in.v[i];
        float d2 = distance(v2.position, pl);
        if (d1 <= 0.0f && d2 <= 0.0f) {
            out.push(v2);
        } else if (d1 <= 0.0f && d2 > 0.0f) {
            float f = d1 / (d1 - d2);
            V3 pos = v1.position + f * (v2.position - v1.position);
            out.push(clipv(pos, v1.inEdge1, kNullEdge, kNullEdge, id));
        } else if (d1 > 0.0f && d2 <= 0.0f) {
            float f = d1 / (d1 - d2);
            V3 pos = v1.position + f * (v2.position - v1.position);
            out.push(clipv(pos, v1.inEdge1, id, kNullEdge, kNullEdge));
            out.push(v2);
        }
        v1 = v2;
        d1 = d2;
    }
}

// clip.cpp:243-280; returns the clipped polygon in `a` (b is scratch).  The two buffers
// alternate between clip planes; the result is copied into `a` once at the end.
template <class H, class P>
__device__ inline void clip_to_face(P& a, P& b, const Xf& xf, float r, uint32_t index, const H& h) {
    uint32_t begin = h.face_edge(index);
    uint32_t e = begin;
    uint32_t guard = 0;
    P* src = &a;
    P* dst = &b;
    bool overflow = a.overflow;
    do {
        uint32_t edgeId = e;  // (edge->twin)->twin
        Plane pl = edge_side_plane(h, edgeId);
        pl.d += r;
        Plane wp = mul(xf, pl);
        dst->overflow = false;
        clip_to_plane(*dst, *src, wp, edgeId);
        overflow = overflow || dst->overflow;
        P* t = src;
        src = dst;
        dst = t;
        if (src->count == 0) break;
        e = h.next(e);
    } while (e != begin && ++guard < 256);
    if (src != &a) {
        a.count = src->count;
        for (uint32_t i = 0; i < src->count; ++i) a.v[i] = src->v[i];
    }
    a.overflow = overflow;
}

// contact_cluster.cpp:22-26
__device__ __forceinline__ bool is_ccw(V3 A, V3 B, V3 C, V3 N) {
    V3 n = cross(B - A, C - A);
    return dot(n, N) > 0.0f;
}

// contact_cluster.cpp:33-56 (the pivot B is read once per i, before any swap)
template <class CP>
__device__ inline void sort_polygon(CP& p, V3 N) {
    V3 A = p.v[0].position;
    for (uint32_t i = 1; i < p.count; ++i) {
        V3 B = p.v[i].position;
        for (uint32_t j = i + 1; j < p.count; ++j) {
            V3 C = p.v[j].position;
            if (!is_ccw(A, B, C, N)) {
                CPV t = p.v[i];
                p.v[i] = p.v[j];
                p.v[j] = t;
            }
        }
    }
}

// contact_cluster.cpp:58-200: reduces `in` into `out` (at most four points).  The chosen
// points (at most three are chosen before the last pick) are skipped by index.
template <class CIn>
__device__ inline void reduce_polygon(CPolyArr<4>& out, const CIn& in, V3 N, uint32_t initial) {
    out.count = 0;
    if (in.count <= 4) {
        out.count = in.count;
        for (uint32_t i = 0; i < in.count; ++i) out.v[i] = in.v[i];
        sort_polygon(out, N);
        return;
    }
    const uint32_t none = 0xFFFFFFFFu;
    uint32_t c1 = none, c2 = none;
    out.v[out.count++] = in.v[initial];
    {
        V3 A = out.v[0].position;
        uint32_t index = 0;
        float mx = -kMaxFloat;
        for (uint32_t i = 0; i < in.count; ++i) {
            if (i == initial) continue;
            V3 d = in.v[i].position - A;
            float dd = dot(d, d);
            if (dd > mx) {
                mx = dd;
                index = i;
            }
        }
        if (mx < kEpsilon * kEpsilon) return;
        out.v[out.count++] = in.v[index];
        c1 = index;
    }
    {
        V3 A = out.v[0].position, B = out.v[1].position;
        uint32_t index = 0;
        float mx = -kMaxFloat;
        for (uint32_t i = 0; i < in.count; ++i) {
            if (i == initial || i == c1) continue;
            V3 C = in.v[i].position;
            V3 n = cross(B - A, C - A);
            float a2 = fabsf(dot(n, N));
            if (a2 > mx) {
                mx = a2;
                index = i;
            }
        }
        const float kAreaTol = 0.01f;
        if (mx < 2.0f * kAreaTol) return;
        V3 C = in.v[index].position;
        if (!is_ccw(A, B, C, N)) {
            CPV t = out.v[0];
            out.v[0] = out.v[1];
            out.v[1] = t;
        }
        out.v[out.count++] = in.v[index];
        c2 = index;
    }
    {
        V3 A = out.v[0].position, B = out.v[1].position;
        uint32_t index = 0;
        float mn = kMaxFloat;
        for (uint32_t i = 0; i < in.count; ++i) {
            if (i == initial || i == c1 || i == c2) continue;
            V3 D = in.v[i].position;
            V3 n = cross(B - A, D - A);
            float sa2 = dot(n, N);
            if (sa2 < mn) {
                mn = sa2;
                index = i;
            }
        }
        const float kAreaTol = 0.01f;
        if (fabsf(mn) < 2.0f * kAreaTol) return;
        out.v[out.count++] = in.v[index];
    }
    sort_polygon(out, N);
}

// collide_hulls.cpp:26-91, and with `rebuild` the segment check of
// b3RebuildEdgeContact (collide_hulls_cache.cpp:35-111).
template <class H1, class H2>
__device__ inline void build_edge_contact(Contact& m, const Xf& xf1, uint32_t index1, const H1& h1,
                                   const Xf& xf2, uint32_t index2, const H2& h2, bool rebuild) {
    V3 C1 = mul(xf1, h1.centroid());
    V3 P1 = mul(xf1, h1.vertex(h1.origin(index1)));
    V3 Q1 = mul(xf1, h1.vertex(h1.origin(index1 + 1)));
    V3 E1 = Q1 - P1;
    V3 N1 = E1;
    normalize_in_place(N1);
    V3 P2 = mul(xf2, h2.vertex(h2.origin(index2)));
    V3 Q2 = mul(xf2, h2.vertex(h2.origin(index2 + 1)));
    V3 E2 = Q2 - P2;
    V3 N2 = E2;
    normalize_in_place(N2);
    float b = dot(N1, N2);
    float den = 1.0f - b * b;
    if (den == 0.0f) return;
    float inv_den = 1.0f / den;
    V3 E3 = P1 - P2;
    float d = dot(N1, E3);
    float e = dot(N2, E3);
    float s = inv_den * (b * e - d);
    float t = inv_den * (e - b * d);
    V3 c1 = P1 + s * N1;
    V3 c2 = P2 + t * N2;
    if (rebuild) {
        float w2[3], w1[3];
        barycentric(w2, P1, Q1, c2);
        barycentric(w1, P2, Q2, c1);
        if (!(w2[1] > 0.0f && w2[1] <= w2[2] && w1[1] > 0.0f && w1[1] <= w1[2])) return;
    }
    V3 N = cross(E1, E2);
    normalize_in_place(N);
    if (dot(N, P1 - C1) < 0.0f) N = -N;
    m.pointCount = 1;
    m.n[0] = mulT(xf1.R, N);
    m.p1[0] = mulT(xf1, c1);
    m.p2[0] = mulT(xf2, c2);
    m.key[0][0] = index1;      // inEdge1
    m.key[0][1] = index2;      // outEdge1
    m.key[0][2] = index1 + 1;  // inEdge2
    m.key[0][3] = index2 + 1;  // outEdge2
}

// collide_hulls.cpp:93-204, in the workspace ws (FaceWs or WideFaceWs).
template <class H1, class H2, class WS>
__device__ inline void build_face_contact(Contact& m, const Xf& xf1, uint32_t index1, const H1& h1, float r1,
                                   const Xf& xf2, const H2& h2, float r2, bool flip, WS& ws) {
    auto& poly = ws.poly;
    auto& scratch = ws.scratch;
    poly.overflow = false;
    scratch.overflow = false;
    float totalRadius = r1 + r2;
    Plane plane1 = mul(xf1, h1.plane(index1));
    V3 normal1 = mulT(xf2.R, plane1.n);
    // hull.inl:47-61 support face of hull 2 along -normal1
    V3 dn = -normal1;
    uint32_t index2 = 0;
    float mp = dot(dn, h2.plane(0).n);
    for (uint32_t i = 1; i < h2.F; ++i) {
        float q = dot(dn, h2.plane(i).n);
        if (q > mp) {
            mp = q;
            index2 = i;
        }
    }
    build_polygon(poly, xf2, index2, h2);
    clip_to_face(poly, scratch, xf1, totalRadius, index1, h1);
    m.overflow = m.overflow || poly.overflow;
    if (poly.count == 0) return;
    auto& p1 = ws.kept;
    CPolyArr<4> red;
    p1.count = 0;
    uint32_t minIndex = 0;
    float minSep = kMaxFloat;
    for (uint32_t i = 0; i < poly.count; ++i) {
        V3 pos = poly.v[i].position;
        float sep = distance(pos, plane1);
        if (sep <= totalRadius) {
            if (sep < minSep) {
                minIndex = p1.count;
                minSep = sep;
            }
            p1.v[p1.count].position = closest_on_plane(pos, plane1);
            p1.v[p1.count].clipIndex = i;
            ++p1.count;
        }
    }
    if (p1.count == 0) return;
    V3 normal = plane1.n;
    V3 s_normal = flip ? -normal : normal;
    reduce_polygon(red, p1, s_normal, minIndex);
    for (uint32_t i = 0; i < red.count; ++i) {
        const ClipV& v2 = poly.v[red.v[i].clipIndex];
        V3 v1 = closest_on_plane(v2.position, plane1);
        if (flip) {
            // b3MakePair(in2, in1, out2, out1) swaps the hull roles of the key.
            m.n[i] = mulT(xf2.R, s_normal);
            m.p1[i] = mulT(xf2, v2.position);
            m.p2[i] = mulT(xf1, v1);
            m.key[i][0] = v2.inEdge2;
            m.key[i][1] = v2.outEdge2;
            m.key[i][2] = v2.inEdge1;
            m.key[i][3] = v2.outEdge1;
        } else {
            m.n[i] = mulT(xf1.R, normal);
            m.p1[i] = mulT(xf1, v1);
            m.p2[i] = mulT(xf2, v2.position);
            m.key[i][0] = v2.inEdge1;
            m.key[i][1] = v2.outEdge1;
            m.key[i][2] = v2.inEdge2;
            m.key[i][3] = v2.outEdge2;
        }
    }
    m.pointCount = red.count;
}

// The rotation test of the cached face rebuild (collide_hulls_cache.cpp:117-126): the
// relative rotation of the bodies since the last step is small.
__device__ inline bool rebuild_face_allowed(Quat q10, Quat q1, Quat q20, Quat q2) {
    Quat dq0 = qmul(conj(q10), q20);
    Quat dq = qmul(conj(q1), q2);
    Quat q = qmul(conj(dq0), dq);
    const float kTol = 0.995f;
    return fabsf(q.w) > kTol;
}

// collide_hulls_cache.cpp:113-153
template <class H1, class H2, class WS>
__device__ inline void rebuild_face_contact(Contact& m, const Xf& xf1, uint32_t index1, const H1& h1, float r1,
                                     const Xf& xf2, const H2& h2, float r2, bool flip,
                                     Quat q10, Quat q1, Quat q20, Quat q2, WS& ws) {
    if (rebuild_face_allowed(q10, q1, q20, q2)) build_face_contact(m, xf1, index1, h1, r1, xf2, h2, r2, flip, ws);
}

// b3FeatureCache::ReadFace (sat.cpp:209-222)
template <class H1, class H2>
__device__ inline uint32_t read_face(uint32_t index1, const Xf& xf1, const H1& h1, const Xf& xf2, const H2& h2,
                              float totalRadius) {
    Xf xf = mulT(xf2, xf1);
    Plane pl = mul(xf, h1.plane(index1));
    V3 d = -pl.n;
    uint32_t mi = 0;
    float mp = dot(d, h2.vertex(0));
    for (uint32_t k = 1; k < h2.V; ++k) {
        float q = dot(d, h2.vertex(k));
        if (q > mp) {
            mp = q;
            mi = k;
        }
    }
    float s = distance(h2.vertex(mi), pl);
    return s > totalRadius ? kCacheSeparation : kCacheOverlap;
}

// b3FeatureCache::ReadEdge (sat.cpp:224-278)
template <class H1, class H2>
__device__ inline uint32_t read_edge(uint32_t i, uint32_t j, const Xf& xf1, const H1& h1, const Xf& xf2,
                              const H2& h2, float totalRadius) {
    Xf xf = mulT(xf2, xf1);
    V3 C1 = mul(xf, h1.centroid());
    V3 P1 = mul(xf, h1.vertex(h1.origin(i)));
    V3 Q1 = mul(xf, h1.vertex(h1.origin(i + 1)));
    V3 E1 = Q1 - P1;
    V3 U1 = mul(xf.R, h1.plane(h1.face(i)).n);
    V3 V1 = mul(xf.R, h1.plane(h1.face(i + 1)).n);
    V3 P2 = h2.vertex(h2.origin(j));
    V3 Q2 = h2.vertex(h2.origin(j + 1));
    V3 E2 = Q2 - P2;
    V3 U2 = h2.plane(h2.face(j)).n;
    V3 V2 = h2.plane(h2.face(j + 1)).n;
    V3 nE1 = -E1, nE2 = -E2, nU2 = -U2, nV2 = -V2;
    float ADC = dot(U1, nE2);
    float BDC = dot(V1, nE2);
    float CBA = dot(nU2, nE1);
    float DBA = dot(nV2, nE1);
    if ((CBA * DBA < 0.0f) && (ADC * BDC < 0.0f) && (CBA * BDC > 0.0f)) {
        float s = project_edges(P1, E1, length(E1), P1 - C1, P2, E2, length(E2));
        return s > totalRadius ? kCacheSeparation : kCacheOverlap;
    }
    return kCacheEmpty;
}

// ------------------------------------------------------------------ sequential SAT

// b3QueryFaceSeparation (sat.cpp:33-58) for one lane.
template <class H1, class H2>
__device__ inline void query_face_seq(const Xf& xf1, const H1& h1, const Xf& xf2, const H2& h2, uint32_t& index,
                                      float& separation) {
    Xf xf = mulT(xf2, xf1);
    index = 0;
    separation = -kMaxFloat;
    // Records are read four at a time (the loads in flight together) and used in index order.
    for (uint32_t i0 = 0; i0 < h1.F; i0 += 4u) {
        Plane raw[4];
#pragma unroll
        for (uint32_t t = 0; t < 4u; ++t) raw[t] = h1.plane(i0 + t < h1.F ? i0 + t : i0);
#pragma unroll
        for (uint32_t t = 0; t < 4u; ++t) {
            const uint32_t i = i0 + t;
            if (i >= h1.F) break;
            Plane pl = mul(xf, raw[t]);
            V3 d = -pl.n;
            uint32_t mi = 0;
            float mp = dot(d, h2.vertex(0));
            uint32_t k = 1;
            for (; k + 4u <= h2.V; k += 4u) {
                V3 w[4];
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) w[u] = h2.vertex(k + u);
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) {
                    float q = dot(d, w[u]);
                    if (q > mp) {
                        mp = q;
                        mi = k + u;
                    }
                }
            }
            for (; k < h2.V; ++k) {
                float q = dot(d, h2.vertex(k));
                if (q > mp) {
                    mp = q;
                    mi = k;
                }
            }
            float s = distance(h2.vertex(mi), pl);
            if (s > separation) {
                index = i;
                separation = s;
            }
        }
    }
}

// b3QueryEdgeSeparation (sat.cpp:110-173) for one lane.
template <class H1, class H2>
__device__ inline void query_edge_seq(const Xf& xf1, const H1& h1, const Xf& xf2, const H2& h2, uint32_t& index1,
                                      uint32_t& index2, float& separation) {
    Xf xf = mulT(xf2, xf1);
    V3 C1 = mul(xf, h1.centroid());
    index1 = 0;
    index2 = 0;
    separation = -kMaxFloat;
    // Hull 1's edge records are read two edges at a time (their loads in flight together).
    V3 nextP = h1.vertex(h1.origin(0)), nextQ = h1.vertex(h1.origin(1));
    V3 nextU = h1.plane(h1.face(0)).n, nextV = h1.plane(h1.face(1)).n;
    for (uint32_t i = 0; i < h1.HE; i += 2) {
        const V3 P1raw = nextP, Q1raw = nextQ, U1raw = nextU, V1raw = nextV;
        if (i + 2 < h1.HE) {
            nextP = h1.vertex(h1.origin(i + 2));
            nextQ = h1.vertex(h1.origin(i + 3));
            nextU = h1.plane(h1.face(i + 2)).n;
            nextV = h1.plane(h1.face(i + 3)).n;
        }
        V3 P1 = mul(xf, P1raw);
        V3 Q1 = mul(xf, Q1raw);
        V3 E1 = Q1 - P1;
        V3 U1 = mul(xf.R, U1raw);
        V3 V1 = mul(xf.R, V1raw);
        V3 nE1 = -E1;
        float L1 = length(E1);
        V3 D1 = P1 - C1;
        for (uint32_t j = 0; j < h2.HE; j += 2) {
            V3 P2 = h2.vertex(h2.origin(j));
            V3 Q2 = h2.vertex(h2.origin(j + 1));
            V3 E2 = Q2 - P2;
            V3 nE2 = -E2, nU2 = -h2.plane(h2.face(j)).n, nV2 = -h2.plane(h2.face(j + 1)).n;
            float ADC = dot(U1, nE2);
            float BDC = dot(V1, nE2);
            float CBA = dot(nU2, nE1);
            float DBA = dot(nV2, nE1);
            if ((CBA * DBA < 0.0f) && (ADC * BDC < 0.0f) && (CBA * BDC > 0.0f)) {
                float s = project_edges(P1, E1, L1, D1, P2, E2, length(E2));
                if (s > separation) {
                    separation = s;
                    index1 = i;
                    index2 = j;
                }
            }
        }
    }
}

__device__ inline void set_feature_cache(FeatureCache* c, uint32_t state, uint32_t type, uint32_t i1, uint32_t i2) {
    if (!c) return;
    c->state = state;
    c->type = type;
    c->index1 = i1;
    c->index2 = i2;
}

// b3CollideCache (collide_hulls_cache.cpp:155-275) for one lane, face contacts built in ws.
// A clip polygon that outgrows ws stops it at once (m.overflow set, the cache untouched
// since), for the wide pass to run the pair again.
template <class H1, class H2, class WS>
__device__ inline void collide_cache_seq(Contact& m, const Xf& xf1, const H1& h1, float r1, const Xf& xf2,
                                         const H2& h2, float r2, FeatureCache* fc, WS& ws) {
    const float totalRadius = r1 + r2;
    uint32_t f1i, f2i, ei, ej;
    float f1s, f2s, es;
    query_face_seq(xf1, h1, xf2, h2, f1i, f1s);
    if (f1s > totalRadius) {
        set_feature_cache(fc, kCacheSeparation, kFeatureFace1, f1i, f1i);
        return;
    }
    query_face_seq(xf2, h2, xf1, h1, f2i, f2s);
    if (f2s > totalRadius) {
        set_feature_cache(fc, kCacheSeparation, kFeatureFace2, f2i, f2i);
        return;
    }
    query_edge_seq(xf1, h1, xf2, h2, ei, ej, es);
    if (es > totalRadius) {
        set_feature_cache(fc, kCacheSeparation, kFeatureEdge1, ei, ej);
        return;
    }
    const float kTol = 0.1f * kLinearSlop;
    const float fmax12 = f1s > f2s ? f1s : f2s;
    if (es > fmax12 + kTol) {
        build_edge_contact(m, xf1, ei, h1, xf2, ej, h2, false);
        if (m.pointCount > 0) {
            set_feature_cache(fc, kCacheOverlap, kFeatureEdge1, ei, ej);
            return;
        }
    } else if (f1s + kTol > f2s) {
        build_face_contact(m, xf1, f1i, h1, r1, xf2, h2, r2, false, ws);
        if (m.overflow) return;
        if (m.pointCount > 0) {
            set_feature_cache(fc, kCacheOverlap, kFeatureFace1, f1i, f1i);
            return;
        }
    } else {
        build_face_contact(m, xf2, f2i, h2, r2, xf1, h1, r1, true, ws);
        if (m.overflow) return;
        if (m.pointCount > 0) {
            set_feature_cache(fc, kCacheOverlap, kFeatureFace2, f2i, f2i);
            return;
        }
    }
    if (es > fmax12) {
        build_edge_contact(m, xf1, ei, h1, xf2, ej, h2, false);
        if (m.pointCount > 0) {
            set_feature_cache(fc, kCacheOverlap, kFeatureEdge1, ei, ej);
            return;
        }
    } else if (f1s > f2s) {
        build_face_contact(m, xf1, f1i, h1, r1, xf2, h2, r2, false, ws);
        if (m.overflow) return;
        if (m.pointCount > 0) {
            set_feature_cache(fc, kCacheOverlap, kFeatureFace1, f1i, f1i);
            return;
        }
    } else {
        build_face_contact(m, xf2, f2i, h2, r2, xf1, h1, r1, true, ws);
        if (m.overflow) return;
        if (m.pointCount > 0) {
            set_feature_cache(fc, kCacheOverlap, kFeatureFace2, f2i, f2i);
            return;
        }
    }
    build_edge_contact(m, xf1, ei, h1, xf2, ej, h2, false);
    set_feature_cache(fc, kCacheOverlap, kFeatureEdge1, ei, ej);
}

// b3CollideHulls with a feature cache (collide_hulls_cache.cpp:279-344) for one lane: the
// cached feature is re-tested (ReadState, sat.cpp:177-278); a separation that still holds,
// or an overlap whose contact can be rebuilt (with the bodies' sweeps for the face rebuild),
// is a hit (b3_convexCacheHits); otherwise the cache is flushed and b3CollideCache runs.
// A cached feature outside the hulls cannot come from this pair and is treated as a miss.
// Face contacts are built in ws; a polygon that outgrows it stops the pair at once
// (m.overflow set, the cache as it was at that point) for the wide pass.
template <class H1, class H2, class WS>
__device__ inline void collide_hulls_cached_seq(Contact& m, const Xf& xf1, const H1& h1, float r1, const Xf& xf2,
                                                const H2& h2, float r2, FeatureCache* fc, Quat q10, Quat q1,
                                                Quat q20, Quat q2, uint32_t& hit, WS& ws) {
    hit = 0;
    const float totalRadius = r1 + r2;
    const uint32_t state0 = fc->state, type = fc->type, i1 = fc->index1, i2 = fc->index2;
    const bool inRange = (type == kFeatureEdge1 && i1 + 1 < h1.HE && i2 + 1 < h2.HE && !(i1 & 1u) && !(i2 & 1u)) ||
                         (type == kFeatureFace1 && i1 < h1.F) || (type == kFeatureFace2 && i1 < h2.F);
    if (inRange && state0 < kCacheEmpty) {
        uint32_t state1;
        switch (type) {
            case kFeatureEdge1: state1 = read_edge(i1, i2, xf1, h1, xf2, h2, totalRadius); break;
            case kFeatureFace1: state1 = read_face(i1, xf1, h1, xf2, h2, totalRadius); break;
            default: state1 = read_face(i1, xf2, h2, xf1, h1, totalRadius); break;
        }
        if (state0 == kCacheSeparation && state1 == kCacheSeparation) {
            hit = 1;
            return;
        }
        if (state0 == kCacheOverlap && state1 == kCacheOverlap) {
            switch (type) {
                case kFeatureEdge1: build_edge_contact(m, xf1, i1, h1, xf2, i2, h2, true); break;
                case kFeatureFace1: rebuild_face_contact(m, xf1, i1, h1, r1, xf2, h2, r2, false, q10, q1, q20, q2, ws); break;
                default: rebuild_face_contact(m, xf2, i1, h2, r2, xf1, h1, r1, true, q20, q2, q10, q1, ws); break;
            }
            if (m.overflow) return;
            if (m.pointCount > 0) {
                hit = 1;
                return;
            }
        }
    }
    fc->state = kCacheEmpty;
    collide_cache_seq(m, xf1, h1, r1, xf2, h2, r2, fc, ws);
}

}  // namespace b3amd

#endif
