// Lane-per-pair sphere/capsule routines, templated over the hull accessor: GJK distance
// (gjk.cpp:32-737) and the closed-form and GJK-based collide functions
// (collide_spheres.cpp, collide_sphere_capsule.cpp, collide_capsules.cpp,
// collide_sphere_hull.cpp, collide_capsule_hull.cpp).
#ifndef B3AMD_CONVEX_BUILD_H
#define B3AMD_CONVEX_BUILD_H

#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "device_common.h"
#include "device_hull.h"
#include "kernels.h"

namespace b3amd {

// ------------------------------------------------------------------ GJK

// A GJK proxy (gjk_proxy.h:26-44): up to three local points, or the vertices of a hull.
// Every index into the local points is a selection, so that a proxy stays in registers.
struct Proxy {
    V3 pts[3];
    uint32_t count;
    const float4* hullVerts;  // non-null for a hull proxy
    __device__ V3 vertex(uint32_t i) const {
        if (hullVerts) {
            float4 q = hullVerts[i];
            return V3{q.x, q.y, q.z};
        }
        return i == 0u ? pts[0] : (i == 1u ? pts[1] : pts[2]);
    }
    // gjk_proxy.h:57-71.  A hull's vertices are read four at a time (the loads in flight
    // together), then compared in index order.
    __device__ uint32_t support(V3 d) const {
        uint32_t maxIndex = 0;
        float maxProjection = dot(d, vertex(0));
        if (!hullVerts) {
#pragma unroll
            for (uint32_t i = 1; i < 3u; ++i) {
                if (i >= count) break;
                float q = dot(d, pts[i]);
                if (q > maxProjection) {
                    maxIndex = i;
                    maxProjection = q;
                }
            }
            return maxIndex;
        }
        uint32_t i = 1;
        for (; i + 4u <= count; i += 4u) {
            float4 w[4];
#pragma unroll
            for (uint32_t k = 0; k < 4u; ++k) w[k] = hullVerts[i + k];
#pragma unroll
            for (uint32_t k = 0; k < 4u; ++k) {
                float q = dot(d, V3{w[k].x, w[k].y, w[k].z});
                if (q > maxProjection) {
                    maxIndex = i + k;
                    maxProjection = q;
                }
            }
        }
        for (; i < count; ++i) {
            float q = dot(d, vertex(i));
            if (q > maxProjection) {
                maxIndex = i;
                maxProjection = q;
            }
        }
        return maxIndex;
    }
};

struct SVert {
    V3 point1, point2, point;
    float weight;
    uint32_t index1, index2;
};

struct Simplex {
    SVert v[4];
    uint32_t count;
};

// gjk.cpp:32-45
__device__ __forceinline__ void bary2(float out[3], V3 A, V3 B, V3 Q) {
    V3 AB = B - A, QA = A - Q, QB = B - Q;
    out[0] = dot(QB, AB);
    out[1] = -dot(QA, AB);
    out[2] = out[0] + out[1];
}

// gjk.cpp:47-72
__device__ __forceinline__ void bary3(float out[4], V3 A, V3 B, V3 C, V3 Q) {
    V3 AB = B - A, AC = C - A;
    V3 QA = A - Q, QB = B - Q, QC = C - Q;
    V3 QB_x_QC = cross(QB, QC), QC_x_QA = cross(QC, QA), QA_x_QB = cross(QA, QB);
    V3 AB_x_AC = cross(AB, AC);
    out[0] = dot(QB_x_QC, AB_x_AC);
    out[1] = dot(QC_x_QA, AB_x_AC);
    out[2] = dot(QA_x_QB, AB_x_AC);
    out[3] = out[0] + out[1] + out[2];
}

__device__ __forceinline__ float det3(V3 a, V3 b, V3 c) { return dot(a, cross(b, c)); }  // vec3.h:170-173
__device__ __forceinline__ float sign_of(float x) { return (float)((0.0f < x) - (x < 0.0f)); }  // math.h:88-92

// gjk.cpp:74-99
__device__ __forceinline__ void bary4(float out[5], V3 A, V3 B, V3 C, V3 D, V3 Q) {
    V3 AB = B - A, AC = C - A, AD = D - A;
    V3 QA = A - Q, QB = B - Q, QC = C - Q, QD = D - Q;
    float divisor = det3(AB, AC, AD);
    float sign = sign_of(divisor);
    out[0] = sign * det3(QB, QC, QD);
    out[1] = sign * det3(QA, QD, QC);
    out[2] = sign * det3(QA, QB, QD);
    out[3] = sign * det3(QA, QC, QB);
    out[4] = sign * divisor;
}

__device__ __forceinline__ void set1(Simplex& s, const SVert& a) {
    s.count = 1;
    s.v[0] = a;
    s.v[0].weight = 1.0f;
}
__device__ __forceinline__ void set2(Simplex& s, const SVert& a, const SVert& b, float w0, float w1, float div) {
    float d = 1.0f / div;
    s.count = 2;
    s.v[0] = a;
    s.v[0].weight = d * w0;
    s.v[1] = b;
    s.v[1].weight = d * w1;
}
__device__ __forceinline__ void set3(Simplex& s, const SVert& a, const SVert& b, const SVert& c, float w0, float w1,
                                     float w2, float div) {
    float d = 1.0f / div;
    s.count = 3;
    s.v[0] = a;
    s.v[0].weight = d * w0;
    s.v[1] = b;
    s.v[1].weight = d * w1;
    s.v[2] = c;
    s.v[2].weight = d * w2;
}

// gjk.cpp:201-234
__device__ inline void solve2(Simplex& s) {
    const V3 Q{0.0f, 0.0f, 0.0f};
    SVert A = s.v[0], B = s.v[1];
    float wAB[3];
    bary2(wAB, A.point, B.point, Q);
    if (wAB[1] <= 0.0f) { set1(s, A); return; }
    if (wAB[0] <= 0.0f) { set1(s, B); return; }
    set2(s, A, B, wAB[0], wAB[1], wAB[2]);
}

// gjk.cpp:236-332
__device__ inline void solve3(Simplex& s) {
    const V3 Q{0.0f, 0.0f, 0.0f};
    SVert A = s.v[0], B = s.v[1], C = s.v[2];
    float wAB[3], wBC[3], wCA[3];
    bary2(wAB, A.point, B.point, Q);
    bary2(wBC, B.point, C.point, Q);
    bary2(wCA, C.point, A.point, Q);
    if (wAB[1] <= 0.0f && wCA[0] <= 0.0f) { set1(s, A); return; }
    if (wAB[0] <= 0.0f && wBC[1] <= 0.0f) { set1(s, B); return; }
    if (wBC[0] <= 0.0f && wCA[1] <= 0.0f) { set1(s, C); return; }
    float wABC[4];
    bary3(wABC, A.point, B.point, C.point, Q);
    float area = wABC[3];
    if (wAB[0] > 0.0f && wAB[1] > 0.0f && area * wABC[2] <= 0.0f) { set2(s, A, B, wAB[0], wAB[1], wAB[2]); return; }
    if (wBC[0] > 0.0f && wBC[1] > 0.0f && area * wABC[0] <= 0.0f) { set2(s, B, C, wBC[0], wBC[1], wBC[2]); return; }
    if (wCA[0] > 0.0f && wCA[1] > 0.0f && area * wABC[1] <= 0.0f) { set2(s, C, A, wCA[0], wCA[1], wCA[2]); return; }
    float divisor = wABC[3];
    if (divisor <= 0.0f) return;  // give up
    set3(s, A, B, C, wABC[0], wABC[1], wABC[2], divisor);
}

// gjk.cpp:334-575
__device__ inline void solve4(Simplex& s) {
    const V3 Q{0.0f, 0.0f, 0.0f};
    SVert A = s.v[0], B = s.v[1], C = s.v[2], D = s.v[3];
    float wAB[3], wAC[3], wAD[3], wBC[3], wCD[3], wDB[3];
    bary2(wAB, A.point, B.point, Q);
    bary2(wBC, B.point, C.point, Q);
    bary2(wAC, A.point, C.point, Q);
    bary2(wAD, A.point, D.point, Q);
    bary2(wCD, C.point, D.point, Q);
    bary2(wDB, D.point, B.point, Q);
    if (wAB[1] <= 0.0f && wAC[1] <= 0.0f && wAD[1] <= 0.0f) { set1(s, A); return; }
    if (wAB[0] <= 0.0f && wDB[0] <= 0.0f && wBC[1] <= 0.0f) { set1(s, B); return; }
    if (wAC[0] <= 0.0f && wBC[0] <= 0.0f && wCD[1] <= 0.0f) { set1(s, C); return; }
    if (wAD[0] <= 0.0f && wCD[0] <= 0.0f && wDB[1] <= 0.0f) { set1(s, D); return; }
    float wACB[4], wABD[4], wADC[4], wBCD[4];
    bary3(wACB, A.point, C.point, B.point, Q);
    bary3(wABD, A.point, B.point, D.point, Q);
    bary3(wADC, A.point, D.point, C.point, Q);
    bary3(wBCD, B.point, C.point, D.point, Q);
    if (wABD[2] <= 0.0f && wACB[1] <= 0.0f && wAB[0] > 0.0f && wAB[1] > 0.0f) { set2(s, A, B, wAB[0], wAB[1], wAB[2]); return; }
    if (wACB[2] <= 0.0f && wADC[1] <= 0.0f && wAC[0] > 0.0f && wAC[1] > 0.0f) { set2(s, A, C, wAC[0], wAC[1], wAC[2]); return; }
    if (wADC[2] <= 0.0f && wABD[1] <= 0.0f && wAD[0] > 0.0f && wAD[1] > 0.0f) { set2(s, A, D, wAD[0], wAD[1], wAD[2]); return; }
    if (wACB[0] <= 0.0f && wBCD[2] <= 0.0f && wBC[0] > 0.0f && wBC[1] > 0.0f) { set2(s, B, C, wBC[0], wBC[1], wBC[2]); return; }
    if (wADC[0] <= 0.0f && wBCD[0] <= 0.0f && wCD[0] > 0.0f && wCD[1] > 0.0f) { set2(s, C, D, wCD[0], wCD[1], wCD[2]); return; }
    if (wABD[0] <= 0.0f && wBCD[1] <= 0.0f && wDB[0] > 0.0f && wDB[1] > 0.0f) { set2(s, D, B, wDB[0], wDB[1], wDB[2]); return; }
    float wABCD[5];
    bary4(wABCD, A.point, B.point, C.point, D.point, Q);
    if (wABCD[3] <= 0.0f && wACB[0] > 0.0f && wACB[1] > 0.0f && wACB[2] > 0.0f) {
        set3(s, A, C, B, wACB[0], wACB[1], wACB[2], wACB[0] + wACB[1] + wACB[2]);
        return;
    }
    if (wABCD[2] <= 0.0f && wABD[0] > 0.0f && wABD[1] > 0.0f && wABD[2] > 0.0f) {
        set3(s, A, B, D, wABD[0], wABD[1], wABD[2], wABD[0] + wABD[1] + wABD[2]);
        return;
    }
    if (wABCD[1] <= 0.0f && wADC[0] > 0.0f && wADC[1] > 0.0f && wADC[2] > 0.0f) {
        set3(s, A, D, C, wADC[0], wADC[1], wADC[2], wADC[0] + wADC[1] + wADC[2]);
        return;
    }
    if (wABCD[0] <= 0.0f && wBCD[0] > 0.0f && wBCD[1] > 0.0f && wBCD[2] > 0.0f) {
        set3(s, B, C, D, wBCD[0], wBCD[1], wBCD[2], wBCD[0] + wBCD[1] + wBCD[2]);
        return;
    }
    float divisor = wABCD[0] + wABCD[1] + wABCD[2] + wABCD[3];
    if (divisor <= 0.0f) return;  // give up
    float d = 1.0f / divisor;
    s.count = 4;
    s.v[0].weight = d * wABCD[0];
    s.v[1].weight = d * wABCD[1];
    s.v[2].weight = d * wABCD[2];
    s.v[3].weight = d * wABCD[3];
}

struct GJKOut {
    V3 point1, point2;
    float distance;
    uint32_t iterations;  // support calls (gjk.cpp:671-673)
};

// b3Simplex::GetMetric (gjk.cpp:806-841): 0 for one vertex, the segment length, the
// triangle's doubled area, the tetrahedron's |det| (b3Det = dot(a, cross(b, c)), b3Sign).
__device__ inline float simplex_metric(const Simplex& s) {
    if (s.count == 2) return length(s.v[0].point - s.v[1].point);
    if (s.count == 3) return length(cross(s.v[1].point - s.v[0].point, s.v[2].point - s.v[0].point));
    if (s.count == 4) {
        const V3 E1 = s.v[1].point - s.v[0].point, E2 = s.v[2].point - s.v[0].point;
        const V3 E3 = s.v[3].point - s.v[0].point;
        const float det = dot(E1, cross(E2, E3));
        const float sign = (float)(0.0f < det) - (float)(det < 0.0f);
        return sign * det;
    }
    return 0.0f;
}

// Byte k (0..19) of a b3SimplexCache record held as five words: metric (0-3), iterations
// (4-7), count (8-9, u16), index1[4] (10-13), index2[4] (14-17), padding.
__device__ __forceinline__ uint32_t cache_byte(const uint32_t* w, uint32_t k) { return (w[k >> 2] >> (8u * (k & 3u))) & 0xFFu; }
__device__ __forceinline__ void set_cache_byte(uint32_t* w, uint32_t k, uint32_t b) {
    const uint32_t sh = 8u * (k & 3u);
    w[k >> 2] = (w[k >> 2] & ~(0xFFu << sh)) | ((b & 0xFFu) << sh);
}

// b3GJK (gjk.cpp:579-737) with b3Simplex::ReadCache / WriteCache (gjk.cpp:743-804) on a
// b3SimplexCache held as five words (null: the no-cache wrapper, gjk.cpp:845-851: an empty
// cache, not written back).  A vertex added by a support step starts with weight 0 (the
// reference leaves it unset; only the degenerate give-up returns could observe it;
// ReadCache zeroes the cached vertices' weights itself).  A cache count above 4 or a cached
// index beyond a proxy's vertices (not a cache the reference could have written for these
// shapes) is read as empty.  cacheHit: ReadCache kept a simplex of two or more vertices
// (b3_gjkCacheHits).
__device__ inline GJKOut gjk_cached(const Xf& xf1, const Proxy& p1, const Xf& xf2, const Proxy& p2, uint32_t* cache,
                                    bool& cacheHit) {
    // Every loop over the simplex's vertices runs over all four slots with a bound test, so
    // that the simplex stays in registers.
    Simplex s;
    cacheHit = false;
    s.count = 0;
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) s.v[i].index1 = s.v[i].index2 = 0u;
    if (cache) {
        s.count = cache[2] & 0xFFu;  // (u8)count
        if (s.count > 4u) s.count = 0;
        bool inRange = true;
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i)
            if (i < s.count && (cache_byte(cache, 10u + i) >= p1.count || cache_byte(cache, 14u + i) >= p2.count))
                inRange = false;
        if (!inRange) s.count = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) {
            if (i >= s.count) break;
            SVert& v = s.v[i];
            v.index1 = cache_byte(cache, 10u + i);
            v.index2 = cache_byte(cache, 14u + i);
            v.point1 = mul(xf1, p1.vertex(v.index1));
            v.point2 = mul(xf2, p2.vertex(v.index2));
            v.point = v.point2 - v.point1;
            v.weight = 0.0f;
        }
        if (s.count > 1u) {
            const float metric1 = __uint_as_float(cache[0]);
            const float metric2 = simplex_metric(s);
            if (metric2 < 0.5f * metric1 || 2.0f * metric1 < metric2 || metric2 < kEpsilon) s.count = 0;  // flush
            else cacheHit = true;
        }
    }
    if (s.count == 0) {
        s.v[0].point1 = mul(xf1, p1.vertex(0));
        s.v[0].point2 = mul(xf2, p2.vertex(0));
        s.v[0].point = s.v[0].point2 - s.v[0].point1;
        s.v[0].weight = 1.0f;
        s.v[0].index1 = 0;
        s.v[0].index2 = 0;
        s.count = 1;
    }
    uint32_t save1[4], save2[4];
    uint32_t iter = 0;
    while (iter < 20) {
        const uint32_t saveCount = s.count;
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) {
            save1[i] = s.v[i].index1;
            save2[i] = s.v[i].index2;
        }
        if (s.count == 2) solve2(s);
        else if (s.count == 3) solve3(s);
        else if (s.count == 4) solve4(s);
        if (s.count == 4) break;
        // Search direction toward the origin (gjk.cpp:101-146).
        V3 d;
        if (s.count == 1) {
            d = V3{0.0f, 0.0f, 0.0f} - s.v[0].point;
        } else if (s.count == 2) {
            V3 A = s.v[0].point, B = s.v[1].point;
            V3 AB = B - A, AQ = V3{0.0f, 0.0f, 0.0f} - A;
            d = cross(cross(AB, AQ), AB);
        } else {
            V3 A = s.v[0].point, B = s.v[1].point, C = s.v[2].point;
            V3 AB = B - A, AC = C - A, AQ = V3{0.0f, 0.0f, 0.0f} - A;
            V3 n = cross(AB, AC);
            d = dot(n, AQ) > 0.0f ? n : -n;
        }
        if (dot(d, d) < kEpsilon * kEpsilon) break;
        SVert vx;
        vx.index1 = p1.support(mulT(xf1.R, -d));
        vx.point1 = mul(xf1, p1.vertex(vx.index1));
        vx.index2 = p2.support(mulT(xf2.R, d));
        vx.point2 = mul(xf2, p2.vertex(vx.index2));
        vx.point = vx.point2 - vx.point1;
        vx.weight = 0.0f;
        // The new vertex takes slot count (below four here), also when it repeats a saved one
        // and the search stops: the output below reads only the first count slots.
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i)
            if (i == s.count) s.v[i] = vx;
        ++iter;
        bool dup = false;
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i)
            if (i < saveCount && vx.index1 == save1[i] && vx.index2 == save2[i]) dup = true;
        if (dup) break;
        ++s.count;
    }
    // gjk.cpp:168-198
    GJKOut o;
    o.iterations = iter;
    const SVert* v = s.v;
    if (s.count == 1) {
        o.point1 = v[0].point1;
        o.point2 = v[0].point2;
    } else if (s.count == 2) {
        o.point1 = v[0].weight * v[0].point1 + v[1].weight * v[1].point1;
        o.point2 = v[0].weight * v[0].point2 + v[1].weight * v[1].point2;
    } else if (s.count == 3) {
        o.point1 = (v[0].weight * v[0].point1 + v[1].weight * v[1].point1) + v[2].weight * v[2].point1;
        o.point2 = (v[0].weight * v[0].point2 + v[1].weight * v[1].point2) + v[2].weight * v[2].point2;
    } else {
        o.point1 = ((v[0].weight * v[0].point1 + v[1].weight * v[1].point1) + v[2].weight * v[2].point1) +
                   v[3].weight * v[3].point1;
        o.point2 = o.point1;
    }
    o.distance = length(o.point1 - o.point2);
    if (cache) {  // WriteCache: metric, count and the used indices; the rest of the record kept
        cache[0] = __float_as_uint(simplex_metric(s));
        cache[2] = (cache[2] & 0xFFFF0000u) | s.count;
#pragma unroll
        for (uint32_t i = 0; i < 4u; ++i) {
            if (i >= s.count) break;
            set_cache_byte(cache, 10u + i, s.v[i].index1);
            set_cache_byte(cache, 14u + i, s.v[i].index2);
        }
    }
    return o;
}

__device__ inline GJKOut gjk(const Xf& xf1, const Proxy& p1, const Xf& xf2, const Proxy& p2) {
    bool hit;
    return gjk_cached(xf1, p1, xf2, p2, nullptr, hit);
}

// b3TestOverlap (collide.cpp:84-95): b3GJK's applyRadius branch (gjk.cpp:709-733) reduced to
// the distance (the witness points it moves do not enter the result), then the 10 FLT_EPSILON
// tolerance.
__device__ __forceinline__ bool overlap_of(const GJKOut& g, float r1, float r2) {
    float distance = g.distance;
    if (g.distance > r1 + r2 && g.distance > kEpsilon) distance -= r1 + r2;
    else distance = 0.0f;
    return distance <= 10.0f * kEpsilon;
}

// ------------------------------------------------------------------ closed forms

__device__ __forceinline__ void zero_key(Contact& m, int i) {
    m.key[i][0] = m.key[i][1] = m.key[i][2] = m.key[i][3] = 0u;
}

__device__ __forceinline__ V3 divv(V3 v, float s) { return V3{v.x / s, v.y / s, v.z / s}; }  // vec3.h:152-155

// collide_spheres.cpp:24-56
__device__ inline void collide_spheres(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2, const ShapeRec& s2) {
    V3 c1 = mul(xf1, V3{s1.v[0], s1.v[1], s1.v[2]});
    V3 c2 = mul(xf2, V3{s2.v[0], s2.v[1], s2.v[2]});
    V3 d = c2 - c1;
    float dd = dot(d, d);
    float totalRadius = s1.radius + s2.radius;
    if (dd > totalRadius * totalRadius) return;
    float distance = length(d);
    V3 normal{0.0f, 1.0f, 0.0f};
    if (distance > kEpsilon) normal = divv(d, distance);
    m.pointCount = 1;
    m.n[0] = mulT(xf1.R, normal);
    m.p1[0] = V3{s1.v[0], s1.v[1], s1.v[2]};
    m.p2[0] = V3{s2.v[0], s2.v[1], s2.v[2]};
    zero_key(m, 0);
}

// collide_sphere_capsule.cpp:24-133
__device__ inline void collide_sphere_capsule(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2,
                                       const ShapeRec& s2) {
    const V3 c0{s2.v[0], s2.v[1], s2.v[2]}, c1{s2.v[3], s2.v[4], s2.v[5]};
    V3 Q = mul(xf1, V3{s1.v[0], s1.v[1], s1.v[2]});
    V3 A = mul(xf2, c0);
    V3 B = mul(xf2, c1);
    V3 AB = B - A;
    float u = dot(B - Q, AB);
    float v = dot(Q - A, AB);
    float radius = s1.radius + s2.radius;
    if (v <= 0.0f || u <= 0.0f) {
        V3 P = v <= 0.0f ? A : B;
        V3 d = P - Q;
        float dd = dot(d, d);
        if (dd > radius * radius) return;
        V3 n{0.0f, 1.0f, 0.0f};
        float len = length(d);
        if (len > kEpsilon) n = divv(d, len);
        m.pointCount = 1;
        m.n[0] = mulT(xf1.R, n);
        m.p1[0] = V3{s1.v[0], s1.v[1], s1.v[2]};
        m.p2[0] = v <= 0.0f ? c0 : c1;
        zero_key(m, 0);
        return;
    }
    float s = dot(AB, AB);
    V3 P;
    if (s < kLinearSlop * kLinearSlop) P = A;
    else P = divv(u * A + v * B, s);
    V3 d = P - Q;
    float dd = dot(d, d);
    if (dd > radius * radius) return;
    V3 QA = A - Q;
    V3 e = cross(AB, QA);
    V3 n = cross(AB, e);
    if (dot(n, QA) < 0.0f) n = -n;
    normalize_in_place(n);
    m.pointCount = 1;
    m.n[0] = mulT(xf1.R, n);
    m.p1[0] = V3{s1.v[0], s1.v[1], s1.v[2]};
    m.p2[0] = mulT(xf2, P);
    zero_key(m, 0);
}

// collide_capsules.cpp:27-57
__device__ inline V3 closest_on_segment(V3 Q, V3 A, V3 B) {
    V3 AB = B - A;
    float u = dot(B - Q, AB);
    float v = dot(Q - A, AB);
    if (v <= 0.0f) return A;
    if (u <= 0.0f) return B;
    float w = dot(AB, AB);
    if (w <= kLinearSlop * kLinearSlop) return A;
    float den = 1.0f / w;
    return den * (u * A + v * B);
}

// collide_capsules.cpp:59-152
__device__ inline void closest_points_segments(V3& C1, V3& C2, V3 P1, V3 Q1, V3 P2, V3 Q2) {
    V3 E1 = Q1 - P1;
    float L1 = length(E1);
    V3 E2 = Q2 - P2;
    float L2 = length(E2);
    if (L1 < kLinearSlop && L2 < kLinearSlop) {
        C1 = P1;
        C2 = P2;
        return;
    }
    if (L1 < kLinearSlop) {
        C1 = P1;
        C2 = closest_on_segment(P1, P2, Q2);
        return;
    }
    if (L2 < kLinearSlop) {
        C1 = closest_on_segment(P2, P1, Q1);
        C2 = P2;
        return;
    }
    V3 N1 = divv(E1, L1);
    V3 N2 = divv(E2, L2);
    float a11 = 1.0f, a12 = -dot(N1, N2);
    float a21 = -a12, a22 = -1.0f;
    // b3Mat22 columns x = (a11, a21), y = (a12, a22); norms are max column abs sums.
    float n0 = fabsf(a11) + fabsf(a21), n1 = fabsf(a12) + fabsf(a22);
    float norm_A = n0 > n1 ? n0 : n1;
    float det_A = 1.0f / (a11 * a22 - a12 * a21);
    float Ixx = det_A * a22, Ixy = -det_A * a21;
    float Iyx = -det_A * a12, Iyy = det_A * a11;
    float m0 = fabsf(Ixx) + fabsf(Ixy), m1 = fabsf(Iyx) + fabsf(Iyy);
    float norm_inv_A = m0 > m1 ? m0 : m1;
    float k_A = norm_A * norm_inv_A;
    const float kMaxConditionNumber = 1000.0f;
    if (!isinf(k_A) && k_A < kMaxConditionNumber) {
        V3 E3 = P1 - P2;
        float bx = -dot(N1, E3);
        float by = -dot(N2, E3);
        float xx = bx * Ixx + by * Iyx;  // b3Mat22 * b3Vec2 = v.x * A.x + v.y * A.y (mat22.h:56-59)
        float xy = bx * Ixy + by * Iyy;
        C1 = P1 + xx * N1;
        C2 = P2 + xy * N2;
    } else {
        C1 = P1;
        C2 = P2;
    }
    C1 = closest_on_segment(C1, P1, Q1);
    C2 = closest_on_segment(C1, P2, Q2);
    C1 = closest_on_segment(C2, P1, Q1);
}

// collide_capsules.cpp:154-174
__device__ inline bool are_parallel(V3 P1, V3 Q1, V3 P2, V3 Q2) {
    V3 E1 = Q1 - P1;
    float L1 = length(E1);
    if (L1 < kLinearSlop) return false;
    V3 E2 = Q2 - P2;
    float L2 = length(E2);
    if (L2 < kLinearSlop) return false;
    const float kTol = 0.005f;
    V3 N = cross(E1, E2);
    return length(N) < kTol * L1 * L2;
}

// clip.cpp:59-97
__device__ inline uint32_t clip_edge_to_plane(ClipV out[2], const ClipV in[2], const Plane& pl, uint32_t id) {
    uint32_t n = 0;
    float d1 = distance(in[0].position, pl);
    float d2 = distance(in[1].position, pl);
    if (d1 <= 0.0f) out[n++] = in[0];
    if (d2 <= 0.0f) out[n++] = in[1];
    if (d1 <= 0.0f && d2 > 0.0f) {
        float f = d1 / (d1 - d2);
        out[n++] = clipv(in[0].position + f * (in[1].position - in[0].position), in[0].inEdge1, kNullEdge, kNullEdge, id);
    } else if (d1 > 0.0f && d2 <= 0.0f) {
        float f = d1 / (d1 - d2);
        out[n++] = clipv(in[0].position + f * (in[1].position - in[0].position), in[1].inEdge1, id, kNullEdge, kNullEdge);
    }
    return n;
}

// clip.cpp:22-31
__device__ __forceinline__ void build_edge(ClipV e[2], V3 A, V3 B) {
    e[0] = clipv(A, 0u, kNullEdge, 0u, kNullEdge);
    e[1] = clipv(B, 1u, kNullEdge, 0u, kNullEdge);
}

// clip.cpp:152-198 (segment side planes)
__device__ inline uint32_t clip_edge_to_segment(ClipV out[2], const ClipV in[2], V3 P1, V3 Q1) {
    ClipV cur[2] = {in[0], in[1]}, tmp[2];
    V3 E1 = Q1 - P1;
    Plane p1;
    p1.n = normalize(E1);
    p1.d = dot(p1.n, Q1);
    uint32_t n = clip_edge_to_plane(tmp, cur, p1, 0u);
    if (n < 2) return n;
    cur[0] = tmp[0];
    cur[1] = tmp[1];
    Plane p2;
    p2.n = -p1.n;
    p2.d = dot(p2.n, P1);
    n = clip_edge_to_plane(out, cur, p2, 1u);
    return n;
}

// clip.cpp:201-240 (hull face side planes)
template <class H>
__device__ inline uint32_t clip_edge_to_hull_face(ClipV out[2], const ClipV in[2], const Xf& xf, float r, uint32_t index,
                                           const H& h) {
    ClipV cur[2] = {in[0], in[1]}, tmp[2];
    uint32_t n = 0;
    uint32_t begin = h.face_edge(index);
    uint32_t e = begin;
    uint32_t guard = 0;
    do {
        uint32_t edgeId = e;  // (edge->twin)->twin
        Plane pl = edge_side_plane(h, edgeId);
        pl.d += r;
        n = clip_edge_to_plane(tmp, cur, mul(xf, pl), edgeId);
        if (n == 0) break;
        cur[0] = tmp[0];
        cur[1] = tmp[1];
        e = h.next(e);
    } while (e != begin && ++guard < 256);
    out[0] = cur[0];
    out[1] = cur[1];
    return n;
}

__device__ __forceinline__ void key_from(Contact& m, int i, const ClipV& c) {
    m.key[i][0] = c.inEdge1;
    m.key[i][1] = c.outEdge1;
    m.key[i][2] = c.inEdge2;
    m.key[i][3] = c.outEdge2;
}

// collide_capsules.cpp:176-259
__device__ inline void collide_capsules(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2, const ShapeRec& s2) {
    V3 A1 = mul(xf1, V3{s1.v[0], s1.v[1], s1.v[2]}), B1 = mul(xf1, V3{s1.v[3], s1.v[4], s1.v[5]});
    V3 A2 = mul(xf2, V3{s2.v[0], s2.v[1], s2.v[2]}), B2 = mul(xf2, V3{s2.v[3], s2.v[4], s2.v[5]});
    V3 point1, point2;
    closest_points_segments(point1, point2, A1, B1, A2, B2);
    float dist = length(point1 - point2);
    float totalRadius = s1.radius + s2.radius;
    if (dist > totalRadius) return;
    if (!(dist > kEpsilon)) return;
    if (are_parallel(A1, B1, A2, B2)) {
        ClipV edge1[2], clip1[2];
        build_edge(edge1, A1, B1);
        uint32_t clipCount = clip_edge_to_segment(clip1, edge1, A2, B2);
        if (clipCount == 2) {
            V3 cp1 = closest_on_segment(clip1[0].position, A2, B2);
            V3 cp2 = closest_on_segment(clip1[1].position, A2, B2);
            float d1 = length(clip1[0].position - cp1);
            float d2 = length(clip1[1].position - cp2);
            if (d1 > kEpsilon && d1 <= totalRadius && d2 > kEpsilon && d2 <= totalRadius) {
                V3 n1 = divv(cp1 - clip1[0].position, d1);
                V3 n2 = divv(cp2 - clip1[1].position, d2);
                m.pointCount = 2;
                m.n[0] = mulT(xf1.R, n1);
                m.p1[0] = mulT(xf1, clip1[0].position);
                m.p2[0] = mulT(xf2, cp1);
                key_from(m, 0, clip1[0]);
                m.n[1] = mulT(xf1.R, n2);
                m.p1[1] = mulT(xf1, clip1[1].position);
                m.p2[1] = mulT(xf2, cp2);
                key_from(m, 1, clip1[1]);
                return;
            }
        }
    }
    V3 normal = divv(point2 - point1, dist);
    m.pointCount = 1;
    m.n[0] = mulT(xf1.R, normal);
    m.p1[0] = mulT(xf1, point1);
    m.p2[0] = mulT(xf2, point2);
    zero_key(m, 0);
}

__device__ __forceinline__ Proxy round_proxy(const ShapeRec& s) {
    Proxy p;
    p.pts[0] = V3{s.v[0], s.v[1], s.v[2]};
    p.pts[1] = V3{s.v[3], s.v[4], s.v[5]};
    p.count = s.type == 0u ? 1u : 2u;
    p.hullVerts = nullptr;
    return p;
}

__device__ __forceinline__ Proxy hull_proxy(const DevHull& h) {
    Proxy p;
    p.count = h.V;
    p.hullVerts = h.base + 2;
    return p;
}

__device__ __forceinline__ Proxy hull_proxy(const TriHull& h) {
    Proxy p;
    p.pts[0] = h.v0;
    p.pts[1] = h.v1;
    p.pts[2] = h.v2;
    p.count = 3;
    p.hullVerts = nullptr;
    return p;
}

// collide_sphere_hull.cpp:26-87 (SAT face query: sat_vertex_and_hull.cpp:29-54)
template <class H2>
__device__ inline void collide_sphere_hull(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2, const H2& h2,
                                    float r2) {
    GJKOut g = gjk(xf1, round_proxy(s1), xf2, hull_proxy(h2));
    m.gjkCalls += 1;
    m.gjkIters += g.iterations;
    const V3 c{s1.v[0], s1.v[1], s1.v[2]};
    float totalRadius = s1.radius + r2;
    if (g.distance > totalRadius) return;
    if (g.distance > 0.0f) {
        V3 normal = divv(g.point2 - g.point1, g.distance);
        m.pointCount = 1;
        m.n[0] = mulT(xf1.R, normal);
        m.p1[0] = c;
        m.p2[0] = mulT(xf2, g.point2);
        zero_key(m, 0);
        return;
    }
    V3 support = mulT(xf2, mul(xf1, c));
    uint32_t maxIndex = 0;
    float maxSep = -kMaxFloat;
    for (uint32_t i = 0; i < h2.F; ++i) {
        float sep = distance(support, h2.plane(i));
        if (sep > maxSep) {
            maxIndex = i;
            maxSep = sep;
        }
    }
    if (maxSep > totalRadius) return;
    Plane plane2 = mul(xf2, h2.plane(maxIndex));
    V3 c1 = mul(xf1, c);
    V3 c2 = closest_on_plane(c1, plane2);
    m.pointCount = 1;
    m.n[0] = mulT(xf1.R, -plane2.n);
    m.p1[0] = c;
    m.p2[0] = mulT(xf2, c2);
    zero_key(m, 0);
}

// collide_capsule_hull.cpp:29-87
template <class H2>
__device__ inline void capsule_edge_contact(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2, uint32_t index2,
                                     const H2& h2) {
    V3 P1 = mul(xf1, V3{s1.v[0], s1.v[1], s1.v[2]});
    V3 Q1 = mul(xf1, V3{s1.v[3], s1.v[4], s1.v[5]});
    V3 E1 = Q1 - P1;
    V3 N1 = E1;
    normalize_in_place(N1);
    V3 C2 = mul(xf2, h2.centroid());
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
    V3 N = cross(E1, E2);
    normalize_in_place(N);
    if (dot(N, P2 - C2) > 0.0f) N = -N;
    m.pointCount = 1;
    m.n[0] = mulT(xf1.R, N);
    m.p1[0] = mulT(xf1, c1);
    m.p2[0] = mulT(xf2, c2);
    m.key[0][0] = 0u;           // inEdge1
    m.key[0][1] = index2;       // outEdge1
    m.key[0][2] = 1u;           // inEdge2
    m.key[0][3] = index2 + 1u;  // outEdge2
}

// collide_capsule_hull.cpp:89-140
template <class H2>
__device__ inline void capsule_face_contact(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2, uint32_t index2,
                                     const H2& h2, float r2) {
    ClipV edge1[2], clip1[2];
    build_edge(edge1, mul(xf1, V3{s1.v[0], s1.v[1], s1.v[2]}), mul(xf1, V3{s1.v[3], s1.v[4], s1.v[5]}));
    uint32_t clipCount = clip_edge_to_hull_face(clip1, edge1, xf2, r2, index2, h2);
    Plane plane2 = mul(xf2, h2.plane(index2));
    V3 n1 = -plane2.n;
    float totalRadius = s1.radius + r2;
    uint32_t pc = 0;
    for (uint32_t i = 0; i < clipCount; ++i) {
        V3 c1 = clip1[i].position;
        float s = distance(c1, plane2);
        if (s <= totalRadius) {
            V3 c2 = closest_on_plane(c1, plane2);
            m.n[pc] = mulT(xf1.R, n1);
            m.p1[pc] = mulT(xf1, c1);
            m.p2[pc] = mulT(xf2, c2);
            key_from(m, pc, clip1[i]);
            ++pc;
        }
    }
    m.pointCount = pc;
}

// collide_capsule_hull.cpp:142-222 (capsule SAT: sat_edge_and_hull.cpp:23-143)
template <class H2>
__device__ inline void collide_capsule_hull(Contact& m, const Xf& xf1, const ShapeRec& s1, const Xf& xf2, const H2& h2,
                                     float r2) {
    GJKOut g = gjk(xf1, round_proxy(s1), xf2, hull_proxy(h2));
    m.gjkCalls += 1;
    m.gjkIters += g.iterations;
    float totalRadius = s1.radius + r2;
    if (g.distance > totalRadius) return;
    const V3 c0{s1.v[0], s1.v[1], s1.v[2]}, c1{s1.v[3], s1.v[4], s1.v[5]};
    if (g.distance > kEpsilon) {
        V3 N1 = divv(g.point2 - g.point1, g.distance);
        V3 localN1 = mulT(xf2.R, N1);
        V3 dn = -localN1;
        uint32_t index2 = 0;
        float mp = dot(dn, h2.plane(0).n);
        for (uint32_t i = 1; i < h2.F; ++i) {
            float q = dot(dn, h2.plane(i).n);
            if (q > mp) {
                mp = q;
                index2 = i;
            }
        }
        V3 N2 = mul(xf2.R, h2.plane(index2).n);
        const float kTol = 0.005f;
        V3 N = cross(N1, N2);
        float L = dot(N, N);
        if (L < kTol * kTol) {
            capsule_face_contact(m, xf1, s1, xf2, index2, h2, r2);
            if (m.pointCount == 2) return;
        }
        m.pointCount = 1;
        m.n[0] = mulT(xf1.R, N1);
        m.p1[0] = mulT(xf1, g.point1);
        m.p2[0] = mulT(xf2, g.point2);
        zero_key(m, 0);
        return;
    }
    // Face query of the hull planes in the capsule frame.
    Xf xf = mulT(xf1, xf2);
    uint32_t fIndex = 0;
    float fSep = -kMaxFloat;
    for (uint32_t i = 0; i < h2.F; ++i) {
        Plane plane = mul(xf, h2.plane(i));
        V3 dn = -plane.n;
        V3 sv = dot(dn, c0) > dot(dn, c1) ? c0 : c1;  // b3Capsule::GetSupportVertex (capsule.h:50-57)
        float sep = distance(sv, plane);
        if (sep > fSep) {
            fIndex = i;
            fSep = sep;
        }
    }
    if (fSep > totalRadius) return;
    // Edge query (sat_edge_and_hull.cpp:99-143).
    Xf xe = mulT(xf2, xf1);
    V3 P1 = mul(xe, c0);
    V3 Q1 = mul(xe, c1);
    V3 E1 = Q1 - P1;
    V3 C2 = h2.centroid();
    uint32_t eIndex = 0;
    float eSep = -kMaxFloat;
    for (uint32_t i = 0; i < h2.HE; i += 2) {
        V3 P2 = h2.vertex(h2.origin(i));
        V3 Q2 = h2.vertex(h2.origin(i + 1));
        V3 E2 = Q2 - P2;
        V3 U2 = h2.plane(h2.face(i)).n;
        V3 V2 = h2.plane(h2.face(i + 1)).n;
        if (dot(E1, U2) * dot(E1, V2) < 0.0f) {
            float sep;
            float L1 = length(E1);
            if (L1 < kLinearSlop) {
                sep = -kMaxFloat;
            } else {
                float L2 = length(E2);
                if (L2 < kLinearSlop) {
                    sep = -kMaxFloat;
                } else {
                    const float kTolP = 0.005f;
                    V3 E1xE2 = cross(E1, E2);
                    float L = length(E1xE2);
                    if (L < kTolP * L1 * L2) {
                        sep = -kMaxFloat;
                    } else {
                        V3 N = (1.0f / L) * E1xE2;
                        if (dot(N, P2 - C2) > 0.0f) N = -N;
                        sep = dot(N, P2 - P1);
                    }
                }
            }
            if (sep > eSep) {
                eSep = sep;
                eIndex = i;
            }
        }
    }
    if (eSep > totalRadius) return;
    const float kTol = 0.1f * kLinearSlop;
    if (eSep > fSep + kTol) capsule_edge_contact(m, xf1, s1, xf2, eIndex, h2);
    else capsule_face_contact(m, xf1, s1, xf2, fIndex, h2, r2);
}

}  // namespace b3amd

#endif
