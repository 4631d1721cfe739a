// Device-resident hull records.
//
// One b3Hull (hull.h:37-64) becomes one contiguous block of 16-byte units in HBM:
//   unit 0        header  x = V | F << 16, y = HE | maxFaceDegree << 16, z = HE / 2
//   unit 1        centroid (x, y, z) and w = bounding radius (no vertex farther from it)
//   units 2..     V vertices (x, y, z, 0)
//   then          F planes (nx, ny, nz, offset)
//   then          HE/2 unique edges, one u32 each: origin(2k) | origin(2k+1) << 8 |
//                 face(2k) << 16 | face(2k+1) << 24  (the SAT edge loop's working set)
//   then          HE half-edges, one u32 each: origin | face << 8 | next << 16; the twin
//                 of half-edge h is h ^ 1 (required by sat.cpp:123-128, checked at upload)
//   then          F u16 first half-edge per face (b3Face::edge)
//   then          the negated Gauss map, 3 units per unique edge k padded to a multiple
//                 of 4 edges: {-U.xyz, -V.x}, {-V.yz, -E.xy}, {-E.z, P.xyz} with U, V the
//                 normals of faces(2k), face(2k+1), P = origin(2k) and E = origin(2k+1) -
//                 P, in the hull's own frame.  -U, -V, -E are exactly the operands hull 2
//                 contributes to b3IsMinkowskiFace in the edge query (sat.cpp:131-156),
//                 and P, -E its edge for b3Project (sat.cpp:158).
//   then          one unit per (padded) edge, {-m.xyz, cos a}: the bounding cap of the
//                 edge's Gauss-map arc from U to V, centre m = (U + V) / |U + V| and angular
//                 radius a = angle(U, V) / 2 + 0.01 rad, which contains the arc with a
//                 margin, so two arcs whose caps are disjoint cannot pass the Gauss-map
//                 test; the SAT kernel tests caps first and runs the exact test only for
//                 pairs whose caps overlap.  (-m, a) is the cap of the negated arc, and a
//                 hull used as hull 1 reads m back as the negation.  Padding: cos a = 1e30.
//   then          the same caps as halves (rounded to nearest), the matrix-core operands of
//                 the cap test: per block b of 32 (padded) edges, 20 units = 320 bytes:
//                 bytes 8t .. 8t + 7 hold (-m.x, -m.y, -m.z, cos a) of edge 32b + t and
//                 bytes 256 + 2t .. 256 + 2t + 1 its sin a; padding edges are zero (the
//                 kernel masks them).
#ifndef B3AMD_DEVICE_HULL_H
#define B3AMD_DEVICE_HULL_H

#include <cstdint>

#include "b3_math.h"

namespace b3amd {

constexpr uint32_t kMaxHullVertices = 255;
constexpr uint32_t kMaxHullFaces = 255;
constexpr uint32_t kMaxHullHalfEdges = 4096;

struct HullLayout {
    uint32_t V, F, HE;
    B3_HD uint32_t E() const { return HE >> 1; }
    B3_HD uint32_t vertOff() const { return 2; }
    B3_HD uint32_t planeOff() const { return 2 + V; }
    B3_HD uint32_t uedgeOff() const { return 2 + V + F; }
    B3_HD uint32_t halfOff() const { return uedgeOff() + (E() + 3) / 4; }
    B3_HD uint32_t faceOff() const { return halfOff() + (HE + 3) / 4; }
    B3_HD uint32_t gaussOff() const { return faceOff() + (F + 7) / 8; }
    B3_HD uint32_t gaussEdges() const { return (E() + 3) & ~3u; }
    B3_HD uint32_t capOff() const { return gaussOff() + 3 * gaussEdges(); }
    B3_HD uint32_t filterBlocks() const { return (gaussEdges() + 31) / 32; }
    B3_HD uint32_t filterOff() const { return capOff() + gaussEdges(); }
    B3_HD uint32_t units() const { return filterOff() + 20 * filterBlocks(); }
};

// Read-only view of one hull record in global memory.
struct DevHull {
    const float4* base;
    uint32_t V, F, HE;

    __device__ static DevHull at(const float4* pool, const uint32_t* offsets, uint32_t h) {
        return at_unit(pool, offsets[h]);
    }
    // The record starting at 16-byte unit `unit` of the pool (ShapeRec::hullUnit).
    __device__ static DevHull at_unit(const float4* pool, uint32_t unit) {
        DevHull r;
        r.base = pool + unit;
        uint4 hd = reinterpret_cast<const uint4*>(r.base)[0];
        r.V = hd.x & 0xFFFFu;
        r.F = hd.x >> 16;
        r.HE = hd.y & 0xFFFFu;
        return r;
    }
    __device__ HullLayout layout() const { return HullLayout{V, F, HE}; }
    __device__ V3 centroid() const {
        float4 q = base[1];
        return V3{q.x, q.y, q.z};
    }
    __device__ V3 vertex(uint32_t i) const {
        float4 q = base[2 + i];
        return V3{q.x, q.y, q.z};
    }
    __device__ Plane plane(uint32_t i) const {
        float4 q = base[2 + V + i];
        return Plane{V3{q.x, q.y, q.z}, q.w};
    }
    __device__ uint32_t uedge(uint32_t k) const {
        return reinterpret_cast<const uint32_t*>(base + layout().uedgeOff())[k];
    }
    __device__ uint32_t half(uint32_t h) const {
        return reinterpret_cast<const uint32_t*>(base + layout().halfOff())[h];
    }
    __device__ uint32_t origin(uint32_t h) const { return half(h) & 0xFFu; }
    __device__ uint32_t face(uint32_t h) const { return (half(h) >> 8) & 0xFFu; }
    __device__ uint32_t next(uint32_t h) const { return half(h) >> 16; }
    __device__ uint32_t face_edge(uint32_t f) const {
        return reinterpret_cast<const uint16_t*>(base + layout().faceOff())[f];
    }
};


// The 2-face hull of one mesh triangle, built in registers exactly as b3TriangleHull::Set
// (triangle_hull.h:38-71): centroid (A + B + C) / 3, planes (N, dot(N, centroid)) and
// (-N, dot(-N, centroid)) with N = normalize((B - A) x (C - A)), and the fixed half-edge
// table (origin, twin, face, next): 0:(0,1,0,2) 1:(1,0,1,3) 2:(1,3,0,4) 3:(2,2,1,5)
// 4:(2,5,0,0) 5:(0,4,1,1).
struct TriHull {
    V3 v0, v1, v2, c;
    Plane p0, p1;
    static constexpr uint32_t V = 3, F = 2, HE = 6;

    __device__ static TriHull make(V3 A, V3 B, V3 C) {
        TriHull t;
        V3 s = (A + B) + C;
        t.c = V3{s.x / 3.0f, s.y / 3.0f, s.z / 3.0f};
        t.v0 = A;
        t.v1 = B;
        t.v2 = C;
        V3 N = normalize(cross(B - A, C - A));
        t.p0 = Plane{N, dot(N, t.c)};
        t.p1 = Plane{-N, dot(-N, t.c)};
        return t;
    }
    __device__ V3 vertex(uint32_t i) const { return i == 0 ? v0 : (i == 1 ? v1 : v2); }
    __device__ Plane plane(uint32_t i) const { return i == 0 ? p0 : p1; }
    __device__ V3 centroid() const { return c; }
    __device__ uint32_t origin(uint32_t h) const { return (0x22110u >> (4u * h)) & 0xFu; }
    __device__ uint32_t face(uint32_t h) const { return h & 1u; }
    __device__ uint32_t next(uint32_t h) const { return h < 4u ? h + 2u : h - 4u; }
    __device__ uint32_t face_edge(uint32_t f) const { return f; }
};

// Per-pair result of the separation queries (sat.cpp:33-173).
struct SatResult {
    uint32_t f1Index;
    float f1Sep;
    uint32_t f2Index;
    float f2Sep;
    uint32_t eIndex1;
    uint32_t eIndex2;
    float eSep;
    uint32_t stage;  // kStage*
};

enum : uint32_t {
    kStageSkip = 0,       // resolved by the feature cache, no SAT
    kStageSepFace1 = 1,   // separated on a face of hull 1
    kStageSepFace2 = 2,   // separated on a face of hull 2
    kStageSepEdge = 3,    // separated on an edge pair
    kStageOverlap = 4,    // overlapping: contact build needed
};

}  // namespace b3amd

#endif
