// Convex hull records in the reference's b3Hull layout (hull.h:37-64), built on the host by
// the quickhull restatement (quickhull.cpp: b3QHull::Set, qhull.cpp:104-354).
#ifndef B3AMD_HULLGEN_H
#define B3AMD_HULLGEN_H

#include <cstdint>
#include <vector>

namespace b3amd {

struct HullRecord {
    float centroid[3];
    std::vector<float> vertices;   // [V][3]
    std::vector<uint32_t> edges;   // [HE][4] origin, twin, face, next
    std::vector<uint32_t> faces;   // [F] first half-edge
    std::vector<float> planes;     // [F][4] normal, offset
};

// b3QHull::Set of `count` points (simplify as its parameter); returns the vertex count, 0 when
// the reference would leave the hull empty.  orphans (may be null): a vertex of qhHull's list
// that no face uses, whose record the reference leaves unwritten.
int quickhull_record(const float* points, uint32_t count, bool simplify, HullRecord* out, bool* orphans);

}  // namespace b3amd

#endif
