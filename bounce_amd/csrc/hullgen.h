// Convex hull records in the reference's b3Hull layout (hull.h:37-64), built on the host.
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

// Triangulated convex hull of n points; returns the vertex count (>= 4) or a
// negative code for degenerate input.
int build_hull(const float* points, int n, HullRecord* out);

}  // namespace b3amd

#endif
