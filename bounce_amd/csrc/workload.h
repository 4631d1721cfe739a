// Seeded, libm-free workload primitives shared by the generators.
#ifndef B3AMD_WORKLOAD_H
#define B3AMD_WORKLOAD_H

#include <cstdint>

#include "hullgen.h"

namespace b3amd {

struct SplitMix {
    uint64_t state;
    uint64_t next();
    double u01();                  // [0, 1) with 53 random bits
    double normal();               // Irwin-Hall(12) approximation of N(0, 1)
    void unit_vector(double v[3]); // rejection in the unit ball, then normalise
    void unit_quat(double q[4]);   // rejection in the 4-ball, then normalise
};

// Independent stream for item `index` of a seeded generator.
SplitMix stream(uint64_t seed, uint64_t index);

// Rotation matrix columns (b3Mat33 layout) of a unit quaternion (x, y, z, w).
void quat_to_rotation(const double q[4], float R[9]);

// Hull of n points on the ellipsoid with semi-axes a, b, c.
int ellipsoid_hull(SplitMix& rng, int n, double a, double b, double c, HullRecord* out);

}  // namespace b3amd

#endif
