// Device helpers shared by the narrowphase kernels: input loads, clip vertices, the
// per-lane contact accumulator and the manifold writer.
#ifndef B3AMD_DEVICE_COMMON_H
#define B3AMD_DEVICE_COMMON_H

#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "device_hull.h"
#include "kernels.h"

namespace b3amd {

__device__ __forceinline__ V3 ld3(const float4& q) { return V3{q.x, q.y, q.z}; }

// (separation, key) arg-max over the wave: larger separation wins, equal separations
// keep the smaller key.  Sequential strict-'>' scans in index order (sat.cpp:47,158)
// pick exactly this element.
__device__ __forceinline__ void wave_argmax(float& s, uint32_t& key) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float os = __shfl_xor(s, off, 64);
        uint32_t ok = __shfl_xor(key, off, 64);
        if (os > s || (os == s && ok < key)) {
            s = os;
            key = ok;
        }
    }
}

// Rank of this lane among the set lanes of a wave mask.
__device__ __forceinline__ uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ Xf load_xf(const float* p) {
    Xf x;
    x.R.c0 = V3{p[0], p[1], p[2]};
    x.R.c1 = V3{p[3], p[4], p[5]};
    x.R.c2 = V3{p[6], p[7], p[8]};
    x.p = V3{p[9], p[10], p[11]};
    return x;
}

struct ClipV {
    V3 position;
    uint32_t inEdge1, outEdge1, inEdge2, outEdge2;
};

struct Contact {
    uint32_t pointCount;
    bool overflow;
    V3 n[4], p1[4], p2[4];
    uint32_t key[4][4];  // inEdge1, outEdge1, inEdge2, outEdge2
    uint32_t gjkCalls = 0, gjkIters = 0;  // b3_gjkCalls / b3_gjkIters contributions (gjk.cpp:583,673)
};

// Adds a per-lane value to a device total with one atomic per wave.  Every lane of the
// wave must be active (callers keep finished lanes alive to the end of the kernel).
__device__ __forceinline__ void wave_count(unsigned long long* dst, uint32_t v) {
    unsigned long long s = v;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63u) == 0 && s) atomicAdd(dst, s);
}
__device__ __forceinline__ void wave_count(unsigned int* dst, uint32_t v) {
    uint32_t s = v;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63u) == 0 && s) atomicAdd(dst, s);
}
__device__ __forceinline__ void wave_max(unsigned int* dst, uint32_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & 63u) == 0 && v) atomicMax(dst, v);
}

// b3MakePair argument order is (inEdge1, inEdge2, outEdge1, outEdge2) (clip.h:36-44).
__device__ __forceinline__ ClipV clipv(V3 p, uint32_t in1, uint32_t in2, uint32_t out1, uint32_t out2) {
    ClipV c;
    c.position = p;
    c.inEdge1 = in1;
    c.inEdge2 = in2;
    c.outEdge1 = out1;
    c.outEdge2 = out2;
    return c;
}

// hull.inl:63-78
template <class H>
__device__ inline Plane edge_side_plane(const H& h, uint32_t index) {
    V3 P = h.vertex(h.origin(index));
    V3 Q = h.vertex(h.origin(index ^ 1u));
    V3 E = Q - P;
    V3 D = cross(E, h.plane(h.face(index)).n);
    Plane r;
    r.n = normalize(D);
    r.d = dot(r.n, P);
    return r;
}

__device__ inline void write_manifold(Manifold* out, const Contact& c) {
    Manifold m;
    m.pointCount = c.pointCount;
    m.flags = c.overflow ? kFlagPolyOverflow : 0u;
    m.reserved[0] = 0;
    m.reserved[1] = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        ManifoldPoint& q = m.points[i];
        if ((uint32_t)i < c.pointCount) {
            q.localNormal1 = c.n[i];
            q.localPoint1 = c.p1[i];
            q.localPoint2 = c.p2[i];
            q.triangleKey = kNullTriangle;
            q.key[0] = c.key[i][0];
            q.key[1] = c.key[i][1];
            q.key[2] = c.key[i][2];
            q.key[3] = c.key[i][3];
        } else {
            q.localNormal1 = V3{0.0f, 0.0f, 0.0f};
            q.localPoint1 = V3{0.0f, 0.0f, 0.0f};
            q.localPoint2 = V3{0.0f, 0.0f, 0.0f};
            q.triangleKey = 0;
            q.key[0] = q.key[1] = q.key[2] = q.key[3] = 0;
        }
    }
    *out = m;
}

}  // namespace b3amd

#endif
