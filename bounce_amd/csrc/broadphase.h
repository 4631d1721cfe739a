// Internal: the device broadphase's state and the device functions the broadphase
// (broadphase.hip) and the contact manager (contact_manager.hip) share.  Not part of the
// public interface.
#ifndef B3AMD_BROADPHASE_H
#define B3AMD_BROADPHASE_H

#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "bounce_amd.h"
#include "context.h"
#include "kernels.h"

namespace b3amd {

constexpr float kAabbExtension = 0.2f;   // B3_AABB_EXTENSION, settings.h:54
constexpr float kAabbMultiplier = 2.0f;  // B3_AABB_MULTIPLIER, settings.h:59
constexpr uint32_t kLive = 1u, kMoved = 2u;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct BpEdit {
    uint32_t id, mode, user, pad;
    float lower[3], upper[3];
};
enum { kEditCreate = 1, kEditDestroy = 2, kEditTouch = 3 };

// Fat AABBs of proxies as two float4 arrays indexed by proxy id (w unused).
struct BpState {
    float4* lower;
    float4* upper;
    uint32_t* user;
    uint32_t* flags;
};

__device__ __forceinline__ bool overlap(float4 al, float4 au, float4 bl, float4 bu) {  // aabb3.h:247-251
    return (al.x <= bu.x) && (al.y <= bu.y) && (al.z <= bu.z) && (au.x >= bl.x) && (au.y >= bl.y) &&
           (au.z >= bl.z);
}

// b3BroadPhase::MoveProxy (broad_phase.cpp:99-149) for live proxy `id`: whether it moved.
__device__ __forceinline__ bool move_proxy(const BpState& st, uint32_t id, V3 lo, V3 hi, V3 disp) {
    const float4 L = st.lower[id], U = st.upper[id];
    // b3AABB3::Contains(aabb) = Contains(m_lower) && Contains(m_upper) (aabb3.h:156-167).
    const bool contains = (L.x <= lo.x && lo.x <= U.x && L.y <= lo.y && lo.y <= U.y && L.z <= lo.z && lo.z <= U.z) &&
                          (L.x <= hi.x && hi.x <= U.x && L.y <= hi.y && hi.y <= U.y && L.z <= hi.z && hi.z <= U.z);
    if (contains) return false;
    float4 nl = make_float4(lo.x - kAabbExtension, lo.y - kAabbExtension, lo.z - kAabbExtension, 0.0f);
    float4 nu = make_float4(hi.x + kAabbExtension, hi.y + kAabbExtension, hi.z + kAabbExtension, 0.0f);
    const float dx = kAabbMultiplier * disp.x, dy = kAabbMultiplier * disp.y, dz = kAabbMultiplier * disp.z;
    if (dx < 0.0f) nl.x += dx; else nu.x += dx;
    if (dy < 0.0f) nl.y += dy; else nu.y += dy;
    if (dz < 0.0f) nl.z += dz; else nu.z += dz;
    st.lower[id] = nl;
    st.upper[id] = nu;
    st.flags[id] |= kMoved;
    return true;
}

// ---- shape AABBs (b3Body::SynchronizeShapes, body.cpp:181-202) ----

// b3Sweep::GetTransform(0.0f) (transform.h:224-234) with the reference's expressions.
__device__ __forceinline__ Xf sweep_transform0(const float* s) {
    // b3Sweep: localCenter s[0..2], orientation0 s[3..6], worldCenter0 s[7..9], t0 s[10],
    // orientation s[11..14], worldCenter s[15..17].
    const float t = 0.0f, u = 1.0f - t;
    const V3 lc{s[0], s[1], s[2]};
    const V3 c{u * s[7] + t * s[15], u * s[8] + t * s[16], u * s[9] + t * s[17]};
    float qx = u * s[3] + t * s[11], qy = u * s[4] + t * s[12], qz = u * s[5] + t * s[13], qw = u * s[6] + t * s[14];
    const float len = sqrtf(qx * qx + qy * qy + qz * qz + qw * qw);  // b3Quat::Normalize (quat.h:88-100)
    if (len > kEpsilon) {
        const float k = 1.0f / len;
        qx *= k, qy *= k, qz *= k, qw *= k;
    }
    Xf xf;
    xf.R = quat_mat33(qx, qy, qz, qw);
    // b3Mul(q, v) (quat.h:225-232): v + w * t + cross(qv, t), t = 2 cross(qv, v).
    const V3 qv{qx, qy, qz};
    const V3 tt = 2.0f * cross(qv, lc);
    const V3 r = (lc + qw * tt) + cross(qv, tt);
    xf.p = c - r;
    return xf;
}

__device__ __forceinline__ V3 vmin(V3 a, V3 b) {  // b3Min (vec3.h:243-246, math.h:54-57)
    return V3{a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z};
}
__device__ __forceinline__ V3 vmax(V3 a, V3 b) {
    return V3{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z};
}

// b3{Sphere,Capsule,Hull}Shape::ComputeAABB (sphere_shape.cpp:53-59, capsule_shape.cpp:130-137,
// hull_shape.cpp:162-166 with b3AABB3::Set, aabb3.h:52-61).
__device__ void shape_aabb(const ShapeRec& s, const float4* pool, const Xf& xf, V3& lo, V3& hi) {
    const V3 r{s.radius, s.radius, s.radius};
    if (s.type == B3AMD_SPHERE) {
        const V3 c = mul(xf, V3{s.v[0], s.v[1], s.v[2]});
        lo = c - r;
        hi = c + r;
    } else if (s.type == B3AMD_CAPSULE) {
        const V3 c1 = mul(xf, V3{s.v[0], s.v[1], s.v[2]});
        const V3 c2 = mul(xf, V3{s.v[3], s.v[4], s.v[5]});
        lo = vmin(c1, c2) - r;
        hi = vmax(c1, c2) + r;
    } else {
        const DevHull h = DevHull::at_unit(pool, s.hullUnit);
        lo = hi = mul(xf, h.vertex(0));
        for (uint32_t i = 1; i < h.V; ++i) {
            const V3 v = mul(xf, h.vertex(i));
            lo = vmin(lo, v);
            hi = vmax(hi, v);
        }
        lo = lo - r;  // Extend(m_radius)
        hi = hi + r;
    }
}

// b3amd_broadphase_create_proxies for AABBs already on the device (d_aabbs, [n][6]): the
// ids come from the host's model of the node pool as there, the fat AABBs are made on the
// device.  Enqueued on the context stream: `records` ([n][2] host words: id, user data) is
// copied to d_records, so the caller keeps both until it synchronises the stream.
int bp_create_proxies_device(b3amd_broadphase* bp, uint32_t n, const float* d_aabbs, const uint32_t* users,
                             uint32_t* ids, uint32_t* records, uint32_t* d_records);
// b3amd_broadphase_find_pairs; with sync false the last launches (the pair sort and the
// move-buffer reset) are left enqueued on the context stream.
int bp_find_pairs(b3amd_broadphase* bp, uint32_t* pair_count, const uint32_t** d_pairs, bool sync);
// b3amd_shape_aabbs without the status read-back, for inputs the caller has checked:
// enqueued on the context stream.
int launch_shape_aabbs(b3amd_context* ctx, uint32_t n, const uint32_t* d_shapes, const uint32_t* d_bodies,
                       uint32_t body_count, const b3amd_transform* d_body_xfs, float* d_aabbs);

}  // namespace b3amd

struct b3amd_broadphase {
    b3amd_context* ctx = nullptr;
    // Node pool of b3DynamicTree as seen through proxy ids: groups pushed by destroys
    // (leaf id, whether a parent node sits under it) above the never-used nodes.
    std::vector<std::pair<uint32_t, bool>> freeGroups;
    uint32_t fresh = 0;
    uint32_t count = 0;
    std::vector<uint8_t> live;
    uint32_t idCap = 0;
    BpState st{nullptr, nullptr, nullptr, nullptr};
    // find_pairs scratch
    uint32_t leafCap = 0;
    uint32_t* dLeaves = nullptr;   // live ids, then the query list after them
    uint32_t* dQueries = nullptr;
    uint64_t* dKeys = nullptr;     // [leafCap] unsorted / sorted Morton keys
    uint64_t* dKeys2 = nullptr;
    uint32_t* dNodes = nullptr;    // left, right, parent, escape, first, last
    float4* dBoxes = nullptr;      // internal node boxes, then the refit's range tables
    uint32_t* dCounts = nullptr;   // [0] leaves, [1] queries, [2] pairs, [3..8] bounds, [9] status
    uint32_t pairCap = 0;
    uint64_t* dPairKeys = nullptr;
    uint64_t* dPairKeys2 = nullptr;
    uint32_t* dPairs = nullptr;    // [pairCap][4]
    void* dSortTmp = nullptr;
    size_t sortTmpBytes = 0;
    BpEdit* dEdits = nullptr;
    uint32_t editCap = 0;
    uint32_t lastPairs = 0;
};


#endif
