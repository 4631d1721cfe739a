// The step after the narrowphase, on the device: warm-start matching of the new manifolds
// against the contact's old ones and b3ContactSolver::InitializeConstraints, so that a
// world's manifolds go from the narrowphase to solver-ready velocity constraints without
// leaving HBM.
//
//   b3Manifold::Initialize() / Initialize(old)   manifold.cpp:21-55, in b3Contact::Update's
//                                                loop over the old manifolds (contact.cpp:63-91)
//   b3WorldManifold(Point)::Initialize           manifold.cpp:57-95
//   InitializeConstraints                        contact_solver.cpp:70-276
//
// One lane per (contact, manifold slot): every output depends only on that slot's new
// manifold, the contact's old manifolds and impulses, the two bodies and the materials.  The
// arithmetic follows the reference's FP32 operation order (b3_math.h), including the old
// manifolds applied in turn so that tangent and motor impulses come from the last one.
#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "kernels.h"

namespace b3amd {
namespace {

__device__ __forceinline__ V3 ld3(const float* p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ void st3(float* p, V3 v) { p[0] = v.x, p[1] = v.y, p[2] = v.z; }

__device__ __forceinline__ M33 ld33(const float* a) {
    M33 m;
    m.c0 = ld3(a);
    m.c1 = ld3(a + 3);
    m.c2 = ld3(a + 6);
    return m;
}

__device__ __forceinline__ bool same_key(const ManifoldPoint& a, const ManifoldPoint& b) {  // manifold.h:30-34
    return a.triangleKey == b.triangleKey && a.key[0] == b.key[0] && a.key[1] == b.key[1] && a.key[2] == b.key[2] &&
           a.key[3] == b.key[3];
}

__global__ __launch_bounds__(256) void contact_init_kernel(ContactInitArgs a) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.n * a.M) return;
    const uint32_t i = s / a.M, j = s - i * a.M;
    if (j >= a.counts[i]) return;
    const Manifold& m = a.manifolds[s];
    // b3Manifold::Initialize(), then Initialize(old) of every old manifold in turn.
    ManifoldImpulses imp;
    for (int k = 0; k < 4; ++k) imp.normalImpulse[k] = 0.0f, imp.persisting[k] = 0u;
    imp.tangentImpulse[0] = imp.tangentImpulse[1] = 0.0f;
    imp.motorImpulse = 0.0f;
    imp.reserved = 0u;
    if (a.warmStarting) {
        const uint32_t oc = a.oldCounts[i];
        for (uint32_t k = 0; k < oc; ++k) {
            const Manifold& om = a.oldManifolds[(size_t)i * a.M + k];
            const ManifoldImpulses& oi = a.oldImpulses[(size_t)i * a.M + k];
            imp.tangentImpulse[0] = oi.tangentImpulse[0];
            imp.tangentImpulse[1] = oi.tangentImpulse[1];
            imp.motorImpulse = oi.motorImpulse;
            for (uint32_t p = 0; p < m.pointCount; ++p)
                for (uint32_t q = 0; q < om.pointCount; ++q)
                    if (same_key(m.points[p], om.points[q])) {
                        imp.normalImpulse[p] = oi.normalImpulse[q];
                        imp.persisting[p] = 1u;
                        break;
                    }
        }
    }
    a.impulses[s] = imp;

    const float* A = a.state + kSolverBodyFloats * (size_t)a.bodies[2 * i];
    const float* B = a.state + kSolverBodyFloats * (size_t)a.bodies[2 * i + 1];
    const float* mt = a.material + 8 * (size_t)i;
    const V3 xA = ld3(A), xB = ld3(B);
    const V3 vA = ld3(A + 7), wA = ld3(A + 10), vB = ld3(B + 7), wB = ld3(B + 10);
    const float mA = A[13], mB = B[13];
    const M33 iA = ld33(A + 14), iB = ld33(B + 14);
    Xf xfA, xfB;
    xfA.R = quat_mat33(A[3], A[4], A[5], A[6]);
    xfA.p = xA - mul(xfA.R, ld3(A + 23));
    xfB.R = quat_mat33(B[3], B[4], B[5], B[6]);
    xfB.p = xB - mul(xfB.R, ld3(B + 23));
    const float radiusA = mt[0], radiusB = mt[1];
    const float friction = sqrtf(mt[2] * mt[3]);          // b3MixFriction (contact_solver.h:93-96)
    const float restitution = mt[4] > mt[5] ? mt[4] : mt[5];  // b3MixRestitution (:87-91)

    VelocityManifold o;
    float* ow = reinterpret_cast<float*>(&o);
    for (uint32_t k = 0; k < sizeof(VelocityManifold) / 4; ++k) ow[k] = 0.0f;
    o.pointCount = m.pointCount;
    o.tangentImpulse[0] = imp.tangentImpulse[0];
    o.tangentImpulse[1] = imp.tangentImpulse[1];
    o.motorImpulse = imp.motorImpulse;
    o.friction = friction;
    o.restitution = restitution;
    if (m.pointCount > 0u) {
        // b3WorldManifold::Initialize: per-point world point (midpoint of the radius-offset
        // points) and normal, their sums in point order, centre and normalised normal.
        V3 center{0.0f, 0.0f, 0.0f}, normal{0.0f, 0.0f, 0.0f};
        V3 wp[4], wn[4];
        for (uint32_t k = 0; k < m.pointCount; ++k) {
            const ManifoldPoint& p = m.points[k];
            const V3 nA = mul(xfA.R, p.localNormal1);
            const V3 cA = mul(xfA, p.localPoint1);
            const V3 cB = mul(xfB, p.localPoint2);
            const V3 pA = cA + radiusA * nA;
            const V3 pB = cB - radiusB * nA;
            wp[k] = 0.5f * (pA + pB);
            wn[k] = nA;
            center = center + wp[k];
            normal = normal + nA;
        }
        {
            const float sc = 1.0f / (float)m.pointCount;  // center /= count
            center = V3{center.x * sc, center.y * sc, center.z * sc};
        }
        normalize_in_place(normal);
        const V3 t1 = fabsf(normal.x) >= 0.57735027f ? V3{normal.y, -normal.x, 0.0f}
                                                     : V3{0.0f, normal.z, -normal.y};  // b3Perp (vec3.h:255-267)
        const V3 t2 = cross(t1, normal);
        st3(o.normal, normal);
        st3(o.tangent1, t1);
        st3(o.tangent2, t2);
        for (uint32_t k = 0; k < m.pointCount; ++k) {
            VelocityPoint& vp = o.points[k];
            const V3 n = wn[k];
            const V3 rA = wp[k] - xA, rB = wp[k] - xB;
            st3(vp.rA, rA);
            st3(vp.rB, rB);
            st3(vp.normal, n);
            vp.normalImpulse = imp.normalImpulse[k];
            const V3 rnA = cross(rA, n), rnB = cross(rB, n);
            const float K = mA + mB + dot(mul(iA, rnA), rnA) + dot(mul(iB, rnB), rnB);
            vp.normalMass = K > 0.0f ? 1.0f / K : 0.0f;
            const V3 dv = ((vB + cross(wB, rB)) - vA) - cross(wA, rA);
            const float vn = dot(n, dv);
            vp.velocityBias = vn < -1.0f ? -restitution * vn : 0.0f;  // B3_VELOCITY_THRESHOLD (settings.h:105)
        }
        const V3 rA = center - xA, rB = center - xB;
        st3(o.rA, rA);
        st3(o.rB, rB);
        const V3 rn1A = cross(rA, t1), rn1B = cross(rB, t1), rn2A = cross(rA, t2), rn2B = cross(rB, t2);
        const float k11 = mA + mB + dot(mul(iA, rn1A), rn1A) + dot(mul(iB, rn1B), rn1B);
        const float k12 = dot(mul(iA, rn1A), rn2A) + dot(mul(iB, rn1B), rn2B);
        const float k22 = mA + mB + dot(mul(iA, rn2A), rn2A) + dot(mul(iB, rn2B), rn2B);
        float det = k11 * k22 - k12 * k12;  // b3Inverse(b3Mat22) (mat22.cpp:77-93)
        if (det != 0.0f) det = 1.0f / det;
        o.tangentMass[0] = det * k22;
        o.tangentMass[1] = -det * k12;
        o.tangentMass[2] = -det * k12;
        o.tangentMass[3] = det * k11;
        M33 iAB;
        iAB.c0 = iA.c0 + iB.c0;
        iAB.c1 = iA.c1 + iB.c1;
        iAB.c2 = iA.c2 + iB.c2;
        const float mass = dot(normal, mul(iAB, normal));
        o.motorMass = mass > 0.0f ? 1.0f / mass : 0.0f;
    }
    a.constraints[s] = o;
}

}  // namespace

hipError_t launch_contact_init(const ContactInitArgs& a, hipStream_t s) {
    const uint64_t lanes = (uint64_t)a.n * a.M;
    if (lanes == 0) return hipSuccess;
    contact_init_kernel<<<(uint32_t)((lanes + 255) / 256), 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace b3amd
