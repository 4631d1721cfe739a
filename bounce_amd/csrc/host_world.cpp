// A b3World step on the host around the device's contact pipeline: b3World::Step / Solve
// (world.cpp:115-337), b3Island::Solve (island.cpp:109-341), b3ContactSolver
// (contact_solver.cpp) and the body state they update (body.cpp), restated in C++ with the
// reference's FP32 operation order, while the broadphase, the contact manager and the
// narrowphase run on the device (b3amd_broadphase_*, b3amd_contact_manager_*).
//
// What the reference's results depend on beyond the arithmetic, and how it is kept:
//  * the body list is pushed at the front (world.cpp:86-92), so Solve seeds islands from the
//    newest body; a body's shapes likewise (body.cpp:101);
//  * a shape's contact edges are pushed at the front when a contact is created
//    (contact_manager.cpp:104-114), so they are the contacts of the world's contact list
//    (also newest first) that involve the shape, in list order;
//  * islands are depth-first with an explicit stack (world.cpp:190-287); body and contact
//    order inside an island is the order of the solver's loops;
//  * b3Contact::Update's warm start (contact.cpp:63-91): the new manifold takes the
//    normal impulse of the old point with the same key and the tangent and motor impulses of
//    the old manifold.
// Joints are not modelled (no joint graph); sensors are refused at shape creation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "bounce_amd.h"
#include "context.h"

namespace b3amd {
namespace {

constexpr float kPi = 3.14159265359f;                                        // settings.h:40
constexpr float kMaxTranslation = 2.0f;                                      // settings.h:85-86
constexpr float kMaxTranslationSq = kMaxTranslation * kMaxTranslation;
constexpr float kMaxRotation = 0.5f * kPi;                                   // settings.h:90-91
constexpr float kMaxRotationSq = kMaxRotation * kMaxRotation;
constexpr float kMaxLinearCorrection = 0.2f;                                 // settings.h:95
constexpr float kBaumgarte = 0.1f;                                           // settings.h:101
constexpr float kVelocityThreshold = 1.0f;                                   // settings.h:105
constexpr float kTimeToSleep = 0.2f;                                         // settings.h:108
constexpr float kSleepLinearTol = 0.05f;                                     // settings.h:109
constexpr float kSleepAngularTol = 2.0f / 180.0f * kPi;                      // settings.h:110

// ---- b3Mat33 / b3Quat / b3Vec2 arithmetic in the reference's order (mat33.h, quat.h,
// vec2.h, mat22.h, math.cpp, time_step.h) ----
inline V3 scale3(float s, V3 v) { return V3{s * v.x, s * v.y, s * v.z}; }
inline M33 madd(const M33& A, const M33& B) { return M33{A.c0 + B.c0, A.c1 + B.c1, A.c2 + B.c2}; }
inline M33 msub(const M33& A, const M33& B) { return M33{A.c0 - B.c0, A.c1 - B.c1, A.c2 - B.c2}; }
inline M33 mscale(float s, const M33& A) { return M33{scale3(s, A.c0), scale3(s, A.c1), scale3(s, A.c2)}; }
inline M33 mmul(const M33& A, const M33& B) { return M33{mul(A, B.c0), mul(A, B.c1), mul(A, B.c2)}; }
inline M33 transpose(const M33& A) {
    return M33{V3{A.c0.x, A.c1.x, A.c2.x}, V3{A.c0.y, A.c1.y, A.c2.y}, V3{A.c0.z, A.c1.z, A.c2.z}};
}
inline M33 mzero() { return M33{V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}}; }
inline M33 diagonal(float x, float y, float z) { return M33{V3{x, 0, 0}, V3{0, y, 0}, V3{0, 0, z}}; }
inline M33 skew(V3 v) { return M33{V3{0.0f, v.z, -v.y}, V3{-v.z, 0.0f, v.x}, V3{v.y, -v.x, 0.0f}}; }
inline float det3(V3 a, V3 b, V3 c) { return dot(a, cross(b, c)); }                     // vec3.h:171-174
inline V3 solve33(const M33& A, V3 b) {                                                   // math.cpp:95-108
    float det = det3(A.c0, A.c1, A.c2);
    if (det != 0.0f) det = 1.0f / det;
    return V3{det * det3(b, A.c1, A.c2), det * det3(A.c0, b, A.c2), det * det3(A.c0, A.c1, b)};
}
inline M33 inverse33(const M33& A) {                                                      // math.cpp:110-133
    float det = det3(A.c0, A.c1, A.c2);
    if (det != 0.0f) det = 1.0f / det;
    const V3 c1 = cross(A.c1, A.c2), c2 = cross(A.c2, A.c0), c3 = cross(A.c0, A.c1);
    const M33 adj{V3{c1.x, c2.x, c3.x}, V3{c1.y, c2.y, c3.y}, V3{c1.z, c2.z, c3.z}};
    return mscale(det, adj);
}
inline M33 steiner(V3 v) {                                                                // time_step.h:58-80
    const float xx = v.x * v.x, yy = v.y * v.y, zz = v.z * v.z;
    M33 S;
    S.c0 = V3{yy + zz, -v.x * v.y, -v.x * v.z};
    S.c1 = V3{S.c0.y, xx + zz, -v.y * v.z};
    S.c2 = V3{S.c0.z, S.c1.z, xx + yy};
    return S;
}
inline M33 rotate_to_frame(const M33& I, const M33& R) { return mmul(mmul(R, I), transpose(R)); }  // time_step.h:86-90
inline M33 rotate_to_frame(const M33& I, Quat q) { return rotate_to_frame(I, quat_mat33(q.x, q.y, q.z, q.w)); }
inline Quat qadd(Quat a, Quat b) { return Quat{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline Quat qsub(Quat a, Quat b) { return Quat{a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline Quat qscale(float s, Quat q) { return Quat{s * q.x, s * q.y, s * q.z, s * q.w}; }
inline void qnormalize(Quat& q) {                                                         // quat.h:88-100
    const float len = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (len > kEpsilon) {
        const float s = 1.0f / len;
        q.x *= s, q.y *= s, q.z *= s, q.w *= s;
    }
}
inline V3 qrotate(Quat q, V3 v) {                                                         // quat.h:225-232
    const V3 qv{q.x, q.y, q.z};
    const V3 t = scale3(2.0f, cross(qv, v));
    return (v + scale3(q.w, t)) + cross(qv, t);
}
inline V3 qrotateT(Quat q, V3 v) { return qrotate(conj(q), v); }                          // quat.h:235-238
inline Quat derivative(Quat q, V3 w) {                                                    // time_step.h:105-111
    return qmul(Quat{0.5f * w.x, 0.5f * w.y, 0.5f * w.z, 0.0f}, q);
}
inline Quat integrate(Quat q, V3 w, float dt) { return qadd(q, qscale(dt, derivative(q, w))); }  // :114-120
inline float fmin_(float a, float b) { return a < b ? a : b; }                            // math.h:54-63
inline float fmax_(float a, float b) { return a > b ? a : b; }
inline float clampf(float a, float lo, float hi) { return fmax_(lo, fmin_(a, hi)); }     // math.h:66-69
inline V3 perp(V3 v) {                                                                    // vec3.h:255-267
    if (fabsf(v.x) >= 0.57735027f) return V3{v.y, -v.x, 0.0f};
    return V3{0.0f, v.z, -v.y};
}

// b3SolveGyro (island.cpp:88-107).
V3 solve_gyro(Quat q, const M33& Ib, V3 w1, float h) {
    V3 w1b = qrotateT(q, w1);
    const M33 J = madd(Ib, mscale(h, msub(mmul(skew(w1b), Ib), skew(mul(Ib, w1b)))));
    const V3 f = scale3(h, cross(w1b, mul(Ib, w1b)));
    w1b = w1b - solve33(J, f);
    return qrotate(q, w1b);
}

}  // namespace
}  // namespace b3amd

struct b3amd_host_world {
    struct Body {
        uint32_t type = 0;
        bool awake = true, island = false;
        uint32_t fixed = 0;
        float mass = 0, invMass = 0;
        M33 I = mzero(), invI = mzero(), worldInvI = mzero();
        V3 localCenter{0, 0, 0}, c0{0, 0, 0}, c{0, 0, 0};
        Quat q0{0, 0, 0, 1}, q{0, 0, 0, 1};
        float t0 = 0.0f;
        Xf xf;
        V3 v{0, 0, 0}, w{0, 0, 0}, force{0, 0, 0}, torque{0, 0, 0};
        float gravityScale = 1.0f, linearDamping = 0.0f, angularDamping = 0.0f, sleepTime = 0.0f;
        std::vector<uint32_t> shapes;  // creation order
    };
    struct Shape {
        uint32_t body, table;
        float density, friction, restitution, radius;
    };
    // Solver-owned state of a contact (b3Manifold impulses, manifold.h:66-93), keyed by the
    // contact's shapes and creation step; manifoldCount is b3Contact::m_manifoldCount.
    struct Contact {
        uint32_t manifoldCount = 0;
        b3amd_manifold m;
        float normalImpulse[4] = {0, 0, 0, 0};
        uint32_t persisting[4] = {0, 0, 0, 0};
        float tangentImpulse[2] = {0, 0};
        float motorImpulse = 0.0f;
        bool overlap = false, island = false;
        uint32_t shapeA = 0, shapeB = 0;
    };

    b3amd_context* ctx = nullptr;
    b3amd_broadphase* bp = nullptr;
    b3amd_contact_manager* cm = nullptr;
    V3 gravity{0.0f, -9.8f, 0.0f};
    bool warmStarting = true, sleeping = false, shapeAdded = false;
    std::vector<Body> bodies;
    std::vector<Shape> shapes;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, Contact> contacts;
    std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> list;  // m_contactList order
    // device copies of the body state
    void* dXf = nullptr;
    void* dQ = nullptr;
    void* dSweep = nullptr;
    void* dIds = nullptr;
    uint32_t cap = 0;
};

namespace b3amd {
namespace {

using HW = b3amd_host_world;

void synchronize_transform(HW::Body& b) {  // body.cpp:176-179: m_xf = m_sweep.GetTransform(1.0f)
    const float t = 1.0f, u = 1.0f - t;
    const V3 c{u * b.c0.x + t * b.c.x, u * b.c0.y + t * b.c.y, u * b.c0.z + t * b.c.z};
    Quat q = qadd(qscale(u, b.q0), qscale(t, b.q));
    qnormalize(q);
    b.xf.R = quat_mat33(q.x, q.y, q.z, q.w);
    b.xf.p = c - qrotate(q, b.localCenter);
}

void set_awake(HW::Body& b, bool flag) {  // body.h:521-540
    if (flag) {
        if (!b.awake) {
            b.awake = true;
            b.sleepTime = 0.0f;
        }
    } else {
        b.awake = false;
        b.sleepTime = 0.0f;
        b.force = b.torque = b.v = b.w = V3{0, 0, 0};
    }
}

// The host view of one hull record in the device arena (device_hull.h layout).
struct HostHull {
    std::vector<float4> units;
    HullLayout L;
    V3 vertex(uint32_t i) const { return V3{units[2 + i].x, units[2 + i].y, units[2 + i].z}; }
    uint32_t half(uint32_t h) const { return reinterpret_cast<const uint32_t*>(units.data() + L.halfOff())[h]; }
    uint32_t face_edge(uint32_t f) const { return reinterpret_cast<const uint16_t*>(units.data() + L.faceOff())[f]; }
};

int fetch_hull(b3amd_context* ctx, uint32_t hull, HostHull& out) {
    if (hull >= ctx->hullCount || ctx->hullOffsets[hull] == kFreeHull)
        return fail(B3AMD_ERR_ARGUMENT, "hull %u is not registered", hull);
    out.L = ctx->layouts[hull];
    out.units.resize(ctx->hullUnits[hull]);
    HIP_TRY(hipMemcpy(out.units.data(), ctx->dPool + ctx->hullOffsets[hull], 16ull * out.units.size(),
                      hipMemcpyDeviceToHost));
    return B3AMD_OK;
}

struct MassData {
    V3 center;
    float mass;
    M33 I;
};

// b3HullShape::ComputeMass (hull_shape.cpp:42-160).
MassData hull_mass(const HostHull& h, float density) {
    V3 s{0, 0, 0};
    for (uint32_t i = 0; i < h.L.V; ++i) s = s + h.vertex(i);
    {
        const float inv = 1.0f / (float)h.L.V;  // b3Vec3::operator/= (vec3.h:70-76)
        s = V3{s.x * inv, s.y * inv, s.z * inv};
    }
    float volume = 0.0f;
    V3 center{0, 0, 0};
    float xx = 0, xy = 0, yy = 0, xz = 0, zz = 0, yz = 0;
    for (uint32_t f = 0; f < h.L.F; ++f) {
        const uint32_t begin = h.face_edge(f);
        uint32_t edge = h.half(begin) >> 16;
        do {
            const uint32_t next = h.half(edge) >> 16;
            const V3 v1 = h.vertex(h.half(begin) & 0xFFu) - s;
            const V3 v2 = h.vertex(h.half(edge) & 0xFFu) - s;
            const V3 v3 = h.vertex(h.half(next) & 0xFFu) - s;
            const float D = det3(v1, v2, v3);
            volume += D;
            const V3 v4 = (v1 + v2) + v3;
            center = center + scale3(D, v4);
            xx += D * (((v1.x * v1.x + v2.x * v2.x) + v3.x * v3.x) + v4.x * v4.x);
            yy += D * (((v1.y * v1.y + v2.y * v2.y) + v3.y * v3.y) + v4.y * v4.y);
            zz += D * (((v1.z * v1.z + v2.z * v2.z) + v3.z * v3.z) + v4.z * v4.z);
            xy += D * (((v1.x * v1.y + v2.x * v2.y) + v3.x * v3.y) + v4.x * v4.y);
            xz += D * (((v1.x * v1.z + v2.x * v2.z) + v3.x * v3.z) + v4.x * v4.z);
            yz += D * (((v1.y * v1.z + v2.y * v2.z) + v3.y * v3.z) + v4.y * v4.z);
            edge = next;
        } while ((h.half(edge) >> 16) != begin);
    }
    M33 I;
    I.c0 = V3{yy + zz, -xy, -xz};
    I.c1 = V3{-xy, xx + zz, -yz};
    I.c2 = V3{-xz, -yz, xx + yy};
    MassData m;
    m.mass = density * volume / 6.0f;
    {
        const float inv = 1.0f / (4.0f * volume);
        center = V3{center.x * inv, center.y * inv, center.z * inv};
    }
    m.center = center + s;
    m.I = mscale(density / 120.0f, I);
    m.I = madd(m.I, mscale(m.mass, msub(steiner(m.center), steiner(center))));
    return m;
}

// b3SphereShape::ComputeMass (sphere_shape.cpp:39-51).
MassData sphere_mass(const ShapeRec& s, float density) {
    const float r = s.radius;
    const V3 c{s.v[0], s.v[1], s.v[2]};
    const float volume = (4.0f / 3.0f) * kPi * r * r * r;
    const float mass = density * volume;
    const float I = mass * (0.4f * r * r + dot(c, c));
    return MassData{c, mass, diagonal(I, I, I)};
}

// b3CapsuleShape::ComputeMass (capsule_shape.cpp:41-128).
MassData capsule_mass(const ShapeRec& s, float density) {
    const V3 A{s.v[0], s.v[1], s.v[2]}, B{s.v[3], s.v[4], s.v[5]};
    const V3 d = B - A;
    const float h = length(d);
    const float h2 = h * h;
    const float r = s.radius, r2 = r * r, r3 = r2 * r;
    const V3 center = scale3(0.5f, A + B);
    float mass = 0.0f;
    M33 I = mzero();
    M33 rotation;
    rotation.c1 = scale3(1.0f / h, d);
    rotation.c0 = perp(rotation.c1);
    rotation.c2 = cross(rotation.c1, rotation.c0);
    {
        const float cylinderVolume = kPi * r2 * h;
        const float cylinderMass = density * cylinderVolume;
        const float Ixx = (1.0f / 12.0f) * cylinderMass * (3.0f * r2 + h2);
        const float Iyy = 0.5f * cylinderMass * r2;
        M33 cylinderI = rotate_to_frame(diagonal(Ixx, Iyy, Ixx), rotation);
        cylinderI = madd(cylinderI, mscale(cylinderMass, steiner(center)));
        mass += cylinderMass;
        I = madd(I, cylinderI);
    }
    {
        const float hemiVolume = (2.0f / 3.0f) * kPi * r3;
        const float hemiMass = density * hemiVolume;
        const float Io = (2.0f / 5.0f) * hemiMass * r2;
        const float coy = (3.0f / 8.0f) * r;
        const float Iyy = Io - hemiMass * coy * coy;
        const float ccy = coy + 0.5f * h;
        const float Ixx = Io + hemiMass * ccy * ccy;
        M33 hemiI = rotate_to_frame(diagonal(Ixx, Iyy, Ixx), rotation);
        hemiI = madd(hemiI, mscale(hemiMass, steiner(center)));
        mass += 2.0f * hemiMass;
        I = madd(I, mscale(2.0f, hemiI));
    }
    return MassData{center, mass, I};
}

// b3Body::ResetMass (body.cpp:230-340).
int reset_mass(HW* w, HW::Body& b) {
    b.mass = 0.0f;
    b.invMass = 0.0f;
    b.I = b.invI = b.worldInvI = mzero();
    b.localCenter = V3{0, 0, 0};
    if (b.type == 0u || b.type == 1u) {
        b.c0 = b.xf.p;
        b.c = b.xf.p;
        b.q0 = b.q;
        return B3AMD_OK;
    }
    V3 localCenter{0, 0, 0};
    for (auto it = b.shapes.rbegin(); it != b.shapes.rend(); ++it) {  // m_shapeList: newest first
        const HW::Shape& s = w->shapes[*it];
        if (s.density == 0.0f) continue;
        const ShapeRec& rec = w->ctx->shapes[s.table];
        MassData md;
        if (rec.type == B3AMD_SPHERE) {
            md = sphere_mass(rec, s.density);
        } else if (rec.type == B3AMD_CAPSULE) {
            md = capsule_mass(rec, s.density);
        } else {
            HostHull h;
            int rc = fetch_hull(w->ctx, rec.hull, h);
            if (rc) return rc;
            md = hull_mass(h, s.density);
        }
        localCenter = localCenter + scale3(md.mass, md.center);
        b.mass += md.mass;
        b.I = madd(b.I, md.I);
    }
    if (b.mass > 0.0f) {
        b.invMass = 1.0f / b.mass;
        localCenter = scale3(b.invMass, localCenter);  // b3Vec3::operator*=
        b.I = msub(b.I, mscale(b.mass, steiner(localCenter)));
        b.invI = inverse33(b.I);
        b.worldInvI = rotate_to_frame(b.invI, b.xf.R);
        // Fixed rotation (body.cpp:286-323).
        auto zero = [](M33& m, int a, int c) { (&m.c0.x)[a * 3 + c] = 0.0f; };
        const int axes[3][4][2] = {{{1, 1}, {2, 1}, {1, 2}, {2, 2}}, {{0, 0}, {0, 2}, {2, 0}, {2, 2}},
                                   {{0, 0}, {0, 1}, {1, 0}, {1, 1}}};
        for (int a = 0; a < 3; ++a)
            if (b.fixed & (1u << a))
                for (int k = 0; k < 4; ++k) {
                    // m.x.y is column x, row y: element (row y, column x)
                    zero(b.invI, axes[a][k][1], axes[a][k][0]);
                    zero(b.worldInvI, axes[a][k][1], axes[a][k][0]);
                }
    } else {
        b.mass = 1.0f;
        b.invMass = 1.0f;
    }
    const V3 oldCenter = b.c;
    b.localCenter = localCenter;
    b.c = mul(b.xf, b.localCenter);
    b.c0 = b.c;
    b.v = b.v + cross(b.w, b.c - oldCenter);
    return B3AMD_OK;
}

int grow_device(HW* w) {
    const uint32_t n = (uint32_t)w->bodies.size();
    if (n <= w->cap) return B3AMD_OK;
    uint32_t cap = w->cap ? w->cap : 256u;
    while (cap < n) cap *= 2u;
    b3amd_context* ctx = w->ctx;
    ctx->release(w->dXf), ctx->release(w->dQ), ctx->release(w->dSweep), ctx->release(w->dIds);
    w->dXf = w->dQ = w->dSweep = w->dIds = nullptr;
    w->cap = 0;
    int rc;
    if ((rc = ctx->alloc(48ull * cap, &w->dXf)) || (rc = ctx->alloc(32ull * cap, &w->dQ)) ||
        (rc = ctx->alloc(72ull * cap, &w->dSweep)) || (rc = ctx->alloc(4ull * cap, &w->dIds)))
        return rc;
    w->cap = cap;
    return B3AMD_OK;
}

// Body transforms, orientation pairs and sweeps to the device, body flags to the manager.
int upload_bodies(HW* w) {
    int rc;
    if ((rc = grow_device(w))) return rc;
    const size_t n = w->bodies.size();
    std::vector<float> xf(12 * n), q(8 * n), sw(18 * n);
    std::vector<uint32_t> flags(n);
    for (size_t i = 0; i < n; ++i) {
        const HW::Body& b = w->bodies[i];
        std::memcpy(&xf[12 * i], &b.xf.R, 36);
        std::memcpy(&xf[12 * i + 9], &b.xf.p, 12);
        std::memcpy(&q[8 * i], &b.q0, 16);
        std::memcpy(&q[8 * i + 4], &b.q, 16);
        float* s = &sw[18 * i];
        std::memcpy(s, &b.localCenter, 12);
        std::memcpy(s + 3, &b.q0, 16);
        std::memcpy(s + 7, &b.c0, 12);
        s[10] = b.t0;
        std::memcpy(s + 11, &b.q, 16);
        std::memcpy(s + 15, &b.c, 12);
        flags[i] = b.type | (b.awake ? B3AMD_BODY_AWAKE : 0u);
    }
    HIP_TRY(hipMemcpy(w->dXf, xf.data(), 4 * xf.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(w->dQ, q.data(), 4 * q.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(w->dSweep, sw.data(), 4 * sw.size(), hipMemcpyHostToDevice));
    return b3amd_contact_manager_set_bodies(w->cm, (uint32_t)n, flags.data());
}

// Bodies woken on the device (AddPair, contact.cpp:94-99) wake on the host: SetAwake(true).
int download_awake(HW* w) {
    std::vector<uint32_t> flags(w->bodies.size());
    int rc = b3amd_contact_manager_get_bodies(w->cm, flags.data());
    if (rc) return rc;
    for (size_t i = 0; i < flags.size(); ++i)
        if (flags[i] & B3AMD_BODY_AWAKE) set_awake(w->bodies[i], true);
    return B3AMD_OK;
}

int find_new_contacts(HW* w) {
    int rc;
    if ((rc = upload_bodies(w))) return rc;
    uint32_t created = 0;
    if ((rc = b3amd_contact_manager_find_new_contacts(w->cm, &created))) return rc;
    return download_awake(w);
}

// UpdateContacts (contact_manager.cpp:167-231) on the device, then the end of
// b3Contact::Update for each contact that ran its narrowphase (contact.cpp:63-91).
int update_contacts(HW* w) {
    int rc;
    if ((rc = upload_bodies(w))) return rc;
    std::vector<bool> active(w->bodies.size());
    for (size_t i = 0; i < w->bodies.size(); ++i)
        active[i] = w->bodies[i].awake && w->bodies[i].type != 0u;
    uint32_t stats[3];
    if ((rc = b3amd_contact_manager_update_contacts(w->cm, (const b3amd_transform*)w->dXf, (const b3amd_quat*)w->dQ,
                                                    stats)))
        return rc;
    uint32_t n = 0;
    if ((rc = b3amd_contact_manager_read(w->cm, 0, &n, nullptr, nullptr, nullptr, nullptr))) return rc;
    std::vector<uint32_t> pairs(2 * (size_t)n + 2), born(n + 1);
    std::vector<b3amd_manifold> man(n + 1);
    if ((rc = b3amd_contact_manager_read(w->cm, n, &n, pairs.data(), born.data(), man.data(), nullptr))) return rc;
    if ((rc = download_awake(w))) return rc;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, HW::Contact> next;
    w->list.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t sa = pairs[2 * i], sb = pairs[2 * i + 1];
        const auto key = std::make_tuple(sa, sb, born[i] & 0x7FFFFFFFu);
        HW::Contact c;
        auto it = w->contacts.find(key);
        if (it != w->contacts.end()) c = it->second;
        c.shapeA = sa, c.shapeB = sb;
        const uint32_t ba = w->shapes[sa].body, bb = w->shapes[sb].body;
        if (active[ba] || active[bb]) {
            // b3Contact::Update: old manifolds, Initialize(), Collide, Initialize(old).
            const uint32_t oldCount = c.manifoldCount;
            const b3amd_manifold old = c.m;
            float oldNormal[4];
            std::memcpy(oldNormal, c.normalImpulse, sizeof(oldNormal));
            const float oldTangent[2] = {c.tangentImpulse[0], c.tangentImpulse[1]};
            const float oldMotor = c.motorImpulse;
            for (int k = 0; k < 4; ++k) c.normalImpulse[k] = 0.0f, c.persisting[k] = 0u;
            c.tangentImpulse[0] = c.tangentImpulse[1] = 0.0f;
            c.motorImpulse = 0.0f;
            c.m = man[i];
            c.manifoldCount = 1;
            if (w->warmStarting) {
                for (uint32_t j = 0; j < oldCount; ++j) {
                    c.tangentImpulse[0] = oldTangent[0], c.tangentImpulse[1] = oldTangent[1];
                    c.motorImpulse = oldMotor;
                    for (uint32_t p = 0; p < c.m.pointCount; ++p)
                        for (uint32_t q = 0; q < old.pointCount; ++q) {
                            const b3amd_manifold_point& a = c.m.points[p];
                            const b3amd_manifold_point& b = old.points[q];
                            if (a.triangleKey == b.triangleKey && a.key[0] == b.key[0] && a.key[1] == b.key[1] &&
                                a.key[2] == b.key[2] && a.key[3] == b.key[3]) {
                                c.normalImpulse[p] = oldNormal[q];
                                c.persisting[p] = 1u;
                                break;
                            }
                        }
                }
            }
            c.overlap = (born[i] >> 31) != 0u;
        }
        next[key] = c;
        w->list.push_back(key);
    }
    w->contacts.swap(next);
    return B3AMD_OK;
}

struct Position {
    V3 x;
    Quat q;
};
struct Velocity {
    V3 v, w;
};

struct VPoint {
    V3 rA, rB, normal;
    float normalMass, normalImpulse, velocityBias;
};
struct VManifold {
    VPoint points[4];
    uint32_t pointCount;
    V3 rA, rB, normal, tangent1, tangent2;
    float tangentMass[4];  // b3Mat22 columns x, y
    float tangentImpulse[2];
    float motorImpulse, motorMass;
};
struct VConstraint {
    uint32_t indexA, indexB;
    float invMassA, invMassB;
    M33 invIA, invIB;
    float friction, restitution;
    uint32_t manifoldCount;
    VManifold m;
};
struct PConstraint {
    uint32_t indexA, indexB;
    float invMassA, invMassB;
    M33 localInvIA, localInvIB;
    V3 localCenterA, localCenterB;
    float radiusA, radiusB;
    uint32_t manifoldCount, pointCount;
    V3 localNormalA[4], localPointA[4], localPointB[4];
};

// b3Island::Solve (island.cpp:109-341) with b3ContactSolver (contact_solver.cpp:70-588).
void island_solve(HW* w, const std::vector<uint32_t>& ib, const std::vector<HW::Contact*>& ic, float dt,
                  uint32_t velocityIterations, uint32_t positionIterations) {
    const float h = dt;
    const size_t nb = ib.size(), nc = ic.size();
    std::vector<Position> pos(nb);
    std::vector<Velocity> vel(nb);
    std::vector<M33> invI(nb);
    std::vector<uint32_t> islandId(w->bodies.size(), 0xFFFFFFFFu);
    for (size_t i = 0; i < nb; ++i) islandId[ib[i]] = (uint32_t)i;
    // 1. Integrate velocities.
    for (size_t i = 0; i < nb; ++i) {
        HW::Body& b = w->bodies[ib[i]];
        V3 v = b.v, wv = b.w;
        const V3 x = b.c;
        const Quat q = b.q;
        b.c0 = b.c;
        b.q0 = b.q;
        if (b.type == 2u) {
            v = v + scale3(h, scale3(b.gravityScale, w->gravity) + scale3(b.invMass, b.force));
            b.force = V3{0, 0, 0};
            const V3 dw1 = mul(mscale(h, b.worldInvI), b.torque);
            const V3 w2 = solve_gyro(q, b.I, wv, h);
            const V3 dw2 = w2 - wv;
            wv = wv + (dw1 + dw2);
            b.torque = V3{0, 0, 0};
            {
                const float s = 1.0f / (1.0f + h * b.linearDamping);
                v = V3{v.x * s, v.y * s, v.z * s};
            }
            {
                const float s = 1.0f / (1.0f + h * b.angularDamping);
                wv = V3{wv.x * s, wv.y * s, wv.z * s};
            }
        }
        vel[i] = Velocity{v, wv};
        pos[i] = Position{x, q};
        invI[i] = b.worldInvI;
    }
    // 2. Initialize constraints (contact_solver.cpp:70-276).
    std::vector<VConstraint> vc(nc);
    std::vector<PConstraint> pc(nc);
    for (size_t i = 0; i < nc; ++i) {
        HW::Contact* c = ic[i];
        const HW::Shape& sa = w->shapes[c->shapeA];
        const HW::Shape& sb = w->shapes[c->shapeB];
        const HW::Body& A = w->bodies[sa.body];
        const HW::Body& B = w->bodies[sb.body];
        PConstraint& p = pc[i];
        VConstraint& v = vc[i];
        p.indexA = islandId[sa.body], p.invMassA = A.invMass, p.localInvIA = A.invI, p.localCenterA = A.localCenter;
        p.radiusA = sa.radius;
        p.indexB = islandId[sb.body], p.invMassB = B.invMass, p.localInvIB = B.invI, p.localCenterB = B.localCenter;
        p.radiusB = sb.radius;
        p.manifoldCount = c->manifoldCount;
        v.indexA = p.indexA, v.invMassA = A.invMass, v.invIA = invI[v.indexA];
        v.indexB = p.indexB, v.invMassB = B.invMass, v.invIB = invI[v.indexB];
        v.friction = sqrtf(sa.friction * sb.friction);                                 // b3MixFriction
        v.restitution = sa.restitution > sb.restitution ? sa.restitution : sb.restitution;  // b3MixRestitution
        v.manifoldCount = c->manifoldCount;
        p.pointCount = c->m.pointCount;
        v.m.pointCount = c->m.pointCount;
        v.m.tangentImpulse[0] = c->tangentImpulse[0], v.m.tangentImpulse[1] = c->tangentImpulse[1];
        v.m.motorImpulse = c->motorImpulse;
        for (uint32_t k = 0; k < c->m.pointCount; ++k) {
            const b3amd_manifold_point& cp = c->m.points[k];
            p.localNormalA[k] = V3{cp.localNormal1.x, cp.localNormal1.y, cp.localNormal1.z};
            p.localPointA[k] = V3{cp.localPoint1.x, cp.localPoint1.y, cp.localPoint1.z};
            p.localPointB[k] = V3{cp.localPoint2.x, cp.localPoint2.y, cp.localPoint2.z};
            v.m.points[k].normalImpulse = c->normalImpulse[k];
        }
    }
    for (size_t i = 0; i < nc; ++i) {
        PConstraint& p = pc[i];
        VConstraint& vcs = vc[i];
        if (vcs.manifoldCount == 0) continue;
        const float mA = vcs.invMassA, mB = vcs.invMassB;
        const M33 iA = vcs.invIA, iB = vcs.invIB;
        const V3 xA = pos[vcs.indexA].x, xB = pos[vcs.indexB].x;
        const Quat qA = pos[vcs.indexA].q, qB = pos[vcs.indexB].q;
        const V3 vA = vel[vcs.indexA].v, wA = vel[vcs.indexA].w, vB = vel[vcs.indexB].v, wB = vel[vcs.indexB].w;
        Xf xfA, xfB;
        xfA.R = quat_mat33(qA.x, qA.y, qA.z, qA.w);
        xfA.p = xA - mul(xfA.R, p.localCenterA);
        xfB.R = quat_mat33(qB.x, qB.y, qB.z, qB.w);
        xfB.p = xB - mul(xfB.R, p.localCenterB);
        VManifold& m = vcs.m;
        // b3WorldManifold::Initialize (manifold.cpp:57-95).
        V3 center{0, 0, 0}, normal{0, 0, 0}, wp[4], wn[4];
        for (uint32_t k = 0; k < m.pointCount; ++k) {
            const V3 nA = mul(xfA.R, p.localNormalA[k]);
            const V3 cA = mul(xfA, p.localPointA[k]);
            const V3 cB = mul(xfB, p.localPointB[k]);
            const V3 pA = cA + scale3(p.radiusA, nA);
            const V3 pB = cB - scale3(p.radiusB, nA);
            wp[k] = scale3(0.5f, pA + pB);
            wn[k] = nA;
            center = center + wp[k];
            normal = normal + nA;
        }
        V3 t1{0, 0, 0}, t2{0, 0, 0};
        if (m.pointCount > 0) {
            const float inv = 1.0f / (float)m.pointCount;
            center = V3{center.x * inv, center.y * inv, center.z * inv};
            normalize_in_place(normal);
            t1 = perp(normal);
            t2 = cross(t1, normal);
        }
        m.normal = normal, m.tangent1 = t1, m.tangent2 = t2;
        for (uint32_t k = 0; k < m.pointCount; ++k) {
            VPoint& vp = m.points[k];
            const V3 n = wn[k];
            const V3 rA = wp[k] - xA, rB = wp[k] - xB;
            vp.rA = rA, vp.rB = rB, vp.normal = n;
            const V3 rnA = cross(rA, n), rnB = cross(rB, n);
            const float K = ((mA + mB) + dot(mul(iA, rnA), rnA)) + dot(mul(iB, rnB), rnB);
            vp.normalMass = K > 0.0f ? 1.0f / K : 0.0f;
            const V3 dv = ((vB + cross(wB, rB)) - vA) - cross(wA, rA);
            const float vn = dot(n, dv);
            vp.velocityBias = 0.0f;
            if (vn < -kVelocityThreshold) vp.velocityBias = -vcs.restitution * vn;
        }
        if (m.pointCount > 0) {
            const V3 rA = center - xA, rB = center - xB;
            m.rA = rA, m.rB = rB;
            const V3 rn1A = cross(rA, t1), rn1B = cross(rB, t1), rn2A = cross(rA, t2), rn2B = cross(rB, t2);
            const float k11 = ((mA + mB) + dot(mul(iA, rn1A), rn1A)) + dot(mul(iB, rn1B), rn1B);
            const float k12 = dot(mul(iA, rn1A), rn2A) + dot(mul(iB, rn1B), rn2B);
            const float k22 = ((mA + mB) + dot(mul(iA, rn2A), rn2A)) + dot(mul(iB, rn2B), rn2B);
            float det = k11 * k22 - k12 * k12;  // b3Inverse(b3Mat22) (math.cpp:77-93)
            if (det != 0.0f) det = 1.0f / det;
            m.tangentMass[0] = det * k22, m.tangentMass[1] = -det * k12;
            m.tangentMass[2] = -det * k12, m.tangentMass[3] = det * k11;
            const float mass = dot(m.normal, mul(madd(iA, iB), m.normal));
            m.motorMass = mass > 0.0f ? 1.0f / mass : 0.0f;
        }
    }
    // Warm start (contact_solver.cpp:278-331).
    if (w->warmStarting) {
        for (size_t i = 0; i < nc; ++i) {
            const VConstraint& vcs = vc[i];
            const float mA = vcs.invMassA, mB = vcs.invMassB;
            const M33 iA = vcs.invIA, iB = vcs.invIB;
            V3 vA = vel[vcs.indexA].v, wA = vel[vcs.indexA].w, vB = vel[vcs.indexB].v, wB = vel[vcs.indexB].w;
            for (uint32_t j = 0; j < vcs.manifoldCount; ++j) {
                const VManifold& m = vcs.m;
                for (uint32_t k = 0; k < m.pointCount; ++k) {
                    const VPoint& vp = m.points[k];
                    const V3 P = scale3(vp.normalImpulse, vp.normal);
                    vA = vA - scale3(mA, P);
                    wA = wA - mul(iA, cross(vp.rA, P));
                    vB = vB + scale3(mB, P);
                    wB = wB + mul(iB, cross(vp.rB, P));
                }
                if (m.pointCount > 0) {
                    const V3 P1 = scale3(m.tangentImpulse[0], m.tangent1);
                    const V3 P2 = scale3(m.tangentImpulse[1], m.tangent2);
                    const V3 P3 = scale3(m.motorImpulse, m.normal);
                    vA = vA - scale3(mA, P1 + P2);
                    wA = wA - mul(iA, cross(m.rA, P1 + P2) + P3);
                    vB = vB + scale3(mB, P1 + P2);
                    wB = wB + mul(iB, cross(m.rB, P1 + P2) + P3);
                }
            }
            vel[vcs.indexA] = Velocity{vA, wA};
            vel[vcs.indexB] = Velocity{vB, wB};
        }
    }
    // 3. Velocity iterations (contact_solver.cpp:333-451).
    for (uint32_t it = 0; it < velocityIterations; ++it) {
        for (size_t i = 0; i < nc; ++i) {
            VConstraint& vcs = vc[i];
            const float mA = vcs.invMassA, mB = vcs.invMassB;
            const M33 iA = vcs.invIA, iB = vcs.invIB;
            V3 vA = vel[vcs.indexA].v, wA = vel[vcs.indexA].w, vB = vel[vcs.indexB].v, wB = vel[vcs.indexB].w;
            for (uint32_t j = 0; j < vcs.manifoldCount; ++j) {
                VManifold& m = vcs.m;
                float normalImpulse = 0.0f;
                for (uint32_t k = 0; k < m.pointCount; ++k) {
                    VPoint& vp = m.points[k];
                    const V3 dv = ((vB + cross(wB, vp.rB)) - vA) - cross(wA, vp.rA);
                    const float Cdot = dot(vp.normal, dv);
                    float impulse = -vp.normalMass * (Cdot - vp.velocityBias);
                    const float oldImpulse = vp.normalImpulse;
                    vp.normalImpulse = fmax_(vp.normalImpulse + impulse, 0.0f);
                    impulse = vp.normalImpulse - oldImpulse;
                    const V3 P = scale3(impulse, vp.normal);
                    vA = vA - scale3(mA, P);
                    wA = wA - mul(iA, cross(vp.rA, P));
                    vB = vB + scale3(mB, P);
                    wB = wB + mul(iB, cross(vp.rB, P));
                    normalImpulse += vp.normalImpulse;
                }
                if (m.pointCount > 0) {
                    {
                        const V3 dv = ((vB + cross(wB, m.rB)) - vA) - cross(wA, m.rA);
                        const float cx = dot(dv, m.tangent1), cy = dot(dv, m.tangent2);
                        // b3Mat22 * b3Vec2 (mat22.h:56-59) of -Cdot.
                        const float nx = -cx, ny = -cy;
                        const float ix = nx * m.tangentMass[0] + ny * m.tangentMass[2];
                        const float iy = nx * m.tangentMass[1] + ny * m.tangentMass[3];
                        const float ox = m.tangentImpulse[0], oy = m.tangentImpulse[1];
                        m.tangentImpulse[0] += ix;
                        m.tangentImpulse[1] += iy;
                        const float maxImpulse = vcs.friction * normalImpulse;
                        if (m.tangentImpulse[0] * m.tangentImpulse[0] + m.tangentImpulse[1] * m.tangentImpulse[1] >
                            maxImpulse * maxImpulse) {
                            // b3Vec2::Normalize (vec2.h:80-89), then *= maxImpulse
                            const float s = sqrtf(m.tangentImpulse[0] * m.tangentImpulse[0] +
                                                  m.tangentImpulse[1] * m.tangentImpulse[1]);
                            if (s > kEpsilon) {
                                m.tangentImpulse[0] /= s;
                                m.tangentImpulse[1] /= s;
                            }
                            m.tangentImpulse[0] *= maxImpulse;
                            m.tangentImpulse[1] *= maxImpulse;
                        }
                        const float jx = m.tangentImpulse[0] - ox, jy = m.tangentImpulse[1] - oy;
                        const V3 P1 = scale3(jx, m.tangent1), P2 = scale3(jy, m.tangent2);
                        const V3 P = P1 + P2;
                        vA = vA - scale3(mA, P);
                        wA = wA - mul(iA, cross(m.rA, P));
                        vB = vB + scale3(mB, P);
                        wB = wB + mul(iB, cross(m.rB, P));
                    }
                    {
                        const float Cdot = dot(m.normal, wB - wA);
                        float impulse = -m.motorMass * Cdot;
                        const float oldImpulse = m.motorImpulse;
                        const float maxImpulse = vcs.friction * normalImpulse;
                        m.motorImpulse = clampf(m.motorImpulse + impulse, -maxImpulse, maxImpulse);
                        impulse = m.motorImpulse - oldImpulse;
                        const V3 P = scale3(impulse, m.normal);
                        wA = wA - mul(iA, P);
                        wB = wB + mul(iB, P);
                    }
                }
            }
            vel[vcs.indexA] = Velocity{vA, wA};
            vel[vcs.indexB] = Velocity{vB, wB};
        }
    }
    // StoreImpulses (contact_solver.cpp:453-479).
    if (w->warmStarting) {
        for (size_t i = 0; i < nc; ++i) {
            HW::Contact* c = ic[i];
            const VManifold& m = vc[i].m;
            if (c->manifoldCount == 0) continue;
            c->tangentImpulse[0] = m.tangentImpulse[0], c->tangentImpulse[1] = m.tangentImpulse[1];
            c->motorImpulse = m.motorImpulse;
            for (uint32_t k = 0; k < c->m.pointCount; ++k) c->normalImpulse[k] = m.points[k].normalImpulse;
        }
    }
    // 4. Integrate positions.
    for (size_t i = 0; i < nb; ++i) {
        const HW::Body& b = w->bodies[ib[i]];
        V3 x = pos[i].x;
        Quat q = pos[i].q;
        V3 v = vel[i].v, wv = vel[i].w;
        const V3 translation = scale3(h, v);
        if (dot(translation, translation) > kMaxTranslationSq) {
            const float ratio = kMaxTranslation / length(translation);
            v = V3{v.x * ratio, v.y * ratio, v.z * ratio};
        }
        const V3 rotation = scale3(h, wv);
        if (dot(rotation, rotation) > kMaxRotationSq) {
            const float ratio = kMaxRotation / length(rotation);
            wv = V3{wv.x * ratio, wv.y * ratio, wv.z * ratio};
        }
        x = x + scale3(h, v);
        q = integrate(q, wv, h);
        pos[i] = Position{x, q};
        vel[i] = Velocity{v, wv};
        invI[i] = rotate_to_frame(b.invI, q);
    }
    // 5. Position iterations (contact_solver.cpp:496-588).
    for (uint32_t it = 0; it < positionIterations; ++it) {
        float minSeparation = 0.0f;
        for (size_t i = 0; i < nc; ++i) {
            const PConstraint& p = pc[i];
            const float mA = p.invMassA, mB = p.invMassB;
            V3 cA = pos[p.indexA].x, cB = pos[p.indexB].x;
            Quat qA = pos[p.indexA].q, qB = pos[p.indexB].q;
            M33 iA = invI[p.indexA], iB = invI[p.indexB];
            for (uint32_t j = 0; j < p.manifoldCount; ++j) {
                for (uint32_t k = 0; k < p.pointCount; ++k) {
                    Xf xfA, xfB;
                    xfA.R = quat_mat33(qA.x, qA.y, qA.z, qA.w);
                    xfA.p = cA - mul(xfA.R, p.localCenterA);
                    xfB.R = quat_mat33(qB.x, qB.y, qB.z, qB.w);
                    xfB.p = cB - mul(xfB.R, p.localCenterB);
                    const V3 normal = mul(xfA.R, p.localNormalA[k]);
                    const V3 c1 = mul(xfA, p.localPointA[k]);
                    const V3 c2 = mul(xfB, p.localPointB[k]);
                    const V3 point = c2;
                    const float separation = (dot(c2 - c1, normal) - p.radiusA) - p.radiusB;
                    minSeparation = fmin_(minSeparation, separation);
                    const float C = clampf(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, 0.0f);
                    const V3 rA = point - cA, rB = point - cB;
                    const V3 rnA = cross(rA, normal), rnB = cross(rB, normal);
                    const float K = ((mA + mB) + dot(rnA, mul(iA, rnA))) + dot(rnB, mul(iB, rnB));
                    const float impulse = K > 0.0f ? -C / K : 0.0f;
                    const V3 P = scale3(impulse, normal);
                    cA = cA - scale3(mA, P);
                    qA = qsub(qA, derivative(qA, mul(iA, cross(rA, P))));
                    qnormalize(qA);
                    iA = rotate_to_frame(p.localInvIA, qA);
                    cB = cB + scale3(mB, P);
                    qB = qadd(qB, derivative(qB, mul(iB, cross(rB, P))));
                    qnormalize(qB);
                    iB = rotate_to_frame(p.localInvIB, qB);
                }
            }
            pos[p.indexA].x = cA, pos[p.indexA].q = qA, invI[p.indexA] = iA;
            pos[p.indexB].x = cB, pos[p.indexB].q = qB, invI[p.indexB] = iB;
        }
        // No joints: the joint solver reports solved, so the contacts decide the early out.
        if (minSeparation >= -3.0f * kLinearSlop) break;
    }
    // 6. Copy state back.
    for (size_t i = 0; i < nb; ++i) {
        HW::Body& b = w->bodies[ib[i]];
        b.c = pos[i].x;
        b.q = pos[i].q;
        qnormalize(b.q);
        b.v = vel[i].v;
        b.w = vel[i].w;
        b.worldInvI = invI[i];
        synchronize_transform(b);
    }
    // 7. Sleep.
    if (w->sleeping) {
        float minSleepTime = kMaxFloat;
        for (size_t i = 0; i < nb; ++i) {
            HW::Body& b = w->bodies[ib[i]];
            if (b.type == 0u) continue;
            const float sqrLinVel = dot(b.v, b.v), sqrAngVel = dot(b.w, b.w);
            if (sqrLinVel > kSleepLinearTol || sqrAngVel > kSleepAngularTol) {
                minSleepTime = 0.0f;
                b.sleepTime = 0.0f;
            } else {
                b.sleepTime += h;
                minSleepTime = fmin_(minSleepTime, b.sleepTime);
            }
        }
        if (minSleepTime >= kTimeToSleep)
            for (size_t i = 0; i < nb; ++i) set_awake(w->bodies[ib[i]], false);
    }
}

// b3World::Solve (world.cpp:146-337).
int solve(HW* w, float dt, uint32_t vi, uint32_t pi) {
    const size_t B = w->bodies.size();
    for (HW::Body& b : w->bodies) b.island = false;
    for (auto& kv : w->contacts) kv.second.island = false;
    // A shape's contact edges: the contacts of the list that involve it, in list order.
    std::vector<std::vector<HW::Contact*>> edges(w->shapes.size());
    for (const auto& key : w->list) {
        HW::Contact* c = &w->contacts[key];
        edges[c->shapeA].push_back(c);
        edges[c->shapeB].push_back(c);
    }
    std::vector<uint32_t> stack;
    stack.reserve(B);
    std::vector<uint32_t> ib;
    std::vector<HW::Contact*> ic;
    for (size_t r = B; r-- > 0;) {  // m_bodyList: newest first
        const uint32_t seed = (uint32_t)r;
        HW::Body& s = w->bodies[seed];
        if (s.island || !s.awake || s.type == 0u) continue;
        ib.clear();
        ic.clear();
        stack.clear();
        stack.push_back(seed);
        s.island = true;
        while (!stack.empty()) {
            const uint32_t bi = stack.back();
            stack.pop_back();
            HW::Body& b = w->bodies[bi];
            ib.push_back(bi);
            b.awake = true;
            if (b.type == 0u) continue;
            for (auto si = b.shapes.rbegin(); si != b.shapes.rend(); ++si) {  // m_shapeList: newest first
                for (HW::Contact* c : edges[*si]) {
                    if (c->island || !c->overlap) continue;
                    ic.push_back(c);
                    c->island = true;
                    const uint32_t other = w->shapes[c->shapeA].body == bi && c->shapeA == *si
                                               ? w->shapes[c->shapeB].body
                                               : w->shapes[c->shapeA].body;
                    HW::Body& o = w->bodies[other];
                    if (o.island) continue;
                    stack.push_back(other);
                    o.island = true;
                }
            }
        }
        island_solve(w, ib, ic, dt, vi, pi);
        for (uint32_t bi : ib)
            if (w->bodies[bi].type == 0u) w->bodies[bi].island = false;
    }
    // SynchronizeShapes of the bodies that were in an island (world.cpp:308-336).
    std::vector<uint32_t> moved;
    for (size_t bi = 0; bi < B; ++bi) {
        const HW::Body& b = w->bodies[bi];
        if (!b.island || b.type == 0u) continue;
        for (uint32_t si : b.shapes) moved.push_back(si);
    }
    int rc;
    if ((rc = upload_bodies(w))) return rc;
    if (!moved.empty()) {
        HIP_TRY(hipMemcpy(w->dIds, moved.data(), 4 * moved.size(), hipMemcpyHostToDevice));
        if ((rc = b3amd_contact_manager_synchronize(w->cm, (uint32_t)moved.size(), (const uint32_t*)w->dIds,
                                                    (const b3amd_transform*)w->dXf, (const b3amd_sweep*)w->dSweep)))
            return rc;
    }
    return find_new_contacts(w);
}

}  // namespace
}  // namespace b3amd

extern "C" {

int b3amd_host_world_create(b3amd_context* ctx, b3amd_host_world** out) {
    if (!ctx || !out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_host_world* w = new b3amd_host_world();
    w->ctx = ctx;
    int rc;
    if ((rc = b3amd_broadphase_create(ctx, &w->bp)) || (rc = b3amd_contact_manager_create(ctx, w->bp, &w->cm))) {
        b3amd_host_world_destroy(w);
        return rc;
    }
    *out = w;
    return B3AMD_OK;
}

int b3amd_host_world_destroy(b3amd_host_world* w) {
    if (!w) return B3AMD_OK;
    b3amd_contact_manager_destroy(w->cm);
    b3amd_broadphase_destroy(w->bp);
    w->ctx->release(w->dXf), w->ctx->release(w->dQ), w->ctx->release(w->dSweep), w->ctx->release(w->dIds);
    delete w;
    return B3AMD_OK;
}

int b3amd_host_world_settings(b3amd_host_world* w, const b3amd_vec3* gravity, int warm_starting, int sleeping) {
    if (!w) return fail(B3AMD_ERR_ARGUMENT, "null world");
    if (gravity) w->gravity = V3{gravity->x, gravity->y, gravity->z};
    w->warmStarting = warm_starting != 0;
    w->sleeping = sleeping != 0;
    return B3AMD_OK;
}

int b3amd_host_world_create_body(b3amd_host_world* w, const b3amd_body_def* def, uint32_t* body) {
    if (!w || !def || !body) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (def->type > 2u) return fail(B3AMD_ERR_ARGUMENT, "body type %u", def->type);
    b3amd_host_world::Body b;  // b3Body::b3Body (body.cpp:25-86)
    b.type = def->type;
    b.awake = def->awake != 0;
    b.fixed = def->fixed_rotation & 7u;
    if (b.type == 2u) b.mass = b.invMass = 1.0f;
    b.v = V3{def->linear_velocity.x, def->linear_velocity.y, def->linear_velocity.z};
    b.w = V3{def->angular_velocity.x, def->angular_velocity.y, def->angular_velocity.z};
    b.c = b.c0 = V3{def->position.x, def->position.y, def->position.z};
    b.q = b.q0 = Quat{def->orientation.x, def->orientation.y, def->orientation.z, def->orientation.w};
    b.xf.p = b.c;
    b.xf.R = quat_mat33(b.q.x, b.q.y, b.q.z, b.q.w);
    b.linearDamping = def->linear_damping;
    b.angularDamping = def->angular_damping;
    b.gravityScale = def->gravity_scale;
    *body = (uint32_t)w->bodies.size();
    w->bodies.push_back(b);
    return B3AMD_OK;
}

int b3amd_host_world_create_shape(b3amd_host_world* w, uint32_t body, const b3amd_shape_def* def, uint32_t* shape) {
    if (!w || !def) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (body >= w->bodies.size()) return fail(B3AMD_ERR_ARGUMENT, "body %u does not exist", body);
    if (def->table_shape >= w->ctx->shapeCount) return fail(B3AMD_ERR_ARGUMENT, "shape %u beyond the table", def->table_shape);
    if (def->sensor) return fail(B3AMD_ERR_UNSUPPORTED, "sensor shapes are not stepped here");
    const ShapeRec& rec = w->ctx->shapes[def->table_shape];
    if (rec.type == B3AMD_MESH) return fail(B3AMD_ERR_UNSUPPORTED, "mesh shapes are not stepped here");
    // b3Body::CreateShape (body.cpp:89-121): add, ResetMass when dense, then the proxy.
    const uint32_t s = (uint32_t)w->shapes.size();
    w->shapes.push_back(b3amd_host_world::Shape{body, def->table_shape, def->density, def->friction, def->restitution,
                                                rec.radius});
    b3amd_host_world::Body& b = w->bodies[body];
    b.shapes.push_back(s);
    int rc;
    if (def->density > 0.0f && (rc = reset_mass(w, b))) return rc;
    if ((rc = upload_bodies(w))) return rc;
    uint32_t proxy;
    if ((rc = b3amd_contact_manager_add_shapes(w->cm, 1, &def->table_shape, &body, (const b3amd_transform*)w->dXf,
                                               &proxy)))
        return rc;
    w->shapeAdded = true;
    if (shape) *shape = s;
    return B3AMD_OK;
}

int b3amd_host_world_step(b3amd_host_world* w, float dt, uint32_t velocity_iterations, uint32_t position_iterations) {
    if (!w) return fail(B3AMD_ERR_ARGUMENT, "null world");
    int rc;
    // b3World::Step (world.cpp:115-144).
    if (w->shapeAdded) {
        if ((rc = find_new_contacts(w))) return rc;
        w->shapeAdded = false;
    }
    if ((rc = update_contacts(w))) return rc;
    if (dt > 0.0f) return solve(w, dt, velocity_iterations, position_iterations);
    return B3AMD_OK;
}

int b3amd_host_world_bodies(b3amd_host_world* w, uint32_t first, uint32_t count, b3amd_body_state* out) {
    if (!w || (count && !out)) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if ((uint64_t)first + count > w->bodies.size()) return fail(B3AMD_ERR_ARGUMENT, "range beyond the bodies");
    for (uint32_t i = 0; i < count; ++i) {
        const b3amd_host_world::Body& b = w->bodies[first + i];
        b3amd_body_state& o = out[i];
        std::memcpy(&o.xf.rotation, &b.xf.R, 36);
        std::memcpy(&o.xf.position, &b.xf.p, 12);
        std::memcpy(&o.sweep.localCenter, &b.localCenter, 12);
        std::memcpy(&o.sweep.orientation0, &b.q0, 16);
        std::memcpy(&o.sweep.worldCenter0, &b.c0, 12);
        o.sweep.t0 = b.t0;
        std::memcpy(&o.sweep.orientation, &b.q, 16);
        std::memcpy(&o.sweep.worldCenter, &b.c, 12);
        std::memcpy(&o.linear_velocity, &b.v, 12);
        std::memcpy(&o.angular_velocity, &b.w, 12);
        o.mass = b.mass;
        o.inv_mass = b.invMass;
        std::memcpy(o.inertia, &b.I, 36);
        std::memcpy(o.inv_inertia, &b.invI, 36);
        std::memcpy(o.world_inv_inertia, &b.worldInvI, 36);
        o.awake = b.awake ? 1u : 0u;
        o.sleep_time = b.sleepTime;
    }
    return B3AMD_OK;
}

int b3amd_host_world_contact_count(b3amd_host_world* w, uint32_t* count) {
    if (!w || !count) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    *count = (uint32_t)w->list.size();
    return B3AMD_OK;
}

}  // extern "C"
