// FP32 vector algebra in the exact association order of the reference math layer.
// Every expression is written out so that, compiled with -ffp-contract=off and IEEE
// division/square root, it rounds identically to the reference on x86-64 SSE2:
//   dot   = (x*x' + y*y') + z*z'                       vec3.h:158-161
//   cross = (y*z' - z*y', z*x' - x*z', x*y' - y*x')      vec3.h:164-167
//   R*v   = (v.x*c0 + v.y*c1) + v.z*c2  (columns)        mat33.h:143-146
//   R^T*v = (dot(c0,v), dot(c1,v), dot(c2,v))            mat33.h:157-160
#ifndef B3AMD_MATH_H
#define B3AMD_MATH_H

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

#define B3_HD __host__ __device__ __forceinline__

namespace b3amd {

constexpr float kLinearSlop = 0.005f;          // settings.h:62
constexpr float kMaxFloat = FLT_MAX;           // settings.h:41
constexpr float kEpsilon = FLT_EPSILON;        // settings.h:42
constexpr uint32_t kNullEdge = 0xFFFFFFFFu;    // clip.h:25
constexpr uint32_t kNullTriangle = 0xFFFFFFFFu;// manifold.h:25

struct V3 {
    float x, y, z;
};
struct Plane {
    V3 n;
    float d;
};
struct M33 {
    V3 c0, c1, c2;  // columns
};
struct Xf {
    M33 R;
    V3 p;
};
struct Quat {
    float x, y, z, w;
};

B3_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
B3_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
B3_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
B3_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
B3_HD V3 operator*(float s, V3 v) { return V3{s * v.x, s * v.y, s * v.z}; }
B3_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
B3_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
B3_HD float length(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
// b3Normalize (vec3.h:189-198): scale by 1/len only when len > epsilon.
B3_HD V3 normalize(V3 v) {
    float len = length(v);
    if (len > kEpsilon) {
        float s = 1.0f / len;
        return s * v;
    }
    return v;
}
// b3Vec3::Normalize (vec3.h:93-104): in place, returns the length.
B3_HD float normalize_in_place(V3& v) {
    float len = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (len > kEpsilon) {
        float s = 1.0f / len;
        v.x *= s;
        v.y *= s;
        v.z *= s;
    }
    return len;
}

B3_HD V3 mul(const M33& A, V3 v) { return (v.x * A.c0 + v.y * A.c1) + v.z * A.c2; }
B3_HD V3 mulT(const M33& A, V3 v) { return V3{dot(A.c0, v), dot(A.c1, v), dot(A.c2, v)}; }
B3_HD M33 mulT(const M33& A, const M33& B) {  // mat33.h:163-169
    return M33{V3{dot(A.c0, B.c0), dot(A.c1, B.c0), dot(A.c2, B.c0)},
               V3{dot(A.c0, B.c1), dot(A.c1, B.c1), dot(A.c2, B.c1)},
               V3{dot(A.c0, B.c2), dot(A.c1, B.c2), dot(A.c2, B.c2)}};
}
B3_HD V3 mul(const Xf& T, V3 v) { return mul(T.R, v) + T.p; }            // transform.h:74-77
B3_HD V3 mulT(const Xf& T, V3 v) { return mulT(T.R, v - T.p); }          // transform.h:106-110
B3_HD Xf mulT(const Xf& A, const Xf& B) {                                 // transform.h:88-94
    return Xf{mulT(A.R, B.R), mulT(A.R, B.p - A.p)};
}
B3_HD Plane mul(const Xf& T, const Plane& p) {                            // geometry.h:77-81
    V3 n = mul(T.R, p.n);
    return Plane{n, p.d + dot(n, T.p)};
}
B3_HD float distance(V3 P, const Plane& p) { return dot(p.n, P) - p.d; } // geometry.h:90-93
B3_HD V3 closest_on_plane(V3 P, const Plane& p) {                         // geometry.h:96-100
    float f = distance(P, p);
    return P - f * p.n;
}
// b3BarycentricCoordinates for a segment (geometry.h:104-116).
B3_HD void barycentric(float out[3], V3 A, V3 B, V3 Q) {
    V3 AB = B - A, QA = A - Q, QB = B - Q;
    float divisor = dot(AB, AB);
    out[0] = dot(QB, AB);
    out[1] = -dot(QA, AB);
    out[2] = divisor;
}

B3_HD Quat qmul(Quat a, Quat b) {  // quat.h:172-179
    return Quat{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
                a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z,
                a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x,
                a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
B3_HD Quat conj(Quat q) { return Quat{-q.x, -q.y, -q.z, q.w}; }  // quat.h:195-198

// b3QuatMat33 (quat.h:306-318).
B3_HD M33 quat_mat33(float x, float y, float z, float w) {
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx = x * x2, xy = x * y2, xz = x * z2;
    const float yy = y * y2, yz = y * z2, zz = z * z2;
    const float wx = w * x2, wy = w * y2, wz = w * z2;
    M33 m;
    m.c0 = V3{1.0f - (yy + zz), xy + wz, xz - wy};
    m.c1 = V3{xy - wz, 1.0f - (xx + zz), yz + wx};
    m.c2 = V3{xz + wy, yz - wx, 1.0f - (xx + yy)};
    return m;
}

}  // namespace b3amd

#endif
