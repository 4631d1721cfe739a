// Seeded synthetic workloads for the narrowphase.  Every random quantity comes from
// a counter-based splitmix64 stream and only IEEE + - * / sqrt in double, so the same
// seed yields bit-identical records on every x86-64 host.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "bounce_amd.h"
#include "hullgen.h"
#include "workload.h"

namespace b3amd {

uint64_t SplitMix::next() {
    uint64_t z = (state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double SplitMix::u01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }

double SplitMix::normal() {
    // Irwin-Hall(12): sum of 12 uniforms minus 6 (mean 0, variance 1).
    double s = 0.0;
    for (int i = 0; i < 12; ++i) s += u01();
    return s - 6.0;
}

void SplitMix::unit_vector(double v[3]) {
    for (;;) {
        double x = 2.0 * u01() - 1.0, y = 2.0 * u01() - 1.0, z = 2.0 * u01() - 1.0;
        double r2 = x * x + y * y + z * z;
        if (r2 > 1e-4 && r2 <= 1.0) {
            double r = std::sqrt(r2);
            v[0] = x / r;
            v[1] = y / r;
            v[2] = z / r;
            return;
        }
    }
}

void SplitMix::unit_quat(double q[4]) {
    for (;;) {
        double a = 2.0 * u01() - 1.0, b = 2.0 * u01() - 1.0, c = 2.0 * u01() - 1.0,
               d = 2.0 * u01() - 1.0;
        double r2 = a * a + b * b + c * c + d * d;
        if (r2 > 1e-4 && r2 <= 1.0) {
            double r = std::sqrt(r2);
            q[0] = a / r;
            q[1] = b / r;
            q[2] = c / r;
            q[3] = d / r;
            return;
        }
    }
}

SplitMix stream(uint64_t seed, uint64_t index) {
    SplitMix s{seed ^ (index * 0xD1B54A32D192ED03ull)};
    s.next();
    return s;
}

void quat_to_rotation(const double q[4], float R[9]) {
    // q = (x, y, z, w); columns of the rotation matrix.
    double x = q[0], y = q[1], z = q[2], w = q[3];
    double xx = x * x, yy = y * y, zz = z * z;
    double xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
    R[0] = (float)(1.0 - 2.0 * (yy + zz));
    R[1] = (float)(2.0 * (xy + wz));
    R[2] = (float)(2.0 * (xz - wy));
    R[3] = (float)(2.0 * (xy - wz));
    R[4] = (float)(1.0 - 2.0 * (xx + zz));
    R[5] = (float)(2.0 * (yz + wx));
    R[6] = (float)(2.0 * (xz + wy));
    R[7] = (float)(2.0 * (yz - wx));
    R[8] = (float)(1.0 - 2.0 * (xx + yy));
}

int ellipsoid_hull(SplitMix& rng, int n, double a, double b, double c, HullRecord* out) {
    std::vector<float> pts(3 * n);
    for (int attempt = 0; attempt < 8; ++attempt) {
        for (int i = 0; i < n; ++i) {
            double d[3];
            rng.unit_vector(d);
            pts[3 * i + 0] = (float)(a * d[0]);
            pts[3 * i + 1] = (float)(b * d[1]);
            pts[3 * i + 2] = (float)(c * d[2]);
        }
        // b3QHull::Set without simplification: the hull keeps every point in convex position.
        bool orphans = false;
        int v = quickhull_record(pts.data(), (uint32_t)n, false, out, &orphans);
        if (v >= 4 && !orphans) return v;
    }
    return -1;
}

namespace {

struct WorldHull {
    std::vector<double> v;   // [V][3] rotated vertices (body origin at 0)
    std::vector<double> n;   // [F][3] rotated face normals
    double c[3];             // rotated centroid
};

WorldHull rotate_hull(const b3amd_hull_pool* pool, uint32_t h, const float R[9]) {
    const uint32_t* d = pool->desc + 6 * (size_t)h;
    WorldHull w;
    auto rotate = [&](const float* x, double* o) {
        o[0] = (double)R[0] * x[0] + (double)R[3] * x[1] + (double)R[6] * x[2];
        o[1] = (double)R[1] * x[0] + (double)R[4] * x[1] + (double)R[7] * x[2];
        o[2] = (double)R[2] * x[0] + (double)R[5] * x[1] + (double)R[8] * x[2];
    };
    w.v.resize(3 * d[1]);
    for (uint32_t k = 0; k < d[1]; ++k) rotate(pool->vertices + 3 * ((size_t)d[0] + k), &w.v[3 * k]);
    w.n.resize(3 * d[5]);
    for (uint32_t k = 0; k < d[5]; ++k) rotate(pool->planes + 4 * ((size_t)d[4] + k), &w.n[3 * k]);
    rotate(pool->centroids + 3 * (size_t)h, w.c);
    return w;
}

inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

}  // namespace

// Distance t along unit u at which B (rotation Rb, origin pA + t u) has SAT signed
// separation g from A (rotation Ra, origin pA).  For a fixed axis the separation is
// linear in t (support points do not move with translation), and the Minkowski-face
// edge pairs depend only on orientation, so sep(t) = max_k (a_k + t b_k) over the face
// axes of both hulls and the Minkowski-face edge axes; the result is the right end of
// the sublevel set {t : sep(t) <= g}, where sep crosses g while increasing.
double place_along(const b3amd_hull_pool* pool, uint32_t ha, const float Ra[9], uint32_t hb,
                   const float Rb[9], const double u[3], double g) {
    WorldHull A = rotate_hull(pool, ha, Ra), B = rotate_hull(pool, hb, Rb);
    const size_t VA = A.v.size() / 3, VB = B.v.size() / 3, FA = A.n.size() / 3, FB = B.n.size() / 3;
    double tR = 1e300, tL = -1e300;
    auto axis = [&](double a, double b) {
        if (b > 1e-12) tR = std::min(tR, (g - a) / b);
        else if (b < -1e-12) tL = std::max(tL, (g - a) / b);
    };
    auto minmax = [](const std::vector<double>& v, size_t n, const double* ax, double* lo, double* hi) {
        *lo = 1e300;
        *hi = -1e300;
        for (size_t k = 0; k < n; ++k) {
            double p = dot3(&v[3 * k], ax);
            *lo = std::min(*lo, p);
            *hi = std::max(*hi, p);
        }
    };
    for (size_t f = 0; f < FA; ++f) {
        const double* n = &A.n[3 * f];
        double loB, hiB, loA, hiA;
        minmax(B.v, VB, n, &loB, &hiB);
        minmax(A.v, VA, n, &loA, &hiA);
        axis(loB - hiA, dot3(n, u));
    }
    for (size_t f = 0; f < FB; ++f) {
        const double* m = &B.n[3 * f];
        double loB, hiB, loA, hiA;
        minmax(B.v, VB, m, &loB, &hiB);
        minmax(A.v, VA, m, &loA, &hiA);
        axis(loA - hiB, -dot3(m, u));
    }
    const uint32_t* da = pool->desc + 6 * (size_t)ha;
    const uint32_t* db = pool->desc + 6 * (size_t)hb;
    const uint32_t* ea = pool->edges + 4 * (size_t)da[2];
    const uint32_t* eb = pool->edges + 4 * (size_t)db[2];
    for (uint32_t i = 0; i < da[3]; i += 2) {
        const double* P1 = &A.v[3 * ea[4 * i]];
        const double* Q1 = &A.v[3 * ea[4 * (i + 1)]];
        double E1[3] = {Q1[0] - P1[0], Q1[1] - P1[1], Q1[2] - P1[2]};
        const double* U1 = &A.n[3 * ea[4 * i + 2]];
        const double* V1 = &A.n[3 * ea[4 * (i + 1) + 2]];
        double D1[3] = {P1[0] - A.c[0], P1[1] - A.c[1], P1[2] - A.c[2]};
        for (uint32_t j = 0; j < db[3]; j += 2) {
            const double* P2 = &B.v[3 * eb[4 * j]];
            const double* Q2 = &B.v[3 * eb[4 * (j + 1)]];
            double E2[3] = {Q2[0] - P2[0], Q2[1] - P2[1], Q2[2] - P2[2]};
            const double* U2 = &B.n[3 * eb[4 * j + 2]];
            const double* V2 = &B.n[3 * eb[4 * (j + 1) + 2]];
            // Gauss-map arcs (U1,V1) and (-U2,-V2) intersect.
            double ADC = -dot3(U1, E2), BDC = -dot3(V1, E2);
            double CBA = dot3(U2, E1), DBA = dot3(V2, E1);
            if (!(CBA * DBA < 0.0 && ADC * BDC < 0.0 && CBA * BDC > 0.0)) continue;
            double N[3] = {E1[1] * E2[2] - E1[2] * E2[1], E1[2] * E2[0] - E1[0] * E2[2],
                           E1[0] * E2[1] - E1[1] * E2[0]};
            double L = std::sqrt(dot3(N, N));
            double L1 = std::sqrt(dot3(E1, E1)), L2 = std::sqrt(dot3(E2, E2));
            if (!(L > 0.005 * L1 * L2)) continue;
            N[0] /= L;
            N[1] /= L;
            N[2] /= L;
            if (dot3(N, D1) < 0.0) { N[0] = -N[0]; N[1] = -N[1]; N[2] = -N[2]; }
            double a = N[0] * (P2[0] - P1[0]) + N[1] * (P2[1] - P1[1]) + N[2] * (P2[2] - P1[2]);
            axis(a, dot3(N, u));
        }
    }
    (void)tL;  // infeasible only for g below the deepest reachable overlap; tR then overshoots g
    return tR;
}

}  // namespace b3amd

using namespace b3amd;

extern "C" int b3amd_gen_hull_pool(uint64_t seed, uint32_t count, uint32_t nmin, uint32_t nmax,
                                   float amin, float amax, uint32_t* desc, float* centroids,
                                   float* vertices, uint32_t* edges, uint32_t* faces,
                                   float* planes, uint64_t* sizes) {
    if (!desc || !centroids || !vertices || !edges || !faces || !planes || !sizes) return B3AMD_ERR_ARGUMENT;
    if (nmin < 4 || nmax < nmin || nmax > 255 || !(amin > 0.0f) || amax < amin) return B3AMD_ERR_ARGUMENT;
    std::vector<HullRecord> recs(count);
    std::vector<int> status(count, 0);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nth; ++t) {
        pool.emplace_back([&, t]() {
            for (uint32_t h = t; h < count; h += nth) {
                SplitMix rng = stream(seed, h);
                uint32_t n = nmin + (uint32_t)(rng.next() % (uint64_t)(nmax - nmin + 1));
                double a = amin + (amax - amin) * rng.u01();
                double b = amin + (amax - amin) * rng.u01();
                double c = amin + (amax - amin) * rng.u01();
                status[h] = ellipsoid_hull(rng, (int)n, a, b, c, &recs[h]);
            }
        });
    }
    for (auto& th : pool) th.join();
    uint64_t nv = 0, ne = 0, nf = 0;
    for (uint32_t h = 0; h < count; ++h) {
        if (status[h] < 4) return B3AMD_ERR_UNSUPPORTED;
        const HullRecord& r = recs[h];
        uint32_t V = (uint32_t)(r.vertices.size() / 3), HE = (uint32_t)(r.edges.size() / 4),
                 F = (uint32_t)r.faces.size();
        desc[6 * h + 0] = (uint32_t)nv;
        desc[6 * h + 1] = V;
        desc[6 * h + 2] = (uint32_t)ne;
        desc[6 * h + 3] = HE;
        desc[6 * h + 4] = (uint32_t)nf;
        desc[6 * h + 5] = F;
        std::memcpy(centroids + 3 * h, r.centroid, 12);
        std::memcpy(vertices + 3 * nv, r.vertices.data(), 12 * V);
        std::memcpy(edges + 4 * ne, r.edges.data(), 16 * HE);
        std::memcpy(faces + nf, r.faces.data(), 4 * F);
        std::memcpy(planes + 4 * nf, r.planes.data(), 16 * F);
        nv += V;
        ne += HE;
        nf += F;
    }
    sizes[0] = nv;
    sizes[1] = ne;
    sizes[2] = nf;
    return B3AMD_OK;
}

extern "C" int b3amd_gen_hull_pairs(uint64_t seed, const b3amd_hull_pool* pool, uint32_t n,
                                    float gap_mean, float gap_sigma, uint32_t* pairs,
                                    b3amd_transform* xfs) {
    if (!pool || !pairs || !xfs || pool->count == 0) return B3AMD_ERR_ARGUMENT;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> threads;
    for (unsigned t = 0; t < nth; ++t) {
        threads.emplace_back([=]() {
            for (uint32_t i = t; i < n; i += nth) {
                SplitMix rng = stream(seed ^ 0x5A17ull, i);
                uint32_t ha = (uint32_t)(rng.next() % pool->count);
                uint32_t hb = (uint32_t)(rng.next() % pool->count);
                double qa[4], qb[4], u[3];
                rng.unit_quat(qa);
                rng.unit_quat(qb);
                rng.unit_vector(u);
                float Ra[9], Rb[9];
                quat_to_rotation(qa, Ra);
                quat_to_rotation(qb, Rb);
                double g = gap_mean + gap_sigma * rng.normal();
                double t_ = place_along(pool, ha, Ra, hb, Rb, u, g);
                double pa[3] = {20.0 * rng.u01() - 10.0, 20.0 * rng.u01() - 10.0, 20.0 * rng.u01() - 10.0};
                pairs[2 * i + 0] = ha;
                pairs[2 * i + 1] = hb;
                b3amd_transform& A = xfs[2 * i];
                b3amd_transform& B = xfs[2 * i + 1];
                std::memcpy(&A.rotation[0], Ra, 36);
                std::memcpy(&B.rotation[0], Rb, 36);
                A.position = {(float)pa[0], (float)pa[1], (float)pa[2]};
                B.position = {(float)(pa[0] + t_ * u[0]), (float)(pa[1] + t_ * u[1]), (float)(pa[2] + t_ * u[2])};
            }
        });
    }
    for (auto& th : threads) th.join();
    return B3AMD_OK;
}

extern "C" int b3amd_gen_place_along(const b3amd_hull_pool* pool, uint32_t n, const uint32_t* ha, const float* ra,
                                     const uint32_t* hb, const float* rb, const double* dirs, const double* gaps,
                                     double* t) {
    if (!pool || (n && (!ha || !ra || !hb || !rb || !dirs || !gaps || !t))) return B3AMD_ERR_ARGUMENT;
    for (uint32_t i = 0; i < n; ++i)
        if (ha[i] >= pool->count || hb[i] >= pool->count) return B3AMD_ERR_ARGUMENT;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> threads;
    for (unsigned k = 0; k < nth; ++k) {
        threads.emplace_back([=]() {
            for (uint32_t i = k; i < n; i += nth)
                t[i] = place_along(pool, ha[i], ra + 9 * (size_t)i, hb[i], rb + 9 * (size_t)i, dirs + 3 * (size_t)i,
                                   gaps[i]);
        });
    }
    for (auto& th : threads) th.join();
    return B3AMD_OK;
}
