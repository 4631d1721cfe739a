// Convex hull construction for the workload generators and for users that need
// b3Hull-layout records (hull.h:37-64) without the reference's quickhull.
//
// Incremental construction in double precision: start from a maximal tetrahedron,
// add points in input order, replace the visible cap by a cone over the horizon.
// Output is triangulated (V = hull vertices, F = 2V - 4, HE = 6V - 12) and laid out
// the way the SAT edge loop requires: half-edges 2k and 2k+1 are twins
// (sat.cpp:123-128), every face lists its half-edges through `next`, planes are
// unit outward normals with offset = dot(n, vertex), and the centroid is the
// volume-weighted centroid.
#include "hullgen.h"

#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace b3amd {
namespace {

struct D3 {
    double x, y, z;
};
inline D3 sub(const D3& a, const D3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 cross(const D3& a, const D3& b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double dot(const D3& a, const D3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

struct Tri {
    int v[3];
    D3 n;      // unnormalised outward normal
    double d;  // dot(n, v0)
    bool alive;
};

Tri make_tri(const std::vector<D3>& p, int a, int b, int c) {
    Tri t;
    t.v[0] = a;
    t.v[1] = b;
    t.v[2] = c;
    t.n = cross(sub(p[b], p[a]), sub(p[c], p[a]));
    t.d = dot(t.n, p[a]);
    t.alive = true;
    return t;
}

double signed_dist(const Tri& t, const D3& q) { return dot(t.n, q) - t.d; }

}  // namespace

int build_hull(const float* points, int n, HullRecord* out) {
    out->vertices.clear();
    out->edges.clear();
    out->faces.clear();
    out->planes.clear();
    if (n < 4) return -1;
    std::vector<D3> p(n);
    for (int i = 0; i < n; ++i) p[i] = {points[3 * i], points[3 * i + 1], points[3 * i + 2]};

    // Scale for tolerances.
    double extent = 0.0;
    for (int i = 0; i < n; ++i) {
        extent = std::fmax(extent, std::fabs(p[i].x));
        extent = std::fmax(extent, std::fabs(p[i].y));
        extent = std::fmax(extent, std::fabs(p[i].z));
    }
    if (extent <= 0.0) return -2;
    const double eps = 1e-9 * extent;

    // Initial tetrahedron.
    int i0 = 0, i1 = -1, i2 = -1, i3 = -1;
    double best = 0.0;
    for (int i = 1; i < n; ++i) {
        D3 d = sub(p[i], p[i0]);
        double l = dot(d, d);
        if (l > best) { best = l; i1 = i; }
    }
    if (i1 < 0 || best <= eps * eps) return -3;
    best = 0.0;
    for (int i = 0; i < n; ++i) {
        D3 c = cross(sub(p[i1], p[i0]), sub(p[i], p[i0]));
        double l = dot(c, c);
        if (l > best) { best = l; i2 = i; }
    }
    if (i2 < 0 || best <= eps * eps * extent * extent) return -4;
    Tri base = make_tri(p, i0, i1, i2);
    best = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = std::fabs(signed_dist(base, p[i]));
        if (s > best) { best = s; i3 = i; }
    }
    if (i3 < 0 || best <= eps * std::sqrt(dot(base.n, base.n))) return -5;
    if (signed_dist(base, p[i3]) > 0.0) std::swap(i1, i2);

    std::vector<Tri> tris;
    tris.push_back(make_tri(p, i0, i1, i2));
    tris.push_back(make_tri(p, i0, i3, i1));
    tris.push_back(make_tri(p, i1, i3, i2));
    tris.push_back(make_tri(p, i2, i3, i0));

    std::vector<char> used(n, 0);
    used[i0] = used[i1] = used[i2] = used[i3] = 1;

    // Directed edge (a,b) -> owning live face, as a dense n*n table.
    std::vector<int> edgeFace((size_t)n * n, -1);
    std::vector<double> nlen;
    auto add_tri = [&](int a, int b, int c) {
        tris.push_back(make_tri(p, a, b, c));
        const Tri& t = tris.back();
        nlen.push_back(std::sqrt(dot(t.n, t.n)));
        int id = (int)tris.size() - 1;
        for (int k = 0; k < 3; ++k) edgeFace[(size_t)t.v[k] * n + t.v[(k + 1) % 3]] = id;
    };
    {
        std::vector<Tri> init;
        init.swap(tris);
        for (const Tri& t : init) add_tri(t.v[0], t.v[1], t.v[2]);
    }

    std::vector<int> visible;
    std::vector<std::pair<int, int>> horizon;
    for (int q = 0; q < n; ++q) {
        if (used[q]) continue;
        visible.clear();
        for (int t = 0; t < (int)tris.size(); ++t) {
            if (!tris[t].alive) continue;
            if (signed_dist(tris[t], p[q]) > eps * nlen[t]) {
                visible.push_back(t);
                tris[t].alive = false;  // removed: the cone over the horizon replaces it
            }
        }
        if (visible.empty()) continue;  // inside or on the hull: dropped
        horizon.clear();
        for (int t : visible) {
            for (int k = 0; k < 3; ++k) {
                int a = tris[t].v[k], b = tris[t].v[(k + 1) % 3];
                int other = edgeFace[(size_t)b * n + a];
                if (other < 0) return -6;
                if (tris[other].alive) horizon.push_back({a, b});
            }
        }
        for (int t : visible)
            for (int k = 0; k < 3; ++k) {
                size_t key = (size_t)tris[t].v[k] * n + tris[t].v[(k + 1) % 3];
                if (edgeFace[key] == t) edgeFace[key] = -1;
            }
        for (auto& e : horizon) add_tri(e.first, e.second, q);
        used[q] = 1;
    }

    // Compact vertices in input order.
    std::vector<int> remap(n, -1);
    for (const Tri& t : tris)
        if (t.alive)
            for (int k = 0; k < 3; ++k) remap[t.v[k]] = 0;
    int V = 0;
    for (int i = 0; i < n; ++i)
        if (remap[i] == 0) remap[i] = V++;
    out->vertices.resize(3 * V);
    for (int i = 0; i < n; ++i)
        if (remap[i] >= 0) {
            out->vertices[3 * remap[i] + 0] = points[3 * i + 0];
            out->vertices[3 * remap[i] + 1] = points[3 * i + 1];
            out->vertices[3 * remap[i] + 2] = points[3 * i + 2];
        }

    // Half-edge layout with adjacent twins.
    std::vector<int> undirected((size_t)V * V, -1);  // (min,max) -> unique edge id
    std::vector<int> faceTris;
    for (int t = 0; t < (int)tris.size(); ++t)
        if (tris[t].alive) faceTris.push_back(t);
    const int F = (int)faceTris.size();
    const int E = 3 * F / 2;
    std::vector<uint32_t> he(4 * 2 * E, 0xFFFFFFFFu);
    out->faces.resize(F);
    int nextEdge = 0;
    for (int f = 0; f < F; ++f) {
        const Tri& t = tris[faceTris[f]];
        int hid[3];
        for (int k = 0; k < 3; ++k) {
            int a = remap[t.v[k]], b = remap[t.v[(k + 1) % 3]];
            size_t key = (size_t)std::min(a, b) * V + std::max(a, b);
            int h;
            if (undirected[key] < 0) {
                int id = nextEdge++;
                if (id >= E) return -7;
                undirected[key] = id;
                h = 2 * id;
            } else {
                h = 2 * undirected[key] + 1;
            }
            he[4 * h + 0] = (uint32_t)a;
            he[4 * h + 1] = (uint32_t)(h ^ 1);
            he[4 * h + 2] = (uint32_t)f;
            hid[k] = h;
        }
        for (int k = 0; k < 3; ++k) he[4 * hid[k] + 3] = (uint32_t)hid[(k + 1) % 3];
        out->faces[f] = (uint32_t)hid[0];
    }
    if (nextEdge != E) return -8;
    for (size_t i = 0; i < he.size(); ++i)
        if (he[i] == 0xFFFFFFFFu) return -9;
    out->edges = he;

    // Planes from the float vertices, in double, rounded once.
    std::vector<D3> fv(V);
    for (int i = 0; i < V; ++i)
        fv[i] = {out->vertices[3 * i], out->vertices[3 * i + 1], out->vertices[3 * i + 2]};
    out->planes.resize(4 * F);
    for (int f = 0; f < F; ++f) {
        uint32_t h0 = out->faces[f];
        uint32_t h1 = he[4 * h0 + 3];
        uint32_t h2 = he[4 * h1 + 3];
        const D3& a = fv[he[4 * h0]];
        const D3& b = fv[he[4 * h1]];
        const D3& c = fv[he[4 * h2]];
        D3 nn = cross(sub(b, a), sub(c, a));
        double l = std::sqrt(dot(nn, nn));
        if (!(l > 0.0)) return -10;
        D3 u = {nn.x / l, nn.y / l, nn.z / l};
        double off = (dot(u, a) + dot(u, b) + dot(u, c)) / 3.0;
        out->planes[4 * f + 0] = (float)u.x;
        out->planes[4 * f + 1] = (float)u.y;
        out->planes[4 * f + 2] = (float)u.z;
        out->planes[4 * f + 3] = (float)off;
    }

    // Volume-weighted centroid (tetrahedra fanned from the vertex mean).
    D3 ref = {0, 0, 0};
    for (int i = 0; i < V; ++i) { ref.x += fv[i].x; ref.y += fv[i].y; ref.z += fv[i].z; }
    ref = {ref.x / V, ref.y / V, ref.z / V};
    double vol = 0.0;
    D3 acc = {0, 0, 0};
    for (int f = 0; f < F; ++f) {
        uint32_t h0 = out->faces[f];
        uint32_t h1 = he[4 * h0 + 3];
        uint32_t h2 = he[4 * h1 + 3];
        D3 a = sub(fv[he[4 * h0]], ref), b = sub(fv[he[4 * h1]], ref), c = sub(fv[he[4 * h2]], ref);
        double v6 = dot(a, cross(b, c));
        vol += v6;
        acc.x += v6 * (a.x + b.x + c.x) / 4.0;
        acc.y += v6 * (a.y + b.y + c.y) / 4.0;
        acc.z += v6 * (a.z + b.z + c.z) / 4.0;
    }
    if (!(vol > 0.0)) return -11;
    out->centroid[0] = (float)(ref.x + acc.x / vol);
    out->centroid[1] = (float)(ref.y + acc.y / vol);
    out->centroid[2] = (float)(ref.z + acc.z / vol);
    return V;
}

}  // namespace b3amd
