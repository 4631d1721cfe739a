// TEST INFRASTRUCTURE ONLY.  extern "C" driver over the reference library, compiled
// together with the reference sources in place (oracle/Makefile) into
// oracle/_ref/libb3ref.so.  It builds b3Hull views over the flat pool, wraps them in
// the reference's shape classes and calls b3CollideShapeAndShape (collide.cpp:162-188)
// exactly as b3ConvexContact::Collide does (convex_contact.cpp:50-63).
//
// The body pointer of a shape is only read by the cached face rebuild
// (collide_hulls_cache.cpp:117-126, GetBody()->GetSweep()); the driver points it at a
// zeroed b3Body image whose m_sweep carries the pair's orientation0/orientation.
#include <cstring>
#include <cstdlib>
#include <unordered_map>
#include <vector>
#include <new>

#define private public
#define protected public
#include <bounce/collision/shapes/qhull.h>
#include <bounce/dynamics/body.h>
#include <bounce/dynamics/shapes/sphere_shape.h>
#include <bounce/dynamics/shapes/capsule_shape.h>
#include <bounce/dynamics/shapes/hull_shape.h>
#include <bounce/dynamics/contacts/manifold.h>
#include <bounce/dynamics/contacts/collide/collide.h>
#include <bounce/collision/shapes/hull.h>
#include <bounce/collision/shapes/mesh.h>
#include <bounce/dynamics/shapes/mesh_shape.h>
#include <bounce/dynamics/contacts/mesh_contact.h>
#include <bounce/dynamics/contacts/convex_contact.h>
#include <bounce/collision/shapes/box_hull.h>
#include <bounce/dynamics/world.h>
#include <bounce/common/draw.h>
#include <bounce/common/memory/stack_allocator.h>
#include <bounce/dynamics/contacts/contact_solver.h>
#undef private
#undef protected

#include "b3oracle.h"

static_assert(sizeof(b3Transform) == sizeof(b3o_xf), "transform layout");
static_assert(sizeof(b3HalfEdge) == sizeof(b3o_halfedge), "half-edge layout");
static_assert(sizeof(b3Plane) == 16, "plane layout");
static_assert(sizeof(b3ConvexCache) == sizeof(b3o_convex_cache), "cache layout");
static_assert(sizeof(b3Quat) == sizeof(b3o_quat), "quat layout");

// Profiling hooks the library expects its user to define (settings.h:147-151).
void b3BeginProfileScope(const char*) {}
void b3EndProfileScope() {}
// Debug-draw globals normally defined by draw_world.cpp (not linked: it pulls in the cloth module).
b3Draw* b3Draw_draw = nullptr;
const b3Color b3Color_red(1.0f, 0.0f, 0.0f);
const b3Color b3Color_green(0.0f, 1.0f, 0.0f);
const b3Color b3Color_yellow(1.0f, 1.0f, 0.0f);
const b3Color b3Color_pink(1.0f, 0.0f, 1.0f);

namespace {

struct Scene {
    std::vector<b3Hull> hulls;
    std::vector<b3Shape*> shapes;
    ~Scene() {
        for (b3Shape* s : shapes) {
            switch (s->m_type) {
            case e_sphereShape: delete (b3SphereShape*)s; break;
            case e_capsuleShape: delete (b3CapsuleShape*)s; break;
            default: delete (b3HullShape*)s; break;
            }
        }
    }
};

void build_hulls(const b3o_hull_pool* pool, std::vector<b3Hull>& out) {
    out.resize(pool->count);
    for (uint32_t h = 0; h < pool->count; ++h) {
        const uint32_t* d = pool->desc + 6 * h;
        b3Hull& hull = out[h];
        hull.centroid = b3Vec3(pool->centroids[3 * h], pool->centroids[3 * h + 1], pool->centroids[3 * h + 2]);
        hull.vertexCount = d[1];
        hull.vertices = (b3Vec3*)(pool->vertices + 3 * (size_t)d[0]);
        hull.edgeCount = d[3];
        hull.edges = (b3HalfEdge*)(pool->edges + 4 * (size_t)d[2]);
        hull.faceCount = d[5];
        hull.faces = (b3Face*)(pool->faces + d[4]);
        hull.planes = (b3Plane*)(pool->planes + 4 * (size_t)d[4]);
    }
}

b3Shape* make_shape(const b3o_shape& s, std::vector<b3Hull>& hulls) {
    switch (s.type) {
    case B3O_SPHERE: {
        b3SphereShape* sp = new b3SphereShape();
        sp->m_center = b3Vec3(s.v[0], s.v[1], s.v[2]);
        sp->m_radius = s.radius;
        return sp;
    }
    case B3O_CAPSULE: {
        b3CapsuleShape* cp = new b3CapsuleShape();
        cp->m_centers[0] = b3Vec3(s.v[0], s.v[1], s.v[2]);
        cp->m_centers[1] = b3Vec3(s.v[3], s.v[4], s.v[5]);
        cp->m_radius = s.radius;
        return cp;
    }
    default: {
        b3HullShape* hp = new b3HullShape();
        hp->m_hull = &hulls[s.hull];
        hp->m_radius = s.radius;
        return hp;
    }
    }
}

struct BodyImage {
    alignas(b3Body) unsigned char bytes[sizeof(b3Body)];
    b3Body* body() { return reinterpret_cast<b3Body*>(bytes); }
};

void to_out(const b3Manifold& m, b3o_manifold* o) {
    std::memset(o, 0, sizeof(*o));
    o->pointCount = m.pointCount;
    for (u32 i = 0; i < m.pointCount; ++i) {
        const b3ManifoldPoint& p = m.points[i];
        b3o_mpoint& q = o->points[i];
        std::memcpy(&q.localNormal1, &p.localNormal1, 12);
        std::memcpy(&q.localPoint1, &p.localPoint1, 12);
        std::memcpy(&q.localPoint2, &p.localPoint2, 12);
        q.triangleKey = p.key.triangleKey;
        std::memcpy(q.key, &p.key.key1, 8);
        std::memcpy(q.key + 2, &p.key.key2, 8);
    }
}

}  // namespace

// The reference's statistics globals (collide_hulls.cpp:291-292, gjk.cpp:27).
extern u32 b3_convexCalls, b3_convexCacheHits;
extern u32 b3_gjkCalls, b3_gjkIters, b3_gjkMaxIters, b3_gjkCacheHits;

extern "C" void b3ref_get_counters(uint64_t out[6]) {
    out[0] = b3_convexCalls;
    out[1] = b3_convexCacheHits;
    out[2] = b3_gjkCalls;
    out[3] = b3_gjkIters;
    out[4] = b3_gjkMaxIters;
    out[5] = b3_gjkCacheHits;
}

extern "C" void b3ref_reset_counters(void) {
    b3_convexCalls = b3_convexCacheHits = 0;
    b3_gjkCalls = b3_gjkIters = b3_gjkMaxIters = b3_gjkCacheHits = 0;
}

// The step after the narrowphase for n contacts of M manifold slots: b3Manifold::Initialize()
// and, with warm starting, Initialize(old) of each old manifold in turn, in
// b3Contact::Update's loop (contact.cpp:63-91); then a b3ContactSolver over those contacts runs
// InitializeConstraints (contact_solver.cpp:70-276) and its velocity constraints are copied out.
// Each contact is a b3ConvexContact between two hull shapes whose bodies carry the island
// index, mass data and local centre the solver reads; its m_manifolds points at its M slots.
extern "C" void b3ref_contact_init(uint32_t n, uint32_t M, const uint32_t* bodies, const b3o_solver_body* state,
                                   const b3o_contact_material* material, const uint32_t* counts,
                                   const b3o_manifold* manifolds, const uint32_t* oldCounts,
                                   const b3o_manifold* oldManifolds, const b3o_manifold_impulses* oldImpulses,
                                   int warmStarting, b3o_manifold_impulses* impulses, b3o_velocity_manifold* out) {
    uint32_t nb = 0;
    for (uint32_t i = 0; i < 2 * n; ++i) nb = bodies[i] + 1 > nb ? bodies[i] + 1 : nb;
    std::vector<BodyImage> images(nb);
    std::vector<b3Position> positions(nb);
    std::vector<b3Velocity> velocities(nb);
    std::vector<b3Mat33> invInertias(nb);
    for (uint32_t b = 0; b < nb; ++b) {
        std::memset(images[b].bytes, 0, sizeof(images[b].bytes));
        b3Body* body = images[b].body();
        const b3o_solver_body& s = state[b];
        body->m_islandID = b;
        body->m_invMass = s.invMass;
        std::memcpy(&body->m_sweep.localCenter, &s.localCenter, 12);
        std::memcpy(&positions[b].x, &s.x, 12);
        std::memcpy(&positions[b].q, &s.q, 16);
        std::memcpy(&velocities[b].v, &s.v, 12);
        std::memcpy(&velocities[b].w, &s.w, 12);
        std::memcpy(&invInertias[b], s.invI, 36);
    }
    std::vector<b3HullShape> shapes(2 * (size_t)n);
    std::vector<b3Manifold> slots((size_t)n * M);
    std::vector<b3ConvexContact*> contacts(n);
    for (uint32_t i = 0; i < n; ++i) {
        const b3o_contact_material& mt = material[i];
        b3HullShape& A = shapes[2 * i];
        b3HullShape& B = shapes[2 * i + 1];
        A.m_radius = mt.radiusA, A.m_friction = mt.frictionA, A.m_restitution = mt.restitutionA;
        B.m_radius = mt.radiusB, B.m_friction = mt.frictionB, B.m_restitution = mt.restitutionB;
        A.m_body = images[bodies[2 * i]].body();
        B.m_body = images[bodies[2 * i + 1]].body();
        b3ConvexContact* c = new b3ConvexContact(&A, &B);
        c->m_pair.shapeA = &A;
        c->m_pair.shapeB = &B;
        c->m_manifolds = slots.data() + (size_t)i * M;
        c->m_manifoldCapacity = M;
        // The old manifolds with their impulses, then the new ones over them (contact.cpp:65-78).
        b3Manifold old[3];
        for (uint32_t k = 0; k < oldCounts[i]; ++k) {
            b3Manifold& m = old[k];
            const b3o_manifold& src = oldManifolds[(size_t)i * M + k];
            const b3o_manifold_impulses& imp = oldImpulses[(size_t)i * M + k];
            m.Initialize();
            m.pointCount = src.pointCount;
            for (uint32_t q = 0; q < src.pointCount; ++q) {
                std::memcpy(&m.points[q].key.triangleKey, &src.points[q].triangleKey, 4);
                std::memcpy(&m.points[q].key.key1, src.points[q].key, 8);
                std::memcpy(&m.points[q].key.key2, src.points[q].key + 2, 8);
                m.points[q].normalImpulse = imp.normalImpulse[q];
            }
            m.tangentImpulse.Set(imp.tangentImpulse[0], imp.tangentImpulse[1]);
            m.motorImpulse = imp.motorImpulse;
        }
        for (uint32_t k = 0; k < M; ++k) c->m_manifolds[k].Initialize();
        c->m_manifoldCount = counts[i];
        for (uint32_t j = 0; j < counts[i]; ++j) {
            b3Manifold& m = c->m_manifolds[j];
            const b3o_manifold& src = manifolds[(size_t)i * M + j];
            m.pointCount = src.pointCount;
            for (uint32_t q = 0; q < src.pointCount; ++q) {
                std::memcpy(&m.points[q].localNormal1, &src.points[q].localNormal1, 12);
                std::memcpy(&m.points[q].localPoint1, &src.points[q].localPoint1, 12);
                std::memcpy(&m.points[q].localPoint2, &src.points[q].localPoint2, 12);
                std::memcpy(&m.points[q].key.triangleKey, &src.points[q].triangleKey, 4);
                std::memcpy(&m.points[q].key.key1, src.points[q].key, 8);
                std::memcpy(&m.points[q].key.key2, src.points[q].key + 2, 8);
            }
        }
        if (warmStarting)
            for (uint32_t j = 0; j < counts[i]; ++j)
                for (uint32_t k = 0; k < oldCounts[i]; ++k) c->m_manifolds[j].Initialize(old[k]);
        for (uint32_t j = 0; j < counts[i]; ++j) {
            const b3Manifold& m = c->m_manifolds[j];
            b3o_manifold_impulses& imp = impulses[(size_t)i * M + j];
            std::memset(&imp, 0, sizeof(imp));
            for (uint32_t q = 0; q < m.pointCount; ++q) {
                imp.normalImpulse[q] = m.points[q].normalImpulse;
                imp.persisting[q] = m.points[q].persisting;
            }
            imp.tangentImpulse[0] = m.tangentImpulse.x;
            imp.tangentImpulse[1] = m.tangentImpulse.y;
            imp.motorImpulse = m.motorImpulse;
        }
        contacts[i] = c;
    }
    b3StackAllocator* allocator = new b3StackAllocator();
    {
        b3ContactSolverDef def;
        def.positions = positions.data();
        def.velocities = velocities.data();
        def.invInertias = invInertias.data();
        def.contacts = reinterpret_cast<b3Contact**>(contacts.data());
        def.count = n;
        def.allocator = allocator;
        def.dt = 1.0f / 60.0f;
        b3ContactSolver solver(&def);
        solver.InitializeConstraints();
        for (uint32_t i = 0; i < n; ++i) {
            const b3ContactVelocityConstraint& vc = solver.m_velocityConstraints[i];
            for (uint32_t j = 0; j < counts[i]; ++j) {
                const b3VelocityConstraintManifold& vcm = vc.manifolds[j];
                b3o_velocity_manifold& o = out[(size_t)i * M + j];
                std::memset(&o, 0, sizeof(o));
                o.pointCount = vcm.pointCount;
                for (uint32_t q = 0; q < vcm.pointCount; ++q) {
                    const b3VelocityConstraintPoint& p = vcm.points[q];
                    std::memcpy(&o.points[q].rA, &p.rA, 12);
                    std::memcpy(&o.points[q].rB, &p.rB, 12);
                    std::memcpy(&o.points[q].normal, &p.normal, 12);
                    o.points[q].normalMass = p.normalMass;
                    o.points[q].normalImpulse = p.normalImpulse;
                    o.points[q].velocityBias = p.velocityBias;
                }
                if (vcm.pointCount > 0) {  // the reference leaves the rest unset for empty manifolds
                    std::memcpy(&o.rA, &vcm.rA, 12);
                    std::memcpy(&o.rB, &vcm.rB, 12);
                    std::memcpy(&o.normal, &vcm.normal, 12);
                    std::memcpy(&o.tangent1, &vcm.tangent1, 12);
                    std::memcpy(&o.tangent2, &vcm.tangent2, 12);
                    o.tangentMass[0] = vcm.tangentMass.x.x, o.tangentMass[1] = vcm.tangentMass.x.y;
                    o.tangentMass[2] = vcm.tangentMass.y.x, o.tangentMass[3] = vcm.tangentMass.y.y;
                    o.motorMass = vcm.motorMass;
                }
                o.tangentImpulse[0] = vcm.tangentImpulse.x;
                o.tangentImpulse[1] = vcm.tangentImpulse.y;
                o.motorImpulse = vcm.motorImpulse;
                o.friction = vc.friction;
                o.restitution = vc.restitution;
            }
        }
    }
    delete allocator;
    for (b3ConvexContact* c : contacts) {
        c->m_manifolds = &c->m_stackManifold;
        delete c;
    }
}

// b3QHull::Set (qhull.cpp:104-354) on `count` points; writes the hull's records and counts
// (vertices, half-edges, faces), all zero when the reference leaves the hull empty.
extern "C" void b3ref_qhull(const float* points, uint32_t count, int simplify, uint32_t counts[3], float* vertices,
                            uint32_t* edges, uint32_t* faces, float* planes, float* centroid) {
    b3QHull hull;
    hull.Set(sizeof(b3Vec3), points, count, simplify != 0);
    counts[0] = hull.vertexCount;
    counts[1] = hull.edgeCount;
    counts[2] = hull.faceCount;
    if (hull.vertexCount == 0) {
        counts[1] = counts[2] = 0;
        return;
    }
    std::memcpy(vertices, hull.vertices, 12 * (size_t)hull.vertexCount);
    std::memcpy(edges, hull.edges, 16 * (size_t)hull.edgeCount);
    std::memcpy(faces, hull.faces, 4 * (size_t)hull.faceCount);
    std::memcpy(planes, hull.planes, 16 * (size_t)hull.faceCount);
    std::memcpy(centroid, &hull.centroid, 12);
}

// b3TestOverlap (collide.cpp:84-95) per pair with the pair's cache, as
// b3ConvexContact::TestOverlap calls it for a sensor contact (convex_contact.cpp:39-48).
extern "C" void b3ref_test_overlap(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t n,
                                   const uint32_t* pairs, const b3o_xf* xfs, b3o_convex_cache* caches,
                                   uint32_t* overlap) {
    Scene scene;
    build_hulls(pool, scene.hulls);
    for (uint32_t i = 0; i < n; ++i) {
        b3Shape* A = make_shape(shapes[pairs[2 * i]], scene.hulls);
        b3Shape* B = make_shape(shapes[pairs[2 * i + 1]], scene.hulls);
        scene.shapes.push_back(A);
        scene.shapes.push_back(B);
        b3ConvexCache cache;
        if (caches) {
            std::memcpy(&cache, caches + i, sizeof(cache));
        } else {
            std::memset(&cache, 0, sizeof(cache));
            cache.featureCache.m_featurePair.state = e_empty;
        }
        b3Transform xfA, xfB;
        std::memcpy(&xfA, xfs + 2 * (size_t)i, sizeof(xfA));
        std::memcpy(&xfB, xfs + 2 * (size_t)i + 1, sizeof(xfB));
        overlap[i] = b3TestOverlap(xfA, 0, A, xfB, 0, B, &cache) ? 1u : 0u;
        if (caches) std::memcpy(caches + i, &cache, sizeof(cache));
        if (scene.shapes.size() > 4096) {
            Scene tmp;
            tmp.shapes.swap(scene.shapes);
        }
    }
}

extern "C" void b3ref_collide_pairs(const b3o_hull_pool* pool, const b3o_shape* shapes,
                                    uint32_t n, const uint32_t* pairs, const b3o_xf* xfs,
                                    const b3o_quat* sweeps, b3o_convex_cache* caches,
                                    b3o_manifold* out) {
    Scene scene;
    build_hulls(pool, scene.hulls);
    // Shapes are created lazily so only referenced records are touched.
    std::vector<b3Shape*> table;
    BodyImage bodyA, bodyB;
    std::memset(bodyA.bytes, 0, sizeof(bodyA.bytes));
    std::memset(bodyB.bytes, 0, sizeof(bodyB.bytes));
    // No sweeps: the bodies did not rotate during the step (orientation0 == orientation),
    // here the identity, which is what a null sweep array means at the boundary.
    const b3Quat identity(0.0f, 0.0f, 0.0f, 1.0f);
    for (BodyImage* b : {&bodyA, &bodyB}) {
        b->body()->m_sweep.orientation0 = identity;
        b->body()->m_sweep.orientation = identity;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const b3o_shape& sa = shapes[pairs[2 * i]];
        const b3o_shape& sb = shapes[pairs[2 * i + 1]];
        b3Shape* A = make_shape(sa, scene.hulls);
        b3Shape* B = make_shape(sb, scene.hulls);
        scene.shapes.push_back(A);
        scene.shapes.push_back(B);
        A->m_body = bodyA.body();
        B->m_body = bodyB.body();
        if (sweeps) {
            const b3o_quat* q = sweeps + 4 * (size_t)i;
            std::memcpy(&bodyA.body()->m_sweep.orientation0, q + 0, 16);
            std::memcpy(&bodyA.body()->m_sweep.orientation, q + 1, 16);
            std::memcpy(&bodyB.body()->m_sweep.orientation0, q + 2, 16);
            std::memcpy(&bodyB.body()->m_sweep.orientation, q + 3, 16);
        }
        b3ConvexCache cache;
        if (caches) {
            std::memcpy(&cache, caches + i, sizeof(cache));
        } else {
            std::memset(&cache, 0, sizeof(cache));
            cache.simplexCache.count = 0;
            cache.featureCache.m_featurePair.state = e_empty;
        }
        b3Transform xfA, xfB;
        std::memcpy(&xfA, xfs + 2 * (size_t)i, sizeof(xfA));
        std::memcpy(&xfB, xfs + 2 * (size_t)i + 1, sizeof(xfB));
        b3Manifold m;
        m.Initialize();
        b3CollideShapeAndShape(m, xfA, A, xfB, B, &cache);
        to_out(m, out + i);
        if (caches) std::memcpy(caches + i, &cache, sizeof(cache));
        // Keep memory bounded on large batches.
        if (scene.shapes.size() > 4096) {
            Scene tmp;
            tmp.shapes.swap(scene.shapes);
        }
    }
}

extern "C" void b3ref_hull_queries(const b3o_hull_pool* pool, uint32_t n, const uint32_t* hullPairs,
                                   const b3o_xf* xfs, uint32_t* out) {
    std::vector<b3Hull> hulls;
    build_hulls(pool, hulls);
    for (uint32_t i = 0; i < n; ++i) {
        const b3Hull* h1 = &hulls[hullPairs[2 * i]];
        const b3Hull* h2 = &hulls[hullPairs[2 * i + 1]];
        b3Transform xf1, xf2;
        std::memcpy(&xf1, xfs + 2 * (size_t)i, sizeof(xf1));
        std::memcpy(&xf2, xfs + 2 * (size_t)i + 1, sizeof(xf2));
        b3FaceQuery f1 = b3QueryFaceSeparation(xf1, h1, xf2, h2);
        b3FaceQuery f2 = b3QueryFaceSeparation(xf2, h2, xf1, h1);
        b3EdgeQuery e = b3QueryEdgeSeparation(xf1, h1, xf2, h2);
        uint32_t* o = out + 8 * (size_t)i;
        o[0] = f1.index; std::memcpy(o + 1, &f1.separation, 4);
        o[2] = f2.index; std::memcpy(o + 3, &f2.separation, 4);
        o[4] = e.index1; o[5] = e.index2; std::memcpy(o + 6, &e.separation, 4);
        o[7] = 0;
    }
}

namespace {

void make_mesh(b3Mesh& mesh, const float* vertices, uint32_t vertexCount, uint32_t triangleCount,
               const uint32_t* triangles) {
    mesh.vertexCount = vertexCount;
    mesh.vertices = (b3Vec3*)vertices;
    mesh.triangleCount = triangleCount;
    mesh.triangles = (b3Triangle*)triangles;
    mesh.BuildTree();
}

uint64_t content_hash(const void* p, size_t bytes, uint64_t h) {  // FNV-1a over 8-byte words
    const unsigned char* b = (const unsigned char*)p;
    size_t i = 0;
    for (; i + 8 <= bytes; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = (h ^ w) * 1099511628211ull;
    }
    for (; i < bytes; ++i) h = (h ^ b[i]) * 1099511628211ull;
    return h;
}

void copy_nodes(const b3Mesh& mesh, float* nodesOut) {
    static_assert(sizeof(b3StaticTree::b3Node) == 32, "static tree node layout");
    std::memcpy(nodesOut, mesh.tree.m_nodes, 32ull * mesh.tree.m_nodeCount);
}

}  // namespace

extern "C" void b3ref_mesh_build_tree(const float* vertices, uint32_t vertexCount, uint32_t triangleCount,
                                      const uint32_t* triangles, float* nodesOut) {
    b3Mesh mesh;
    make_mesh(mesh, vertices, vertexCount, triangleCount, triangles);
    copy_nodes(mesh, nodesOut);
}

// Freshly created b3MeshContacts (contact_manager.cpp:246-268 + :130-141): construct,
// FindNewPairs, Collide, and read the triangle cache list and clustered manifolds.
extern "C" void b3ref_collide_mesh(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* m,
                                   uint32_t n, const uint32_t* shapeA, const b3o_xf* xfs, uint32_t cap,
                                   uint32_t* triCounts, uint32_t* triLists, b3o_manifold* out,
                                   uint32_t* manifoldCounts) {
    Scene scene;
    build_hulls(pool, scene.hulls);
    // The tree of the last mesh (same counts and content hash) is kept: repeated calls on
    // the same mesh (timing samples) do not rebuild it.
    static b3Mesh mesh;
    static uint64_t meshKey[3] = {0, 0, 0};
    const uint64_t key[3] = {m->vertexCount, m->triangleCount,
                             content_hash(m->vertices, 12ull * m->vertexCount,
                                          content_hash(m->triangles, 12ull * m->triangleCount, 1469598103934665603ull))};
    if (std::memcmp(key, meshKey, sizeof(key)) != 0) {
        mesh.~b3Mesh();
        new (&mesh) b3Mesh();
        make_mesh(mesh, m->vertices, m->vertexCount, m->triangleCount, m->triangles);
        std::memcpy(meshKey, key, sizeof(key));
    }
    // b3World's constructor and destructor reset the statistics globals
    // (world.cpp:39-45, 67-71): keep the totals across the world's lifetime.
    uint64_t before[6];
    b3ref_get_counters(before);
    b3World* world = new b3World();
    b3BodyDef bdA;
    bdA.type = e_dynamicBody;
    b3Body* bodyA = world->CreateBody(bdA);
    b3BodyDef bdB;
    b3Body* bodyB = world->CreateBody(bdB);
    b3MeshShape meshShape;
    meshShape.m_mesh = &mesh;
    meshShape.m_body = bodyB;
    for (uint32_t i = 0; i < n; ++i) {
        b3Shape* A = make_shape(shapes[shapeA[i]], scene.hulls);
        scene.shapes.push_back(A);
        A->m_body = bodyA;
        std::memcpy(&bodyA->m_xf, xfs + 2 * (size_t)i, sizeof(b3Transform));
        std::memcpy(&bodyB->m_xf, xfs + 2 * (size_t)i + 1, sizeof(b3Transform));
        b3MeshContact* mc = new b3MeshContact(A, &meshShape);
        mc->m_pair.shapeA = A;
        mc->m_pair.shapeB = &meshShape;
        mc->FindNewPairs();
        triCounts[i] = mc->m_triangleCount;
        for (uint32_t k = 0; k < mc->m_triangleCount && k < cap; ++k) triLists[(size_t)i * cap + k] = mc->m_triangles[k].index;
        mc->Collide();
        manifoldCounts[i] = mc->m_manifoldCount;
        for (uint32_t k = 0; k < 3; ++k) {
            if (k < mc->m_manifoldCount) to_out(mc->m_manifolds[k], out + 3 * (size_t)i + k);
            else std::memset(out + 3 * (size_t)i + k, 0, sizeof(b3o_manifold));
        }
        delete mc;
    }
    uint64_t during[6];
    b3ref_get_counters(during);
    delete world;
    b3_convexCalls = (u32)(before[0] + during[0]);
    b3_convexCacheHits = (u32)(before[1] + during[1]);
    b3_gjkCalls = (u32)(before[2] + during[2]);
    b3_gjkIters = (u32)(before[3] + during[3]);
    b3_gjkMaxIters = (u32)(before[4] > during[4] ? before[4] : during[4]);
}

// ---------------------------------------------------------------------------------
// A whole reference world, stepped by b3World::Step (world.cpp:115-144), recording the
// narrowphase calls of each step: the box pyramid (20 levels = 210 unit boxes,
// b3BoxHull_identity, collision.cpp:29) on a static 100 x 1 x 100 m box.  Bounce world
// defaults (world.cpp:47-50: gravity -9.8 y, warm starting on, sleeping off), density 1,
// default friction/restitution (shape.h:50-58).
namespace {

struct Pyramid {
    b3World* world = nullptr;
    b3BoxHull ground;
    std::vector<b3Body*> bodies;
    std::unordered_map<const b3Shape*, uint32_t> bodyOf;
    std::unordered_map<const b3Contact*, uint32_t> ids;  // live contacts -> stable id
    std::vector<b3Contact*> gathered;                     // the last gather's contacts
    uint32_t nextId = 0;
    uint64_t saved[6];
};

uint32_t contact_id(Pyramid* p, const b3Contact* c) {
    auto it = p->ids.find(c);
    if (it != p->ids.end()) return it->second;
    const uint32_t id = p->nextId++;
    p->ids[c] = id;
    return id;
}

}  // namespace

extern "C" void* b3ref_pyramid_create(uint32_t levels) {
    Pyramid* p = new Pyramid();
    b3ref_get_counters(p->saved);
    p->world = new b3World();
    p->ground.Set(50.0f, 0.5f, 50.0f);
    {
        b3BodyDef bd;  // static at the origin
        b3Body* g = p->world->CreateBody(bd);
        b3HullShape hs;
        hs.m_hull = &p->ground;
        b3ShapeDef sd;
        sd.shape = &hs;
        b3Shape* s = g->CreateShape(sd);
        p->bodyOf[s] = 0;
        p->bodies.push_back(g);
    }
    for (uint32_t l = 0; l < levels; ++l) {
        for (uint32_t j = 0; j + l < levels; ++j) {
            b3BodyDef bd;
            bd.type = e_dynamicBody;
            bd.position.x = 2.1f * float32(j) + float32(l) * 1.0f - 0.5f * float32(levels) * 2.0f;
            bd.position.y = 3.0f * float32(l + 1);
            bd.position.z = 0.0f;
            b3Body* b = p->world->CreateBody(bd);
            b3HullShape hs;
            hs.m_hull = &b3BoxHull_identity;
            b3ShapeDef sd;
            sd.shape = &hs;
            sd.density = 1.0f;
            b3Shape* s = b->CreateShape(sd);
            p->bodyOf[s] = (uint32_t)p->bodies.size();
            p->bodies.push_back(b);
        }
    }
    return p;
}

extern "C" void b3ref_pyramid_destroy(void* handle) {
    Pyramid* p = (Pyramid*)handle;
    // b3World's constructor, Step and destructor reset the statistics globals
    // (world.cpp:39-45, 122-127, 62-71): leave them as they were before the world.
    delete p->world;
    b3_convexCalls = (u32)p->saved[0];
    b3_convexCacheHits = (u32)p->saved[1];
    b3_gjkCalls = (u32)p->saved[2];
    b3_gjkIters = (u32)p->saved[3];
    b3_gjkMaxIters = (u32)p->saved[4];
    b3_gjkCacheHits = (u32)p->saved[5];
    delete p;
}

extern "C" uint32_t b3ref_pyramid_body_count(void* handle) { return (uint32_t)((Pyramid*)handle)->bodies.size(); }

// Hull records of the world: which = 0 the ground box, 1 the unit box.  counts = V, HE, F.
extern "C" void b3ref_pyramid_hull(void* handle, uint32_t which, uint32_t counts[3], float* centroid,
                                   float* vertices, uint32_t* edges, uint32_t* faces, float* planes) {
    Pyramid* p = (Pyramid*)handle;
    const b3Hull* h = which == 0 ? (const b3Hull*)&p->ground : (const b3Hull*)&b3BoxHull_identity;
    counts[0] = h->vertexCount;
    counts[1] = h->edgeCount;
    counts[2] = h->faceCount;
    if (centroid) std::memcpy(centroid, &h->centroid, 12);
    if (vertices) std::memcpy(vertices, h->vertices, 12ull * h->vertexCount);
    if (edges) std::memcpy(edges, h->edges, 16ull * h->edgeCount);
    if (faces) std::memcpy(faces, h->faces, 4ull * h->faceCount);
    if (planes) std::memcpy(planes, h->planes, 16ull * h->faceCount);
}

// Body transforms and sweep orientations (orientation0, orientation) as the next step's
// narrowphase will read them.
extern "C" void b3ref_pyramid_bodies(void* handle, b3o_xf* xfs, b3o_quat* sweeps) {
    Pyramid* p = (Pyramid*)handle;
    for (size_t i = 0; i < p->bodies.size(); ++i) {
        const b3Body* b = p->bodies[i];
        std::memcpy(xfs + i, &b->m_xf, sizeof(b3o_xf));
        std::memcpy(sweeps + 2 * i, &b->m_sweep.orientation0, 16);
        std::memcpy(sweeps + 2 * i + 1, &b->m_sweep.orientation, 16);
    }
}

// The contacts the next b3World::Step will collide, in m_contactList order, as
// UpdateContacts selects them (contact_manager.cpp:167-231: bodies may collide, one of them
// active, fat AABBs overlapping), after what Step does first when shapes were added
// (world.cpp:129-134).  Writes, for the first cap of them, a stable id, the body pair, the
// transforms and sweeps the collider reads and the cache it starts from; keeps the list in
// p->gathered and returns its length.
extern "C" uint32_t b3ref_pyramid_gather(void* handle, uint32_t cap, uint32_t* ids, uint32_t* bodyPairs,
                                         b3o_xf* xfs, b3o_quat* sweeps, b3o_convex_cache* cacheIn) {
    Pyramid* p = (Pyramid*)handle;
    b3World* w = p->world;
    if (w->m_flags & b3World::e_shapeAddedFlag) {
        w->m_contactMan.FindNewContacts();
        w->m_flags &= ~b3World::e_shapeAddedFlag;
    }
    p->gathered.clear();
    for (b3Contact* c = w->m_contactMan.m_contactList.m_head; c; c = c->m_next) {
        b3Shape* sA = c->m_pair.shapeA;
        b3Shape* sB = c->m_pair.shapeB;
        b3Body* bA = sA->m_body;
        b3Body* bB = sB->m_body;
        const bool activeA = bA->IsAwake() && bA->m_type != e_staticBody;
        const bool activeB = bB->IsAwake() && bB->m_type != e_staticBody;
        if (!bA->ShouldCollide(bB) || !w->m_contactMan.m_broadPhase.TestOverlap(sA->m_broadPhaseID, sB->m_broadPhaseID)) {
            if (activeA || activeB || !bA->ShouldCollide(bB)) p->ids.erase(c);  // destroyed by UpdateContacts
            continue;
        }
        if (!activeA && !activeB) continue;
        const uint32_t k = (uint32_t)p->gathered.size();
        p->gathered.push_back(c);
        if (k >= cap) continue;
        b3ConvexContact* cc = (b3ConvexContact*)c;
        ids[k] = contact_id(p, c);
        bodyPairs[2 * k] = p->bodyOf[sA];
        bodyPairs[2 * k + 1] = p->bodyOf[sB];
        std::memcpy(xfs + 2 * k, &bA->m_xf, sizeof(b3o_xf));
        std::memcpy(xfs + 2 * k + 1, &bB->m_xf, sizeof(b3o_xf));
        std::memcpy(sweeps + 4 * k, &bA->m_sweep.orientation0, 16);
        std::memcpy(sweeps + 4 * k + 1, &bA->m_sweep.orientation, 16);
        std::memcpy(sweeps + 4 * k + 2, &bB->m_sweep.orientation0, 16);
        std::memcpy(sweeps + 4 * k + 3, &bB->m_sweep.orientation, 16);
        std::memcpy(cacheIn + k, &cc->m_cache, sizeof(b3o_convex_cache));
    }
    return (uint32_t)p->gathered.size();
}

// One b3World::Step(1/60, 8, 2) with the reference's own narrowphase.  Records the gathered
// contacts (b3ref_pyramid_gather) and, after the step, each one's cache and manifold.
extern "C" uint32_t b3ref_pyramid_step(void* handle, uint32_t cap, uint32_t* ids, uint32_t* bodyPairs,
                                       b3o_xf* xfs, b3o_quat* sweeps, b3o_convex_cache* cacheIn,
                                       b3o_convex_cache* cacheOut, b3o_manifold* out) {
    Pyramid* p = (Pyramid*)handle;
    const uint32_t n = b3ref_pyramid_gather(handle, cap, ids, bodyPairs, xfs, sweeps, cacheIn);
    p->world->Step(1.0f / 60.0f, 8, 2);
    for (uint32_t k = 0; k < n && k < cap; ++k) {
        b3ConvexContact* cc = (b3ConvexContact*)p->gathered[k];
        std::memcpy(cacheOut + k, &cc->m_cache, sizeof(b3o_convex_cache));
        to_out(cc->m_stackManifold, out + k);
    }
    return n;
}

// ---------------------------------------------------------------------------------
// The closed loop: the same world stepped by b3World::Step, but every b3ConvexContact::
// Collide (convex_contact.cpp:52-63) takes its manifold and cache from a batch the caller
// computed, before the step, for the contacts b3ref_pyramid_gather listed.  The library is
// linked with -Wl,--wrap on b3CollideShapeAndShape (collide.cpp:162-186), so the reference's
// own call sites reach wrap_collide; outside a loop step it forwards to the reference's
// function.  Everything else in the step (b3Contact::Update's tail, contact.cpp:63-140,
// the islands and the solver) is the reference's unchanged code.
namespace {

struct LoopStep {
    std::unordered_map<const b3ConvexCache*, uint32_t> slot;
    const b3o_xf* xfs;
    const b3o_manifold* manifolds;
    const b3o_convex_cache* caches;
    uint32_t served = 0, missed = 0, mismatched = 0;
};
LoopStep* g_loop = nullptr;

}  // namespace

void real_collide(b3Manifold& manifold, const b3Transform& xfA, const b3Shape* shapeA, const b3Transform& xfB,
                  const b3Shape* shapeB, b3ConvexCache* cache)
    __asm__("__real__Z22b3CollideShapeAndShapeR10b3ManifoldRK11b3TransformPK7b3ShapeS3_S6_P13b3ConvexCache");
void wrap_collide(b3Manifold& manifold, const b3Transform& xfA, const b3Shape* shapeA, const b3Transform& xfB,
                  const b3Shape* shapeB, b3ConvexCache* cache)
    __asm__("__wrap__Z22b3CollideShapeAndShapeR10b3ManifoldRK11b3TransformPK7b3ShapeS3_S6_P13b3ConvexCache");

void wrap_collide(b3Manifold& manifold, const b3Transform& xfA, const b3Shape* shapeA, const b3Transform& xfB,
                  const b3Shape* shapeB, b3ConvexCache* cache) {
    LoopStep* L = g_loop;
    if (!L) {
        real_collide(manifold, xfA, shapeA, xfB, shapeB, cache);
        return;
    }
    auto it = L->slot.find(cache);
    if (it == L->slot.end()) {  // a contact the gather did not list: the step stays defined
        ++L->missed;
        real_collide(manifold, xfA, shapeA, xfB, shapeB, cache);
        return;
    }
    const uint32_t k = it->second;
    if (std::memcmp(&xfA, L->xfs + 2 * k, sizeof(b3Transform)) != 0 ||
        std::memcmp(&xfB, L->xfs + 2 * k + 1, sizeof(b3Transform)) != 0)
        ++L->mismatched;
    // What the reference's collider writes: the point count and, per point, the local
    // normal and points and the key; the impulses stay as b3Manifold::Initialize set them.
    const b3o_manifold& m = L->manifolds[k];
    manifold.pointCount = m.pointCount;
    for (u32 i = 0; i < m.pointCount; ++i) {
        b3ManifoldPoint& q = manifold.points[i];
        const b3o_mpoint& s = m.points[i];
        std::memcpy(&q.localNormal1, &s.localNormal1, 12);
        std::memcpy(&q.localPoint1, &s.localPoint1, 12);
        std::memcpy(&q.localPoint2, &s.localPoint2, 12);
        q.key.triangleKey = s.triangleKey;
        std::memcpy(&q.key.key1, s.key, 8);
        std::memcpy(&q.key.key2, s.key + 2, 8);
    }
    std::memcpy(cache, L->caches + k, sizeof(b3ConvexCache));
    ++L->served;
}

// One b3World::Step(1/60, 8, 2) in the closed loop.  n, xfs [n][2], manifolds [n] and
// caches [n] describe the contacts of the last b3ref_pyramid_gather in its order (xfs as it
// wrote them; a contact whose transforms differ at Collide time is counted as mismatched).
// stats = served, missed (collided but not gathered), mismatched.  Returns -1 when n is not
// the gathered count.
extern "C" int b3ref_pyramid_loop_step(void* handle, uint32_t n, const b3o_xf* xfs, const b3o_manifold* manifolds,
                                       const b3o_convex_cache* caches, uint32_t stats[3]) {
    Pyramid* p = (Pyramid*)handle;
    if (n != p->gathered.size()) return -1;
    LoopStep L;
    L.xfs = xfs, L.manifolds = manifolds, L.caches = caches;
    for (uint32_t k = 0; k < n; ++k) L.slot[&((b3ConvexContact*)p->gathered[k])->m_cache] = k;
    g_loop = &L;
    p->world->Step(1.0f / 60.0f, 8, 2);
    g_loop = nullptr;
    stats[0] = L.served, stats[1] = L.missed, stats[2] = L.mismatched;
    return 0;
}

// Per body, 32 floats: transform (12), sweep world centre and orientation (3 + 4), their
// values at the step's start (3 + 4), linear and angular velocity (3 + 3).
extern "C" void b3ref_pyramid_state(void* handle, float* out) {
    Pyramid* p = (Pyramid*)handle;
    for (size_t i = 0; i < p->bodies.size(); ++i) {
        const b3Body* b = p->bodies[i];
        float* o = out + 32 * i;
        std::memcpy(o, &b->m_xf, 48);
        std::memcpy(o + 12, &b->m_sweep.worldCenter, 12);
        std::memcpy(o + 15, &b->m_sweep.orientation, 16);
        std::memcpy(o + 19, &b->m_sweep.worldCenter0, 12);
        std::memcpy(o + 22, &b->m_sweep.orientation0, 16);
        std::memcpy(o + 26, &b->m_linearVelocity, 12);
        std::memcpy(o + 29, &b->m_angularVelocity, 12);
    }
}

// ---------------------------------------------------------------------------------
// Persistent mesh contacts: real b3MeshContact objects kept across calls and driven as
// b3World does each step: Collide in UpdateContacts (contact.cpp:41-140), then after the
// solver SynchronizeShapes and FindNewPairs (contact_manager.cpp:144-162).  One body image
// per side is re-posed before each contact's call (a contact keeps only m_aabbA, its
// triangle caches and manifolds).
namespace {

struct RefMeshContacts {
    Scene scene;
    b3Mesh mesh;
    b3World* world = nullptr;
    b3Body* bodyA = nullptr;
    b3Body* bodyB = nullptr;
    b3MeshShape meshShape;
    std::vector<b3MeshContact*> contacts;
    uint64_t saved[6];
};

void pose(b3Body* b, const b3o_xf* xf) { std::memcpy(&b->m_xf, xf, sizeof(b3Transform)); }

}  // namespace

extern "C" void* b3ref_mesh_contacts_create(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* m,
                                            uint32_t n, const uint32_t* shapeA, const b3o_xf* xfs) {
    RefMeshContacts* r = new RefMeshContacts();
    build_hulls(pool, r->scene.hulls);
    make_mesh(r->mesh, m->vertices, m->vertexCount, m->triangleCount, m->triangles);
    b3ref_get_counters(r->saved);
    r->world = new b3World();
    b3BodyDef bdA;
    bdA.type = e_dynamicBody;
    r->bodyA = r->world->CreateBody(bdA);
    b3BodyDef bdB;
    r->bodyB = r->world->CreateBody(bdB);
    r->meshShape.m_mesh = &r->mesh;
    r->meshShape.m_body = r->bodyB;
    for (uint32_t i = 0; i < n; ++i) {
        b3Shape* A = make_shape(shapes[shapeA[i]], r->scene.hulls);
        r->scene.shapes.push_back(A);
        A->m_body = r->bodyA;
        pose(r->bodyA, xfs + 2 * (size_t)i);
        pose(r->bodyB, xfs + 2 * (size_t)i + 1);
        b3MeshContact* mc = new b3MeshContact(A, &r->meshShape);
        mc->m_pair.shapeA = A;
        mc->m_pair.shapeB = &r->meshShape;
        mc->FindNewPairs();
        r->contacts.push_back(mc);
    }
    // The world's constructor reset the statistics globals: keep the totals running.
    b3_convexCalls = (u32)r->saved[0];
    b3_convexCacheHits = (u32)r->saved[1];
    b3_gjkCalls = (u32)r->saved[2];
    b3_gjkIters = (u32)r->saved[3];
    b3_gjkMaxIters = (u32)r->saved[4];
    b3_gjkCacheHits = (u32)r->saved[5];
    return r;
}

extern "C" void b3ref_mesh_contacts_collide(void* handle, const b3o_xf* xfs, const b3o_quat* sweeps, b3o_manifold* out,
                                            uint32_t* manifoldCounts) {
    RefMeshContacts* r = (RefMeshContacts*)handle;
    const b3Quat identity(0.0f, 0.0f, 0.0f, 1.0f);
    for (size_t i = 0; i < r->contacts.size(); ++i) {
        b3MeshContact* mc = r->contacts[i];
        pose(r->bodyA, xfs + 2 * i);
        pose(r->bodyB, xfs + 2 * i + 1);
        b3Sweep& sa = r->bodyA->m_sweep;
        b3Sweep& sb = r->bodyB->m_sweep;
        if (sweeps) {
            std::memcpy(&sa.orientation0, sweeps + 4 * i, 16);
            std::memcpy(&sa.orientation, sweeps + 4 * i + 1, 16);
            std::memcpy(&sb.orientation0, sweeps + 4 * i + 2, 16);
            std::memcpy(&sb.orientation, sweeps + 4 * i + 3, 16);
        } else {
            sa.orientation0 = sa.orientation = sb.orientation0 = sb.orientation = identity;
        }
        mc->m_manifoldCount = 0;
        for (u32 k = 0; k < mc->m_manifoldCapacity; ++k) mc->m_manifolds[k].Initialize();
        mc->Collide();
        manifoldCounts[i] = mc->m_manifoldCount;
        for (uint32_t k = 0; k < 3; ++k) {
            if (k < mc->m_manifoldCount) to_out(mc->m_manifolds[k], out + 3 * i + k);
            else std::memset(out + 3 * i + k, 0, sizeof(b3o_manifold));
        }
    }
}

extern "C" void b3ref_mesh_contacts_test_overlap(void* handle, const b3o_xf* xfs, uint32_t* overlap) {
    RefMeshContacts* r = (RefMeshContacts*)handle;
    for (size_t i = 0; i < r->contacts.size(); ++i) {
        pose(r->bodyA, xfs + 2 * i);
        pose(r->bodyB, xfs + 2 * i + 1);
        overlap[i] = r->contacts[i]->TestOverlap() ? 1u : 0u;
    }
}

extern "C" void b3ref_mesh_contacts_synchronize(void* handle, const b3o_xf* xfs, const float* centers0) {
    RefMeshContacts* r = (RefMeshContacts*)handle;
    for (size_t i = 0; i < r->contacts.size(); ++i) {
        b3MeshContact* mc = r->contacts[i];
        pose(r->bodyA, xfs + 2 * i);
        pose(r->bodyB, xfs + 2 * i + 1);
        std::memcpy(&r->bodyA->m_sweep.worldCenter0, centers0 + 3 * i, 12);
        mc->SynchronizeShapes();
        mc->FindNewPairs();
    }
}

extern "C" void b3ref_mesh_contacts_read(void* handle, uint32_t cap, uint32_t* triCounts, uint32_t* triLists,
                                         b3o_convex_cache* caches, float* fat) {
    RefMeshContacts* r = (RefMeshContacts*)handle;
    for (size_t i = 0; i < r->contacts.size(); ++i) {
        const b3MeshContact* mc = r->contacts[i];
        triCounts[i] = mc->m_triangleCount;
        for (uint32_t k = 0; k < mc->m_triangleCount && k < cap; ++k) {
            triLists[i * cap + k] = mc->m_triangles[k].index;
            std::memcpy(caches + i * cap + k, &mc->m_triangles[k].cache, sizeof(b3o_convex_cache));
        }
        std::memcpy(fat + 6 * i, &mc->m_aabbA.m_lower, 12);
        std::memcpy(fat + 6 * i + 3, &mc->m_aabbA.m_upper, 12);
    }
}

extern "C" void b3ref_mesh_contacts_destroy(void* handle) {
    RefMeshContacts* r = (RefMeshContacts*)handle;
    uint64_t now[6];
    b3ref_get_counters(now);
    for (b3MeshContact* mc : r->contacts) delete mc;
    delete r->world;  // resets the statistics globals: keep them
    b3_convexCalls = (u32)now[0];
    b3_convexCacheHits = (u32)now[1];
    b3_gjkCalls = (u32)now[2];
    b3_gjkIters = (u32)now[3];
    b3_gjkMaxIters = (u32)now[4];
    b3_gjkCacheHits = (u32)now[5];
    delete r;
}

// ---------------------------------------------------------------- the step before
// The reference's own b3BroadPhase, driven call by call.  User data is carried as an integer
// in the void* the reference stores; FindPairs' callback records the (userData1, userData2)
// it would pass to b3ContactManager::AddPair, with the proxy ids of the pair.
namespace {
struct PairRecorder {
    b3BroadPhase* bp;
    std::vector<uint32_t> out;
    void AddPair(void* a, void* b) {
        out.push_back((uint32_t)(uintptr_t)a);
        out.push_back((uint32_t)(uintptr_t)b);
    }
};
b3AABB3 ref_aabb(const b3o_aabb* a) {
    b3AABB3 r;
    r.m_lower = b3Vec3(a->lower.x, a->lower.y, a->lower.z);
    r.m_upper = b3Vec3(a->upper.x, a->upper.y, a->upper.z);
    return r;
}
// user data 0 is a null pointer in the reference; keep every value distinct from it
void* user_ptr(uint32_t u) { return (void*)((uintptr_t)u + 1u); }
uint32_t user_of(void* p) { return (uint32_t)((uintptr_t)p - 1u); }
}  // namespace

extern "C" void* b3ref_bp_create(void) { return new b3BroadPhase(); }
extern "C" void b3ref_bp_destroy(void* bp) { delete (b3BroadPhase*)bp; }
extern "C" uint32_t b3ref_bp_create_proxy(void* bp, const b3o_aabb* aabb, uint32_t user) {
    return ((b3BroadPhase*)bp)->CreateProxy(ref_aabb(aabb), user_ptr(user));
}
extern "C" void b3ref_bp_destroy_proxy(void* bp, uint32_t id) { ((b3BroadPhase*)bp)->DestroyProxy(id); }
extern "C" void b3ref_bp_touch_proxy(void* bp, uint32_t id) { ((b3BroadPhase*)bp)->TouchProxy(id); }
extern "C" int b3ref_bp_move_proxy(void* bp, uint32_t id, const b3o_aabb* aabb, const b3o_vec3* d) {
    return ((b3BroadPhase*)bp)->MoveProxy(id, ref_aabb(aabb), b3Vec3(d->x, d->y, d->z)) ? 1 : 0;
}
extern "C" uint32_t b3ref_bp_find_pairs(void* handle, uint32_t cap, uint32_t* out) {
    b3BroadPhase* bp = (b3BroadPhase*)handle;
    PairRecorder rec;
    rec.bp = bp;
    bp->FindPairs(&rec);
    // The proxy ids of each reported pair are in the broadphase's sorted, unique pair buffer
    // (broad_phase.h:188-207), one entry per AddPair call, in the same order.
    const uint32_t n = (uint32_t)(rec.out.size() / 2);
    uint32_t k = 0, index = 0;
    while (index < bp->m_pairCount && k < n) {
        const b3Pair* p = bp->m_pairs + index;
        if (k < cap) {
            out[4 * k] = p->proxy1;
            out[4 * k + 1] = p->proxy2;
            out[4 * k + 2] = rec.out[2 * k] - 1u;
            out[4 * k + 3] = rec.out[2 * k + 1] - 1u;
        }
        ++k;
        ++index;
        while (index < bp->m_pairCount && bp->m_pairs[index].proxy1 == p->proxy1 && bp->m_pairs[index].proxy2 == p->proxy2)
            ++index;
    }
    return n;
}
extern "C" int b3ref_bp_test_overlap(void* bp, uint32_t a, uint32_t b) {
    return ((b3BroadPhase*)bp)->TestOverlap(a, b) ? 1 : 0;
}
extern "C" void b3ref_bp_get_aabb(void* bp, uint32_t id, b3o_aabb* out) {
    const b3AABB3& a = ((b3BroadPhase*)bp)->GetAABB(id);
    std::memcpy(&out->lower, &a.m_lower, 12);
    std::memcpy(&out->upper, &a.m_upper, 12);
}

// A b3World whose bodies (one convex shape each, density 0 so no mass update moves the
// local centre) are moved by a script: add bodies (b3World::CreateBody + b3Body::CreateShape,
// world.cpp:86-92, body.cpp:89-121), FindNewContacts and UpdateContacts of its own
// b3ContactManager (contact_manager.cpp:156-231), and a move that sets m_xf and m_sweep and
// runs b3Body::SynchronizeShapes (body.cpp:181-202) as b3World::Solve does after the solver
// (world.cpp:308-336).
namespace {
struct KWorld {
    b3World* world = nullptr;
    std::vector<uint32_t> desc, edges, faces;
    std::vector<float> centroids, vertices, planes;
    std::vector<b3Hull> hulls;
    std::vector<b3o_shape> shapes;
    std::vector<b3Body*> bodies;
    std::unordered_map<const b3Shape*, uint32_t> bodyOf;
    uint64_t saved[6];
};
}  // namespace

extern "C" void* b3ref_kworld_create(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t shapeCount) {
    KWorld* k = new KWorld();
    b3ref_get_counters(k->saved);
    uint32_t nv = 0, ne = 0, nf = 0;
    for (uint32_t h = 0; h < pool->count; ++h) {
        const uint32_t* d = pool->desc + 6 * h;
        nv = std::max(nv, d[0] + d[1]);
        ne = std::max(ne, d[2] + d[3]);
        nf = std::max(nf, d[4] + d[5]);
    }
    k->desc.assign(pool->desc, pool->desc + 6 * (size_t)pool->count);
    k->centroids.assign(pool->centroids, pool->centroids + 3 * (size_t)pool->count);
    k->vertices.assign(pool->vertices, pool->vertices + 3 * (size_t)nv);
    k->edges.assign(pool->edges, pool->edges + 4 * (size_t)ne);
    k->faces.assign(pool->faces, pool->faces + nf);
    k->planes.assign(pool->planes, pool->planes + 4 * (size_t)nf);
    b3o_hull_pool own = {pool->count, k->desc.data(), k->centroids.data(), k->vertices.data(),
                         k->edges.data(), k->faces.data(), k->planes.data()};
    build_hulls(&own, k->hulls);
    k->shapes.assign(shapes, shapes + shapeCount);
    k->world = new b3World();
    return k;
}

extern "C" void b3ref_kworld_destroy(void* h) {
    KWorld* k = (KWorld*)h;
    delete k->world;
    b3_convexCalls = (u32)k->saved[0];
    b3_convexCacheHits = (u32)k->saved[1];
    b3_gjkCalls = (u32)k->saved[2];
    b3_gjkIters = (u32)k->saved[3];
    b3_gjkMaxIters = (u32)k->saved[4];
    b3_gjkCacheHits = (u32)k->saved[5];
    delete k;
}

extern "C" void b3ref_kworld_add_bodies(void* h, uint32_t n, const uint32_t* shapeIdx, const uint32_t* types,
                                        const b3o_xf* xfs, const b3o_sweep* sweeps, uint32_t* proxyIds) {
    KWorld* k = (KWorld*)h;
    static_assert(sizeof(b3Sweep) == sizeof(b3o_sweep), "sweep layout");
    for (uint32_t i = 0; i < n; ++i) {
        b3BodyDef bd;
        bd.type = (b3BodyType)types[i];
        b3Body* b = k->world->CreateBody(bd);
        std::memcpy(&b->m_xf, xfs + i, sizeof(b3Transform));
        std::memcpy(&b->m_sweep, sweeps + i, sizeof(b3Sweep));
        Scene tmp;  // a template shape; CreateShape clones it (b3Shape::Create, shape.cpp)
        b3Shape* proto = make_shape(k->shapes[shapeIdx[i]], k->hulls);
        b3ShapeDef sd;
        sd.shape = proto;
        b3Shape* s = b->CreateShape(sd);
        tmp.shapes.push_back(proto);
        k->bodyOf[s] = (uint32_t)k->bodies.size();
        k->bodies.push_back(b);
        if (proxyIds) proxyIds[i] = s->m_broadPhaseID;
    }
}

extern "C" uint32_t b3ref_kworld_find_new_contacts(void* h) {
    KWorld* k = (KWorld*)h;
    const uint32_t before = k->world->m_contactMan.m_contactList.m_count;
    k->world->m_contactMan.FindNewContacts();
    k->world->m_flags &= ~b3World::e_shapeAddedFlag;
    return k->world->m_contactMan.m_contactList.m_count - before;
}

extern "C" void b3ref_kworld_update_contacts(void* h, uint32_t stats[3]) {
    KWorld* k = (KWorld*)h;
    b3ContactManager& cm = k->world->m_contactMan;
    uint32_t before = cm.m_contactList.m_count, active = 0, skipped = 0;
    for (b3Contact* c = cm.m_contactList.m_head; c; c = c->m_next) {
        b3Body* bA = c->m_pair.shapeA->m_body;
        b3Body* bB = c->m_pair.shapeB->m_body;
        const bool aA = bA->IsAwake() && bA->m_type != e_staticBody;
        const bool aB = bB->IsAwake() && bB->m_type != e_staticBody;
        if (bA->ShouldCollide(bB) && !aA && !aB) ++skipped;
        (void)active;
    }
    cm.UpdateContacts();
    const uint32_t after = cm.m_contactList.m_count;
    if (stats) stats[0] = before - after, stats[1] = after - skipped, stats[2] = skipped;
}

extern "C" void b3ref_kworld_move_bodies(void* h, uint32_t n, const uint32_t* bodies, const b3o_xf* xfs,
                                         const b3o_sweep* sweeps) {
    KWorld* k = (KWorld*)h;
    for (uint32_t i = 0; i < n; ++i) {
        b3Body* b = k->bodies[bodies[i]];
        std::memcpy(&b->m_xf, xfs + i, sizeof(b3Transform));
        std::memcpy(&b->m_sweep, sweeps + i, sizeof(b3Sweep));
        b->SynchronizeShapes();
    }
}

extern "C" uint32_t b3ref_kworld_contacts(void* h, uint32_t cap, uint32_t* bodyPairs, b3o_manifold* out,
                                          b3o_convex_cache* caches) {
    KWorld* k = (KWorld*)h;
    uint32_t i = 0;
    for (b3Contact* c = k->world->m_contactMan.m_contactList.m_head; c; c = c->m_next, ++i) {
        if (i >= cap) continue;
        if (bodyPairs) {
            bodyPairs[2 * i] = k->bodyOf[c->m_pair.shapeA];
            bodyPairs[2 * i + 1] = k->bodyOf[c->m_pair.shapeB];
        }
        b3ConvexContact* cc = (b3ConvexContact*)c;
        if (out) to_out(cc->m_stackManifold, out + i);
        if (caches) std::memcpy(caches + i, &cc->m_cache, sizeof(b3o_convex_cache));
    }
    return i;
}

extern "C" void b3ref_kworld_proxies(void* h, uint32_t n, const uint32_t* bodies, uint32_t* ids, b3o_aabb* fat) {
    KWorld* k = (KWorld*)h;
    for (uint32_t i = 0; i < n; ++i) {
        const b3Shape* s = k->bodies[bodies[i]]->m_shapeList.m_head;
        if (ids) ids[i] = s->m_broadPhaseID;
        if (fat) b3ref_bp_get_aabb(&k->world->m_contactMan.m_broadPhase, s->m_broadPhaseID, fat + i);
    }
}
