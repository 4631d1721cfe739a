/*
 * TEST INFRASTRUCTURE ONLY - the oracle for the contact narrowphase.
 *
 * This header declares (1) the plain-C restatement of the reference narrowphase
 * (b3oracle.c -> liboracle.so) and (2) the extern "C" driver over the compiled
 * reference (ref_driver.cpp -> _ref/libb3ref.so).  Only tests, the smoke check and
 * the CPU-baseline leg of bench.py load these libraries, and only as checkers.
 *
 * Exchange records are byte-compatible with the reference types:
 *   b3o_vec3      == b3Vec3            (vec3.h:106)
 *   b3o_xf        == b3Transform       (transform.h:56-57; rotation columns x,y,z then position)
 *   b3o_quat      == b3Quat            (quat.h:141)
 *   b3o_halfedge  == b3HalfEdge        (hull.h:29-35)
 *   b3o_plane     == b3Plane           (geometry.h:72-73)
 *   b3o_feature_cache == b3SATFeaturePair (sat.h:74-80)
 *   b3o_simplex_cache == b3SimplexCache   (gjk.h:85-92)
 * The manifold record keeps the fields the narrowphase writes (manifold.h:66-93):
 * the three local vectors, the feature key (triangleKey + b3FeaturePair words in
 * struct order inEdge1,outEdge1,inEdge2,outEdge2, clip.h:28-34) and pointCount.
 */
#ifndef B3ORACLE_H
#define B3ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z; } b3o_vec3;
typedef struct { b3o_vec3 c[3]; b3o_vec3 p; } b3o_xf;
typedef struct { float x, y, z, w; } b3o_quat;
typedef struct { b3o_vec3 n; float d; } b3o_plane;
typedef struct { uint32_t origin, twin, face, next; } b3o_halfedge;

/* Flat hull pool. Hull h: desc[6h..6h+5] = vertexOffset, vertexCount, edgeOffset,
 * edgeCount (half-edges), faceOffset, faceCount. Edge/face indices are local to the hull. */
typedef struct {
    uint32_t count;
    const uint32_t* desc;
    const float* centroids;       /* [count][3] */
    const float* vertices;        /* [*][3]     */
    const uint32_t* edges;        /* [*][4]     */
    const uint32_t* faces;        /* [*]        */
    const float* planes;          /* [*][4]     */
} b3o_hull_pool;

/* Shape types follow b3ShapeType (shape.h:32-39). */
enum { B3O_SPHERE = 0, B3O_CAPSULE = 1, B3O_HULL = 2, B3O_MESH = 3 };

/* A convex shape: sphere (v[0..2] = centre), capsule (v[0..5] = centres), hull (hull index). */
typedef struct {
    uint32_t type;
    uint32_t hull;
    float radius;
    float v[6];
    uint32_t pad[3];
} b3o_shape;

typedef struct {
    b3o_vec3 localNormal1;
    b3o_vec3 localPoint1;
    b3o_vec3 localPoint2;
    uint32_t triangleKey;
    uint32_t key[4];
} b3o_mpoint;

typedef struct {
    uint32_t pointCount;
    uint32_t pad[3];
    b3o_mpoint points[4];
} b3o_manifold;

typedef struct { uint32_t state, type, index1, index2; } b3o_feature_cache;
typedef struct {
    float metric;
    uint32_t iterations;
    uint16_t count;
    uint8_t index1[4];
    uint8_t index2[4];
} b3o_simplex_cache;
typedef struct { b3o_simplex_cache simplex; b3o_feature_cache feature; } b3o_convex_cache;

/* Feature cache enums (sat.h:58-72). */
enum { B3O_SEPARATION = 0, B3O_OVERLAP = 1, B3O_EMPTY = 2 };
enum { B3O_EDGE1 = 0, B3O_FACE1 = 1, B3O_FACE2 = 2 };

/* ---- the step after the narrowphase: warm-start matching and contact constraints ---- */
/* A body as b3ContactSolver sees it: island position (sweep world centre, orientation) and
 * velocity (b3Position / b3Velocity, time_step.h), b3Body::m_invMass, the world inverse
 * inertia b3Body::m_worldInvI and m_sweep.localCenter. */
typedef struct {
    b3o_vec3 x;
    b3o_quat q;
    b3o_vec3 v, w;
    float invMass;
    float invI[9];            /* b3Mat33 columns */
    b3o_vec3 localCenter;
} b3o_solver_body;
/* The two shapes' m_radius, m_friction and m_restitution (mixed by b3MixFriction /
 * b3MixRestitution, contact_solver.h:88-96). */
typedef struct { float radiusA, radiusB, frictionA, frictionB, restitutionA, restitutionB, reserved[2]; } b3o_contact_material;
/* The solver-owned fields of b3Manifold (manifold.h:66-93). */
typedef struct {
    float normalImpulse[4];
    uint32_t persisting[4];
    float tangentImpulse[2];
    float motorImpulse;
    uint32_t reserved;
} b3o_manifold_impulses;
/* b3VelocityConstraintPoint / b3VelocityConstraintManifold (contact_solver.h:58-84) with the
 * contact's mixed friction and restitution. */
typedef struct { b3o_vec3 rA, rB, normal; float normalMass, normalImpulse, velocityBias; } b3o_velocity_point;
typedef struct {
    b3o_velocity_point points[4];
    b3o_vec3 rA, rB, normal, tangent1, tangent2;
    float tangentMass[4];     /* b3Mat22 columns x, y */
    float tangentImpulse[2];
    float motorImpulse, motorMass;
    uint32_t pointCount;
    float friction, restitution;
    uint32_t reserved[2];
} b3o_velocity_manifold;

/* For n contacts of M manifold slots each: the new manifolds' impulses from
 * b3Manifold::Initialize() and, with warm starting, Initialize(old) of every old manifold in
 * turn (contact.cpp:63-91, manifold.cpp:21-55), then b3ContactSolver::InitializeConstraints
 * (contact_solver.cpp:70-276) for each new manifold.  Contact i joins bodies bodies[2i],
 * bodies[2i+1] (indices into state); counts / oldCounts give its manifold counts. */
void b3o_contact_init(uint32_t n, uint32_t M, const uint32_t* bodies, const b3o_solver_body* state,
                      const b3o_contact_material* material, const uint32_t* counts, const b3o_manifold* manifolds,
                      const uint32_t* oldCounts, const b3o_manifold* oldManifolds,
                      const b3o_manifold_impulses* oldImpulses, int warmStarting, b3o_manifold_impulses* impulses,
                      b3o_velocity_manifold* out);
void b3ref_contact_init(uint32_t n, uint32_t M, const uint32_t* bodies, const b3o_solver_body* state,
                        const b3o_contact_material* material, const uint32_t* counts, const b3o_manifold* manifolds,
                        const uint32_t* oldCounts, const b3o_manifold* oldManifolds,
                        const b3o_manifold_impulses* oldImpulses, int warmStarting, b3o_manifold_impulses* impulses,
                        b3o_velocity_manifold* out);

/* ---- plain-C restatement (liboracle.so) ---- */

/* Collide n shape pairs as b3CollideShapeAndShape (collide.cpp:162-188) with the
 * global convex cache on (collide_hulls.cpp:291).  pairs[2i], pairs[2i+1] index
 * shapes[]; xfs[2i], xfs[2i+1] are the transforms; sweeps (may be NULL) gives per pair
 * orientation0/orientation of body A then body B (4 quats) for the cached face
 * rebuild; caches (may be NULL: fresh empty caches) are read and updated in place. */
void b3o_collide_pairs(const b3o_hull_pool* pool, const b3o_shape* shapes,
                       uint32_t n, const uint32_t* pairs, const b3o_xf* xfs,
                       const b3o_quat* sweeps, b3o_convex_cache* caches,
                       b3o_manifold* out);

/* Sensor overlap of n convex shape pairs as b3TestOverlap (collide.cpp:84-95), the test
 * b3ConvexContact::TestOverlap runs for contacts with a sensor shape (convex_contact.cpp:39-48):
 * GJK with the radii applied and each pair's simplex cache read and written (gjk.cpp:579-841).
 * overlap[i] = 1 when the distance is within 10 FLT_EPSILON. */
void b3o_test_overlap(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t n, const uint32_t* pairs,
                      const b3o_xf* xfs, b3o_convex_cache* caches, uint32_t* overlap);

/* Separation queries only (sat.cpp:33-58, 110-173) for hull pairs, for kernel-level checks.
 * out[8i..8i+7] = face1.index, face1.sep(bits), face2.index, face2.sep(bits),
 *                 edge.index1, edge.index2, edge.sep(bits), 0 */
void b3o_hull_queries(const b3o_hull_pool* pool, uint32_t n, const uint32_t* hullPairs,
                      const b3o_xf* xfs, uint32_t* out);

/* A b3Mesh (mesh.h:41-99): vertices [V][3], triangles [T][3] (v1, v2, v3) and the
 * b3StaticTree nodes [2T-1][8] = aabb lower xyz, upper xyz, child1, child2|index
 * (static_tree.h:62-77; child1 = 0xFFFFFFFF marks a leaf). */
typedef struct {
    uint32_t vertexCount;
    const float* vertices;
    uint32_t triangleCount;
    const uint32_t* triangles;
    uint32_t nodeCount;
    const float* nodes;
} b3o_mesh;

/* Static tree over the triangle AABBs (static_tree.cpp:42-180); nodesOut holds 2T-1 nodes. */
void b3o_mesh_build_tree(const float* vertices, uint32_t triangleCount, const uint32_t* triangles, float* nodesOut);

/* n freshly created mesh contacts (b3MeshContact, mesh_contact.cpp): shape A = shapes[shapeA[i]],
 * xfs[2i] = body A, xfs[2i+1] = mesh body.  Midphase triangle lists (DFS order) go to
 * triLists[i*cap ...] with their full count in triCounts[i]; the clustered output is
 * out[3i .. 3i+2] with manifoldCounts[i] manifolds. */
void b3o_collide_mesh(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* mesh, uint32_t n,
                      const uint32_t* shapeA, const b3o_xf* xfs, uint32_t cap, uint32_t* triCounts,
                      uint32_t* triLists, b3o_manifold* out, uint32_t* manifoldCounts);
void b3ref_collide_mesh(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* mesh, uint32_t n,
                        const uint32_t* shapeA, const b3o_xf* xfs, uint32_t cap, uint32_t* triCounts,
                        uint32_t* triLists, b3o_manifold* out, uint32_t* manifoldCounts);
/* The reference's own tree: b3Mesh::BuildTree, nodes copied out in the same 8-word layout. */
void b3ref_mesh_build_tree(const float* vertices, uint32_t vertexCount, uint32_t triangleCount,
                           const uint32_t* triangles, float* nodesOut);

/* Persistent mesh contacts (b3MeshContact across steps, mesh_contact.cpp:30-275): create is
 * the constructor and FindNewPairs; collide is Collide with every triangle's cache (sweeps may
 * be NULL: identity orientations; per contact body A's orientation0/orientation then the mesh
 * body's); synchronize is SynchronizeShapes, MoveAABB and FindNewPairs with body A's sweep
 * worldCenter0 in centers0[3i..]; read copies the first cap triangles and caches of each list
 * and the fat AABBs [n][6].  The oracle version reads the tree nodes from mesh->nodes; the
 * compiled reference's builds its own b3Mesh tree.  The arrays passed to create must outlive
 * the set. */
typedef struct b3o_mesh_contacts b3o_mesh_contacts;
b3o_mesh_contacts* b3o_mesh_contacts_create(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* mesh,
                                            uint32_t n, const uint32_t* shapeA, const b3o_xf* xfs);
void b3o_mesh_contacts_collide(b3o_mesh_contacts* h, const b3o_xf* xfs, const b3o_quat* sweeps, b3o_manifold* out,
                               uint32_t* manifoldCounts);
void b3o_mesh_contacts_synchronize(b3o_mesh_contacts* h, const b3o_xf* xfs, const float* centers0);
/* b3MeshContact::TestOverlap (mesh_contact.cpp:191-215) for each contact: the sensor test
 * against the listed triangles in order, each with its cache, stopping at the first
 * overlap.  overlap[i] = 1 / 0. */
void b3o_mesh_contacts_test_overlap(b3o_mesh_contacts* h, const b3o_xf* xfs, uint32_t* overlap);
void b3ref_mesh_contacts_test_overlap(void* handle, const b3o_xf* xfs, uint32_t* overlap);
void b3o_mesh_contacts_read(b3o_mesh_contacts* h, uint32_t cap, uint32_t* triCounts, uint32_t* triLists,
                            b3o_convex_cache* caches, float* fat);
void b3o_mesh_contacts_destroy(b3o_mesh_contacts* h);
void* b3ref_mesh_contacts_create(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* mesh, uint32_t n,
                                 const uint32_t* shapeA, const b3o_xf* xfs);
void b3ref_mesh_contacts_collide(void* h, const b3o_xf* xfs, const b3o_quat* sweeps, b3o_manifold* out,
                                 uint32_t* manifoldCounts);
void b3ref_mesh_contacts_synchronize(void* h, const b3o_xf* xfs, const float* centers0);
void b3ref_mesh_contacts_read(void* h, uint32_t cap, uint32_t* triCounts, uint32_t* triLists, b3o_convex_cache* caches,
                              float* fat);
void b3ref_mesh_contacts_destroy(void* h);

/* Totals of the reference's statistics globals since the last reset: b3_convexCalls,
 * b3_convexCacheHits, b3_gjkCalls, b3_gjkIters, b3_gjkMaxIters. */
void b3o_get_counters(uint64_t out[6]);
void b3o_reset_counters(void);

/* FP32 ops executed by the last b3o_collide_pairs call (liboracle_count.so only). */
uint64_t b3o_flop_count(void);
/* Per phase: 0 feature-cache read/rebuild, 1 separation queries, 2 contact building. */
uint64_t b3o_flop_count_phase(int phase);
void b3o_flop_reset(void);

/* ---- the step before the narrowphase: broadphase and contact manager ---- */
typedef struct { b3o_vec3 lower, upper; } b3o_aabb;                       /* b3AABB3 aabb3.h:26-28 */
typedef struct {                                                          /* b3Sweep transform.h:202-222 */
    b3o_vec3 localCenter;
    b3o_quat orientation0;
    b3o_vec3 worldCenter0;
    float t0;
    b3o_quat orientation;
    b3o_vec3 worldCenter;
} b3o_sweep;

/* b3BroadPhase over b3DynamicTree, restated (oracle) and the reference's own (b3ref_bp_*).
 * find_pairs writes up to cap pairs of 4 words (proxy1, proxy2, user1, user2) in the order
 * b3ContactManager::AddPair receives them and returns how many there are. */
typedef struct b3o_broadphase b3o_broadphase;
b3o_broadphase* b3o_bp_create(void);
void b3o_bp_destroy(b3o_broadphase* bp);
uint32_t b3o_bp_create_proxy(b3o_broadphase* bp, const b3o_aabb* aabb, uint32_t user);
void b3o_bp_destroy_proxy(b3o_broadphase* bp, uint32_t id);
void b3o_bp_touch_proxy(b3o_broadphase* bp, uint32_t id);
int b3o_bp_move_proxy(b3o_broadphase* bp, uint32_t id, const b3o_aabb* aabb, const b3o_vec3* displacement);
uint32_t b3o_bp_find_pairs(b3o_broadphase* bp, uint32_t cap, uint32_t* out);
int b3o_bp_test_overlap(const b3o_broadphase* bp, uint32_t a, uint32_t b);
void b3o_bp_get_aabb(const b3o_broadphase* bp, uint32_t id, b3o_aabb* out);
void* b3ref_bp_create(void);
void b3ref_bp_destroy(void* bp);
uint32_t b3ref_bp_create_proxy(void* bp, const b3o_aabb* aabb, uint32_t user);
void b3ref_bp_destroy_proxy(void* bp, uint32_t id);
void b3ref_bp_touch_proxy(void* bp, uint32_t id);
int b3ref_bp_move_proxy(void* bp, uint32_t id, const b3o_aabb* aabb, const b3o_vec3* displacement);
uint32_t b3ref_bp_find_pairs(void* bp, uint32_t cap, uint32_t* out);
int b3ref_bp_test_overlap(void* bp, uint32_t a, uint32_t b);
void b3ref_bp_get_aabb(void* bp, uint32_t id, b3o_aabb* out);

/* b3Body::SynchronizeShapes' AABB and displacement for shape shapeIdx[i] on body bodies[i]
 * (body.cpp:181-202); with sweeps null, the AABB at xfs alone (b3Body::CreateShape). */
void b3o_shape_aabbs(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t n, const uint32_t* shapeIdx,
                     const uint32_t* bodies, const b3o_xf* xfs, const b3o_sweep* sweeps, b3o_aabb* out,
                     b3o_vec3* disp);

/* A world of bodies with one convex shape each, moved by a script (no solver):
 * b3ContactManager::FindNewContacts / UpdateContacts and b3Body::SynchronizeShapes, restated
 * (b3o_kworld_*) and the reference's own b3World (b3ref_kworld_*).  Body types follow
 * b3BodyType (0 static, 1 kinematic, 2 dynamic).  stats = destroyed, collided, skipped. */
typedef struct b3o_kworld b3o_kworld;
b3o_kworld* b3o_kworld_create(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t shapeCount);
void b3o_kworld_destroy(b3o_kworld* w);
void b3o_kworld_add_bodies(b3o_kworld* w, uint32_t n, const uint32_t* shapeIdx, const uint32_t* types,
                           const b3o_xf* xfs, const b3o_sweep* sweeps, uint32_t* proxyIds);
uint32_t b3o_kworld_find_new_contacts(b3o_kworld* w);
void b3o_kworld_update_contacts(b3o_kworld* w, uint32_t stats[3]);
void b3o_kworld_move_bodies(b3o_kworld* w, uint32_t n, const uint32_t* bodies, const b3o_xf* xfs,
                            const b3o_sweep* sweeps);
uint32_t b3o_kworld_contacts(const b3o_kworld* w, uint32_t cap, uint32_t* bodyPairs, b3o_manifold* out,
                             b3o_convex_cache* caches);
void b3o_kworld_proxies(const b3o_kworld* w, uint32_t n, const uint32_t* bodies, uint32_t* ids, b3o_aabb* fat);
void* b3ref_kworld_create(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t shapeCount);
void b3ref_kworld_destroy(void* w);
void b3ref_kworld_add_bodies(void* w, uint32_t n, const uint32_t* shapeIdx, const uint32_t* types,
                             const b3o_xf* xfs, const b3o_sweep* sweeps, uint32_t* proxyIds);
uint32_t b3ref_kworld_find_new_contacts(void* w);
void b3ref_kworld_update_contacts(void* w, uint32_t stats[3]);
void b3ref_kworld_move_bodies(void* w, uint32_t n, const uint32_t* bodies, const b3o_xf* xfs,
                              const b3o_sweep* sweeps);
uint32_t b3ref_kworld_contacts(void* w, uint32_t cap, uint32_t* bodyPairs, b3o_manifold* out,
                               b3o_convex_cache* caches);
void b3ref_kworld_proxies(void* w, uint32_t n, const uint32_t* bodies, uint32_t* ids, b3o_aabb* fat);

/* ---- compiled reference driver (_ref/libb3ref.so) ---- */
void b3ref_collide_pairs(const b3o_hull_pool* pool, const b3o_shape* shapes,
                         uint32_t n, const uint32_t* pairs, const b3o_xf* xfs,
                         const b3o_quat* sweeps, b3o_convex_cache* caches,
                         b3o_manifold* out);
void b3ref_hull_queries(const b3o_hull_pool* pool, uint32_t n, const uint32_t* hullPairs,
                        const b3o_xf* xfs, uint32_t* out);

#ifdef __cplusplus
}
#endif
#endif
