/*
 * bounce_amd - MI355X (gfx950) contact narrowphase for the Bounce physics engine.
 *
 * C ABI of the drop-in boundary.  Every entry point returns an int status
 * (B3AMD_OK = 0, negative = error; b3amd_last_error() gives the message) and never
 * aborts: the reference has no status returns (asserts only, settings.h:115), so an
 * error here corresponds to what the reference would have asserted on or silently
 * mis-handled.  All records are plain C structs byte-compatible with the reference
 * types they replace; there are no GPU-runtime types in any signature.
 *
 * Reference call surface replaced (see the per-function notes):
 *   b3CollideShapeAndShape      collide.h:63-66, collide.cpp:162-188
 *   b3Collide{Sphere,Capsule,Hull}And*   collide.h:68-97
 *   b3ConvexCache               collide.h:38-42 (per-contact, persistent)
 *   b3Manifold / b3ManifoldPoint manifold.h:66-93 (output)
 *   b3Hull                      hull.h:37-64 (borrowed, immutable geometry)
 *   b3MeshContact::Collide      mesh_contact.cpp:217-275 (mesh entry)
 * Caller: b3Contact::Update (contact.cpp:41-140) from
 * b3ContactManager::UpdateContacts (contact_manager.cpp:167-231).
 */
#ifndef BOUNCE_AMD_H
#define BOUNCE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define B3AMD_OK 0
#define B3AMD_ERR_ARGUMENT (-1)   /* bad pointer, size or index */
#define B3AMD_ERR_DEVICE (-2)     /* HIP runtime error */
#define B3AMD_ERR_BUDGET (-3)     /* HBM budget of the context exceeded (admission refused) */
#define B3AMD_ERR_UNSUPPORTED (-4)/* record outside the supported limits */
#define B3AMD_ERR_NO_DEVICE (-5)  /* no GPU visible */
#define B3AMD_ERR_OVERFLOW (-6)   /* a per-pair bounded buffer overflowed (see flags) */

/* ---- records (byte-compatible with the reference) ---- */
typedef struct { float x, y, z; } b3amd_vec3;                       /* b3Vec3 vec3.h:106 */
typedef struct { b3amd_vec3 rotation[3]; b3amd_vec3 position; } b3amd_transform; /* b3Transform transform.h:56-57 (rotation columns) */
typedef struct { float x, y, z, w; } b3amd_quat;                   /* b3Quat quat.h:141 */

/* Flat pool of b3Hull records.  Hull h: desc[6h..6h+5] = vertexOffset, vertexCount,
 * edgeOffset, edgeCount (half-edges), faceOffset, faceCount; indices inside a hull are
 * hull-local, exactly as in b3Hull.  Half-edges 2k and 2k+1 must be twins
 * (sat.cpp:123-128).  Limits: vertexCount <= 255, faceCount <= 255, edgeCount <= 4096. */
typedef struct {
    uint32_t count;
    const uint32_t* desc;
    const float* centroids;   /* [count][3]   b3Hull::centroid */
    const float* vertices;    /* [*][3]       b3Vec3 */
    const uint32_t* edges;    /* [*][4]       b3HalfEdge origin, twin, face, next */
    const uint32_t* faces;    /* [*]          b3Face::edge */
    const float* planes;      /* [*][4]       b3Plane normal, offset */
} b3amd_hull_pool;

/* Shape types follow b3ShapeType (shape.h:32-39). */
enum { B3AMD_SPHERE = 0, B3AMD_CAPSULE = 1, B3AMD_HULL = 2, B3AMD_MESH = 3 };

/* One convex shape: b3SphereShape (m_center, m_radius), b3CapsuleShape (m_centers[2],
 * m_radius) or b3HullShape (m_hull -> pool index, m_radius). */
typedef struct {
    uint32_t type;
    uint32_t hull;
    float radius;
    float v[6];
    uint32_t pad[3];          /* ignored: the library fills in its own copy */
} b3amd_shape;

/* b3ManifoldPoint fields written by the narrowphase (manifold.h:66-76); key[] holds the
 * b3FeaturePair words in struct order inEdge1, outEdge1, inEdge2, outEdge2 (clip.h:28-34),
 * i.e. the two u64 of b3ManifoldPointKey::key1/key2 (manifold.h:41-63). */
typedef struct {
    b3amd_vec3 localNormal1;
    b3amd_vec3 localPoint1;
    b3amd_vec3 localPoint2;
    uint32_t triangleKey;
    uint32_t key[4];
} b3amd_manifold_point;

/* b3Manifold (manifold.h:79-93) without the solver-owned impulses; flags carries
 * per-pair status bits (B3AMD_FLAG_*). */
typedef struct {
    uint32_t pointCount;
    uint32_t flags;
    uint32_t reserved[2];
    b3amd_manifold_point points[4];
} b3amd_manifold;

#define B3AMD_FLAG_POLY_OVERFLOW 1u   /* clip polygon exceeded the device bound */

typedef struct { uint32_t state, type, index1, index2; } b3amd_feature_cache; /* b3SATFeaturePair sat.h:74-80 */
typedef struct {
    float metric;
    uint32_t iterations;
    uint16_t count;
    uint8_t index1[4];
    uint8_t index2[4];
} b3amd_simplex_cache;                                                      /* b3SimplexCache gjk.h:85-92 */
typedef struct { b3amd_simplex_cache simplex; b3amd_feature_cache feature; } b3amd_convex_cache; /* b3ConvexCache collide.h:38-42 */

/* b3SATCacheType / b3SATFeatureType (sat.h:58-72). */
enum { B3AMD_CACHE_SEPARATION = 0, B3AMD_CACHE_OVERLAP = 1, B3AMD_CACHE_EMPTY = 2 };
enum { B3AMD_FEATURE_EDGE1 = 0, B3AMD_FEATURE_FACE1 = 1, B3AMD_FEATURE_FACE2 = 2 };

/* ---- context ---- */
typedef struct b3amd_context b3amd_context;

/* One context per caller thread; owns one HIP stream on `device`.  hbm_budget_bytes
 * caps the device memory the context may hold (0 = 64 GiB). */
int b3amd_create(int device, uint64_t hbm_budget_bytes, b3amd_context** out);
int b3amd_destroy(b3amd_context* ctx);
/* Message of the last error on the calling thread. */
const char* b3amd_last_error(void);
/* Bytes of device memory currently held by the context. */
uint64_t b3amd_memory_in_use(const b3amd_context* ctx);

/* ---- geometry registry ----
 * Replaces the borrowed `const b3Hull*` of b3HullShape (hull_shape.h:44): hulls are
 * validated, converted to the device record layout and kept in one device arena; shapes
 * refer to hulls by handle (b3amd_shape.hull).  b3amd_upload_hulls replaces every hull with
 * the pool's, handle i = pool hull i. */
int b3amd_upload_hulls(b3amd_context* ctx, const b3amd_hull_pool* pool);
/* Adds the pool's hulls to the registry without touching the others and writes their handles
 * (freed handles are reused, lowest first).  Only the new records are uploaded; the arena
 * grows by doubling when full, and only then do graphs captured before refuse to launch.
 * The b3Hull a b3HullShape borrows (hull_shape.h:44) becomes this handle. */
int b3amd_hull_add(b3amd_context* ctx, const b3amd_hull_pool* hulls, uint32_t* handles);
/* Frees a hull's record and handle.  Fails while a shape of the current shape table refers
 * to it (the reference frees a b3Hull only when no shape borrows it). */
int b3amd_hull_free(b3amd_context* ctx, uint32_t handle);
/* Registry state: handle slots, live hulls, bytes of live records, arena capacity. */
int b3amd_hull_registry_info(const b3amd_context* ctx, uint32_t* slots, uint32_t* live, uint64_t* used_bytes,
                             uint64_t* capacity_bytes);
/* The shape table that pair indices refer to (b3Shape records of the world). */
int b3amd_upload_shapes(b3amd_context* ctx, uint32_t count, const b3amd_shape* shapes);

/* ---- batch narrowphase ----
 * For each pair i: b3CollideShapeAndShape(out[i], xfs[2i], shapes[pairs[2i]],
 * xfs[2i+1], shapes[pairs[2i+1]], &caches[i]) with the global convex cache on
 * (collide_hulls.cpp:291).  Shape types must satisfy typeA <= typeB
 * (contact_manager.cpp:238-242).  sweeps (may be NULL) gives per pair the
 * orientation0/orientation of body A then body B (4 quats), read only by the cached
 * face rebuild (collide_hulls_cache.cpp:117-126).  caches (may be NULL: fresh empty
 * caches) are read and written back.  out[i] must not alias inputs. */
int b3amd_collide(b3amd_context* ctx, uint32_t n, const uint32_t* pairs,
                  const b3amd_transform* xfs, const b3amd_quat* sweeps,
                  b3amd_convex_cache* caches, b3amd_manifold* out);

/* Same on device-resident buffers (pointers from b3amd_device_alloc), enqueued on the
 * context stream without synchronizing.  d_caches may be NULL only if d_sweeps is NULL. */
int b3amd_collide_device(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs,
                         const b3amd_transform* d_xfs, const b3amd_quat* d_sweeps,
                         b3amd_convex_cache* d_caches, b3amd_manifold* d_out);

/* One pair, synchronously, with the signature of b3CollideShapeAndShape (collide.h:63-66):
 * manifold, transform and shape of A, transform and shape of B, cache.  Shapes are indices
 * into the uploaded shape table; sweeps (may be NULL) are the four orientations the cached
 * face rebuild reads from the bodies (orientation0, orientation of A, then of B,
 * collide_hulls_cache.cpp:117-126).  For call sites that collide one contact at a time;
 * batches belong to b3amd_collide. */
int b3amd_collide_shape_and_shape(b3amd_context* ctx, b3amd_manifold* manifold, const b3amd_transform* xfA,
                                  uint32_t shapeA, const b3amd_transform* xfB, uint32_t shapeB,
                                  const b3amd_quat* sweeps, b3amd_convex_cache* cache);

/* ---- the step after the narrowphase ----
 * Warm-start matching and contact-constraint setup on device-resident manifolds, so that a
 * world's contacts go from the narrowphase to the solver without a host round trip.  For each
 * contact i of n and manifold slot j < counts[i] (of M slots, M = 1 for convex contacts,
 * up to 3 for mesh contacts): the new manifold's impulses are b3Manifold::Initialize() and,
 * when warm_starting, Initialize(old) of each of the contact's oldCounts[i] old manifolds in
 * turn (contact.cpp:63-91, manifold.cpp:21-55: a point whose key an old manifold holds takes
 * that point's normal impulse and is marked persisting; tangent and motor impulses come from
 * the last old manifold); then b3ContactSolver::InitializeConstraints for that manifold
 * (contact_solver.cpp:70-276): the world manifold (manifold.cpp:57-95), per point the
 * anchors, normal effective mass and restitution bias, and the friction anchors, 2x2 tangent
 * mass and twist mass.  Slots j >= counts[i] are not written.  Contact i joins bodies
 * bodies[2i], bodies[2i+1], indices into the solver-body array (the island's b3Position and
 * b3Velocity with the body's mass data). */
typedef struct {
    b3amd_vec3 x;             /* b3Position::x (sweep world centre) */
    b3amd_quat q;             /* b3Position::q */
    b3amd_vec3 v, w;          /* b3Velocity */
    float invMass;            /* b3Body::m_invMass */
    float invI[9];            /* b3Body::m_worldInvI, columns */
    b3amd_vec3 localCenter;   /* b3Body::m_sweep.localCenter */
} b3amd_solver_body;
typedef struct { float radiusA, radiusB, frictionA, frictionB, restitutionA, restitutionB, reserved[2]; } b3amd_contact_material;
/* The solver-owned fields of b3Manifold (manifold.h:66-93). */
typedef struct {
    float normalImpulse[4];
    uint32_t persisting[4];
    float tangentImpulse[2];
    float motorImpulse;
    uint32_t reserved;
} b3amd_manifold_impulses;
/* b3VelocityConstraintPoint / b3VelocityConstraintManifold (contact_solver.h:58-84) with the
 * contact's b3MixFriction / b3MixRestitution. */
typedef struct { b3amd_vec3 rA, rB, normal; float normalMass, normalImpulse, velocityBias; } b3amd_velocity_point;
typedef struct {
    b3amd_velocity_point points[4];
    b3amd_vec3 rA, rB, normal, tangent1, tangent2;
    float tangentMass[4];     /* b3Mat22 columns */
    float tangentImpulse[2];
    float motorImpulse, motorMass;
    uint32_t pointCount;
    float friction, restitution;
    uint32_t reserved[2];
} b3amd_velocity_manifold;
int b3amd_contact_init(b3amd_context* ctx, uint32_t n, uint32_t M, const uint32_t* d_bodies,
                       const b3amd_solver_body* d_state, const b3amd_contact_material* d_material,
                       const uint32_t* d_counts, const b3amd_manifold* d_manifolds, const uint32_t* d_old_counts,
                       const b3amd_manifold* d_old_manifolds, const b3amd_manifold_impulses* d_old_impulses,
                       int warm_starting, b3amd_manifold_impulses* d_impulses,
                       b3amd_velocity_manifold* d_constraints);

/* ---- sensors ----
 * For each pair i: b3TestOverlap(xfs[2i], 0, shapes[pairs[2i]], xfs[2i+1], 0,
 * shapes[pairs[2i+1]], &caches[i]) (collide.cpp:84-95), the test b3Contact::Update runs
 * instead of the narrowphase when either shape is a sensor (contact.cpp:55-61, through
 * b3ConvexContact::TestOverlap, convex_contact.cpp:39-48): GJK distance between the shapes
 * (b3ShapeGJKProxy, collide.cpp:30-82) with the radii applied (gjk.cpp:709-733), warm-started
 * from and written back to each pair's b3SimplexCache (gjk.cpp:743-841; u8 vertex indices),
 * overlapping when within 10 FLT_EPSILON.  Any convex pair with typeA <= typeB, hull-hull
 * included.  overlap[i] = 1 overlapping, 0 apart, B3AMD_OVERLAP_INVALID for a pair the
 * shape table cannot resolve.  caches may be NULL (fresh simplex caches, not kept); only the
 * simplex part of each record is read and written. */
#define B3AMD_OVERLAP_INVALID 2u
int b3amd_test_overlap(b3amd_context* ctx, uint32_t n, const uint32_t* pairs, const b3amd_transform* xfs,
                       b3amd_convex_cache* caches, uint32_t* overlap);
/* Same on device-resident buffers, enqueued on the context stream without synchronizing. */
int b3amd_test_overlap_device(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                              b3amd_convex_cache* d_caches, uint32_t* d_overlap);

/* ---- the step before the narrowphase: broadphase ----
 * b3BroadPhase (broad_phase.h:36-119, broad_phase.cpp) with the fat AABBs in HBM.  Proxy ids
 * are the ones b3BroadPhase::CreateProxy returns for the same sequence of creates and
 * destroys (b3DynamicTree node ids, dynamic_tree.cpp:41-114; moves do not change them), and
 * find_pairs reports exactly the pairs FindPairs hands to b3ContactManager::AddPair, in the
 * same order (broad_phase.h:160-208: sorted by (proxy1, proxy2), unique).  Destroy a
 * broadphase before its context. */
typedef struct { b3amd_vec3 lower, upper; } b3amd_aabb;                    /* b3AABB3 aabb3.h:26-28 */
/* b3Sweep (transform.h:202-222), field for field. */
typedef struct {
    b3amd_vec3 localCenter;
    b3amd_quat orientation0;
    b3amd_vec3 worldCenter0;
    float t0;
    b3amd_quat orientation;
    b3amd_vec3 worldCenter;
} b3amd_sweep;
typedef struct b3amd_broadphase b3amd_broadphase;
int b3amd_broadphase_create(b3amd_context* ctx, b3amd_broadphase** out);
int b3amd_broadphase_destroy(b3amd_broadphase* bp);
/* CreateProxy (broad_phase.cpp:78-90) for each AABB in order: the fat AABB is the AABB
 * extended by b3_aabbExtension; the proxy is buffered as moved.  user_data[i] (null: i) is
 * what find_pairs reports for the proxy (the reference's userData, a b3Shape*). */
int b3amd_broadphase_create_proxies(b3amd_broadphase* bp, uint32_t n, const b3amd_aabb* aabbs,
                                    const uint32_t* user_data, uint32_t* proxy_ids);
/* DestroyProxy (broad_phase.cpp:92-97) for each id in order. */
int b3amd_broadphase_destroy_proxies(b3amd_broadphase* bp, uint32_t n, const uint32_t* proxy_ids);
/* TouchProxy (broad_phase.cpp:151-154). */
int b3amd_broadphase_touch_proxies(b3amd_broadphase* bp, uint32_t n, const uint32_t* proxy_ids);
/* MoveProxy (broad_phase.cpp:99-149) on device arrays, enqueued on the context stream:
 * d_moved[i] (may be null) = whether proxy d_proxy_ids[i] moved.  An id that is not live
 * moves nothing and makes the next find_pairs return B3AMD_ERR_ARGUMENT. */
int b3amd_broadphase_move_proxies(b3amd_broadphase* bp, uint32_t n, const uint32_t* d_proxy_ids,
                                  const b3amd_aabb* d_aabbs, const b3amd_vec3* d_displacements, uint32_t* d_moved);
/* FindPairs (broad_phase.h:160-208): *pair_count pairs at *d_pairs (device memory of the
 * broadphase, valid until the next call), 4 words each: proxy1 < proxy2 and their user
 * data, in the order AddPair receives them; the move buffer is cleared.  Synchronizes. */
int b3amd_broadphase_find_pairs(b3amd_broadphase* bp, uint32_t* pair_count, const uint32_t** d_pairs);
/* TestOverlap (broad_phase.cpp:73-76) of proxy pairs d_proxy_pairs[2i], [2i+1]: 1, 0, or
 * B3AMD_OVERLAP_INVALID for an id that is not live. */
int b3amd_broadphase_test_overlap(b3amd_broadphase* bp, uint32_t n, const uint32_t* d_proxy_pairs,
                                  uint32_t* d_overlap);
/* Live proxies and the current id capacity (ids are < id_capacity). */
int b3amd_broadphase_info(const b3amd_broadphase* bp, uint32_t* proxy_count, uint32_t* id_capacity);
/* GetAABB / GetUserData for ids first .. first+count-1 (host copies; flags bit 0 live, bit 1
 * moved; any pointer may be null). */
int b3amd_broadphase_read(b3amd_broadphase* bp, uint32_t first, uint32_t count, b3amd_aabb* aabbs,
                          uint32_t* user_data, uint32_t* flags);
/* The AABB b3Body::SynchronizeShapes moves each shape's proxy to (body.cpp:181-202): shape
 * d_shapes[i] of the context's table on body d_bodies[i] (< body_count); the shape's ComputeAABB
 * (sphere_shape.cpp:53-59, capsule_shape.cpp:130-137, hull_shape.cpp:162-166) at the body's
 * m_xf combined with the one at m_sweep.GetTransform(0) (transform.h:224-234), and the
 * displacement m_xf.position - that transform's position.  With d_body_sweeps null: the
 * AABB at m_xf alone (b3Body::CreateShape, body.cpp:110-115) and zero displacement.
 * d_displacements may be null.  Synchronizes (validates the shape entries). */
int b3amd_shape_aabbs(b3amd_context* ctx, uint32_t n, const uint32_t* d_shapes, const uint32_t* d_bodies,
                      uint32_t body_count, const b3amd_transform* d_body_xfs, const b3amd_sweep* d_body_sweeps,
                      b3amd_aabb* d_aabbs, b3amd_vec3* d_displacements);

/* ---- the step before: the contact manager ----
 * b3ContactManager (contact_manager.cpp) for convex contacts with the contact list, each
 * contact's b3ConvexCache and manifold in HBM, over a device broadphase.  Shapes are added
 * with their body (b3Body::CreateShape, body.cpp:89-121: the proxy's AABB is the shape's at
 * the body's m_xf, its user data the manager's shape index, returned in order from 0); body
 * flags are b3BodyType (0 static, 1 kinematic, 2 dynamic) | B3AMD_BODY_AWAKE.  Joints,
 * contact filters, sensors and mesh shapes are not modelled here (ShouldCollide is the
 * body-type test, body.cpp:204-211).  Destroy a manager before its broadphase. */
#define B3AMD_BODY_AWAKE 4u
typedef struct b3amd_contact_manager b3amd_contact_manager;
int b3amd_contact_manager_create(b3amd_context* ctx, b3amd_broadphase* bp, b3amd_contact_manager** out);
int b3amd_contact_manager_destroy(b3amd_contact_manager* m);
/* Replaces the body flags (host array of body_count words). */
int b3amd_contact_manager_set_bodies(b3amd_contact_manager* m, uint32_t body_count, const uint32_t* flags);
/* Current body flags (AddPair and touch changes set B3AMD_BODY_AWAKE, contact_manager.cpp:116-121,
 * contact.cpp:94-99). */
int b3amd_contact_manager_get_bodies(b3amd_contact_manager* m, uint32_t* flags);
/* b3Body::CreateShape for n shapes: context shape-table entry table_shapes[i] on body
 * bodies[i]; the AABB comes from d_body_xfs (device, one per body).  proxy_ids may be null. */
int b3amd_contact_manager_add_shapes(b3amd_contact_manager* m, uint32_t n, const uint32_t* table_shapes,
                                     const uint32_t* bodies, const b3amd_transform* d_body_xfs, uint32_t* proxy_ids);
/* b3Body::SynchronizeShapes (body.cpp:181-202) for the manager shapes d_shape_ids[0..n):
 * the swept AABB between m_sweep.GetTransform(0) and m_xf, then MoveProxy.  Enqueued; a bad
 * shape id makes the next find_new_contacts return B3AMD_ERR_ARGUMENT. */
int b3amd_contact_manager_synchronize(b3amd_contact_manager* m, uint32_t n, const uint32_t* d_shape_ids,
                                      const b3amd_transform* d_body_xfs, const b3amd_sweep* d_body_sweeps);
/* FindNewContacts (contact_manager.cpp:156-165): FindPairs, then AddPair for each pair in
 * order (:34-142); new contacts start with a new b3ConvexContact's cache
 * (convex_contact.cpp:31-36) and join the list front.  Synchronizes. */
int b3amd_contact_manager_find_new_contacts(b3amd_contact_manager* m, uint32_t* created);
/* UpdateContacts (contact_manager.cpp:167-231) on device body state (one transform and one
 * orientation pair (orientation0, orientation) per body, 16-byte aligned; sweeps may be
 * null): contacts whose bodies may not collide or whose fat AABBs stopped overlapping are
 * destroyed, contacts with no active body are skipped, the others run the narrowphase with
 * their cache (b3Contact::Update, contact.cpp:41-114).  stats (may be null) = destroyed,
 * collided, skipped.  The narrowphase is left enqueued on the context stream. */
int b3amd_contact_manager_update_contacts(b3amd_contact_manager* m, const b3amd_transform* d_body_xfs,
                                          const b3amd_quat* d_body_sweeps, uint32_t stats[3]);
/* The contact list in m_contactList order (host copies of up to cap entries; any output may
 * be null): shape pairs (manager shape indices, typeA <= typeB), born[i] = the number of
 * update_contacts calls before the contact was created | overlap flag << 31, manifolds and
 * caches. */
int b3amd_contact_manager_read(b3amd_contact_manager* m, uint32_t cap, uint32_t* count, uint32_t* shape_pairs,
                               uint32_t* born, b3amd_manifold* manifolds, b3amd_convex_cache* caches);
/* Device view: the list as slot indices, and the manifold and cache arrays by slot. */
int b3amd_contact_manager_view(b3amd_contact_manager* m, uint32_t* count, const uint32_t** d_order,
                               const b3amd_manifold** d_manifolds, const b3amd_convex_cache** d_caches);

/* ---- a whole step: b3World on the host around the device pipeline ----
 * b3World::Step (world.cpp:115-144): FindNewContacts for new shapes, UpdateContacts, and
 * Solve (:146-337: depth-first islands, b3Island::Solve with b3ContactSolver, island.cpp:109-341,
 * contact_solver.cpp:70-588, then SynchronizeShapes and FindNewContacts), bit for bit, with
 * the broadphase, contact manager and narrowphase on the device and the islands and the
 * sequential-impulse solver on the host.  World defaults as b3World's constructor
 * (world.cpp:33-50): gravity (0, -9.8, 0), warm starting on, sleeping off.  Joints, sensors
 * and mesh shapes are not stepped here. */
typedef struct b3amd_host_world b3amd_host_world;
typedef struct {                 /* b3BodyDef (body.h:47-100) */
    uint32_t type;               /* b3BodyType */
    uint32_t awake;
    uint32_t fixed_rotation;     /* bit 0 x, bit 1 y, bit 2 z */
    b3amd_vec3 position;
    b3amd_quat orientation;
    b3amd_vec3 linear_velocity, angular_velocity;
    float gravity_scale, linear_damping, angular_damping;
} b3amd_body_def;
typedef struct {                 /* b3ShapeDef (shape.h:48-66): the geometry is a shape-table entry */
    uint32_t table_shape;
    float density, friction, restitution;
    uint32_t sensor;             /* must be 0 */
} b3amd_shape_def;
typedef struct {                 /* b3Body state after a step (body.h) */
    b3amd_transform xf;          /* m_xf */
    b3amd_sweep sweep;           /* m_sweep */
    b3amd_vec3 linear_velocity, angular_velocity;
    float mass, inv_mass;
    float inertia[9], inv_inertia[9], world_inv_inertia[9];  /* m_I, m_invI, m_worldInvI (columns) */
    uint32_t awake;
    float sleep_time;
} b3amd_body_state;
int b3amd_host_world_create(b3amd_context* ctx, b3amd_host_world** out);
int b3amd_host_world_destroy(b3amd_host_world* w);
/* m_gravity (null: unchanged), m_warmStarting, m_sleeping. */
int b3amd_host_world_settings(b3amd_host_world* w, const b3amd_vec3* gravity, int warm_starting, int sleeping);
/* b3World::CreateBody (world.cpp:86-92); bodies are numbered from 0 in creation order. */
int b3amd_host_world_create_body(b3amd_host_world* w, const b3amd_body_def* def, uint32_t* body);
/* b3Body::CreateShape (body.cpp:89-121): ResetMass with the shape's b3*Shape::ComputeMass when
 * its density is not 0, then its proxy. */
int b3amd_host_world_create_shape(b3amd_host_world* w, uint32_t body, const b3amd_shape_def* def, uint32_t* shape);
/* b3World::Step(dt, velocity_iterations, position_iterations). */
int b3amd_host_world_step(b3amd_host_world* w, float dt, uint32_t velocity_iterations, uint32_t position_iterations);
int b3amd_host_world_bodies(b3amd_host_world* w, uint32_t first, uint32_t count, b3amd_body_state* out);
int b3amd_host_world_contact_count(b3amd_host_world* w, uint32_t* count);

/* b3amd_collide_device without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops
 * and harnesses that forbid hipMalloc).  The batch's scratch lives in a file-scope device
 * array of the library, one copy per GPU, about 11.5 MB, held until the process ends; n
 * above B3AMD_NOALLOC_MAX_PAIRS returns B3AMD_ERR_ARGUMENT.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (other contexts or streams) race.  The hull
 * pool and shape table must already be uploaded (those uploads allocate). */
#define B3AMD_NOALLOC_MAX_PAIRS 262144u
int b3amd_collide_device_noalloc(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs,
                                 const b3amd_transform* d_xfs, const b3amd_quat* d_sweeps,
                                 b3amd_convex_cache* d_caches, b3amd_manifold* d_out);

/* One world step of b3ContactManager::UpdateContacts (contact_manager.cpp:167-231) on
 * device-resident body state: for each contact i, b3ConvexContact::Collide
 * (convex_contact.cpp:50-62) with bodies A = d_contacts[2i], B = d_contacts[2i+1], i.e.
 * b3CollideShapeAndShape(d_out[i], xf[A], shape[A], xf[B], shape[B], &d_caches[i]) with
 * shape[b] = d_body_shapes[b] (an index into the shape table) and xf[b] = d_body_xfs[b]
 * (b3Body::m_xf).  d_body_sweeps (may be NULL) holds per body (orientation0,
 * orientation), the sweep the cached face rebuild reads.  The caller passes the contacts
 * UpdateContacts would collide: those it destroys or skips (filtered, no AABB overlap, both
 * bodies inactive, :186-222) are left out of the list.  Shape types must satisfy
 * typeA <= typeB; a body index >= body_count marks the contact B3AMD_FLAG_INVALID_PAIR.
 * Only the body transforms change per step, so only they need to cross PCIe.  Enqueued on
 * the context stream without synchronizing; transforms and sweeps 16-byte aligned. */
int b3amd_update_contacts(b3amd_context* ctx, uint32_t body_count, const uint32_t* d_body_shapes,
                          const b3amd_transform* d_body_xfs, const b3amd_quat* d_body_sweeps,
                          uint32_t contact_count, const uint32_t* d_contacts, b3amd_convex_cache* d_caches,
                          b3amd_manifold* d_out);

int b3amd_synchronize(b3amd_context* ctx);

/* ---- launch shape of the hull SAT (no reference counterpart: the reference runs the
 * queries of sat.cpp:33-173 one pair at a time on one thread) ----
 * The SAT kernel is one persistent grid over the batch's hull pairs.  waves_per_simd
 * (0 = 4, at most 32) is the number of resident waves per SIMD the grid is sized for; it
 * also picks the kernel's register budget (the variant built for min(max(w, 4), 6) waves).
 * Each wave takes a static share of the pair list, then the last sixteenth goes out from a
 * device counter tail_chunk pairs at a time (1..64, 0 = 8). */
int b3amd_set_sat_tuning(b3amd_context* ctx, uint32_t waves_per_simd, uint32_t tail_chunk);
int b3amd_get_sat_tuning(const b3amd_context* ctx, uint32_t* waves_per_simd, uint32_t* tail_chunk);
/* Times every launch shape on a device-resident batch (as b3amd_collide_device, without
 * feature caches or sweeps, so every hull pair runs the full SAT), `repeats` batches per
 * shape, keeps the fastest in the context and returns its mean batch time in best_ms
 * (may be NULL).  d_out is overwritten. */
int b3amd_autotune(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                   b3amd_manifold* d_out, uint32_t repeats, float* best_ms);

/* ---- captured batches (no reference counterpart: the reference calls
 * b3CollideShapeAndShape per contact, contact.cpp:41-140) ----
 * A HIP graph of one b3amd_collide_device batch on fixed device buffers, captured with the
 * current launch shape and its own scratch memory.  Each launch re-runs the batch on
 * whatever the buffers hold at that time, on the context stream, without synchronizing.
 * A graph refuses to launch after the hull pool or shape table has been re-uploaded.
 * Destroy graphs before their context (b3amd_destroy destroys any left). */
typedef struct b3amd_graph b3amd_graph;
int b3amd_graph_create(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                       const b3amd_quat* d_sweeps, b3amd_convex_cache* d_caches, b3amd_manifold* d_out,
                       b3amd_graph** out);
int b3amd_graph_launch(b3amd_graph* graph);
int b3amd_graph_destroy(b3amd_graph* graph);

/* ---- serving many worlds (no reference counterpart: the reference runs one b3World's
 * b3ContactManager::UpdateContacts, contact_manager.cpp:167-231, per call) ----
 * A server continuously batches the narrowphase steps of many worlds on one context.  A
 * world is a set of bodies (each an index into the context's shape table) and a fixed
 * contact list of body-index pairs (typeA <= typeB, in b3ContactManager list order); its
 * contact list, feature caches (created empty, as b3ConvexContact's constructor does,
 * convex_contact.cpp:35-36) and manifolds stay in HBM.  A request is one step: the body
 * transforms [bodyCount] (b3Body::m_xf) and optionally the sweeps [bodyCount][2]
 * (orientation0, orientation) read by the cached face rebuild.  A worker thread packs the
 * oldest pending requests, at most one per world, into batches of at most
 * max_batch_pairs contacts and max_batch_bodies bodies.  Up to three batches are in flight
 * at once, each on its own pipeline lane (two streams, device inputs and scratch), so
 * batches overlap on the device; a batch holding a world's step waits on the device for
 * the batch holding that world's previous step, so each world's steps run in order.
 * Admission: a world needs pair_count slots of the server's slab of slab_slots contacts
 * (created with the server and counted in the context's HBM budget).  b3amd_world_create
 * returns B3AMD_ERR_BUDGET at once when they do not fit; b3amd_world_admit waits, in
 * arrival order among admissions, until destroyed worlds free enough slots.  All entry
 * points are thread-safe with respect to each other; while a server exists it uses the
 * context's stream for world uploads and reads: call no other entry point on that context
 * concurrently. */
typedef struct b3amd_server b3amd_server;
typedef struct b3amd_server_stats {
    uint64_t batches;         /* device batches launched */
    uint64_t requests;        /* world steps packed into them */
    uint64_t pairs;           /* contacts computed */
    uint32_t max_batch_pairs; /* largest batch */
    uint32_t reserved;
    double gpu_ms;            /* device time summed over batches (copies, gather, kernels; batches
                               * on different lanes overlap, so this can exceed wall time) */
    double pack_ms;           /* host time spent packing batches */
} b3amd_server_stats;
int b3amd_server_create(b3amd_context* ctx, uint32_t max_batch_pairs, uint32_t max_batch_bodies,
                        uint32_t slab_slots, b3amd_server** out);
int b3amd_server_destroy(b3amd_server* server);
int b3amd_world_create(b3amd_server* server, uint32_t body_count, const uint32_t* body_shapes, uint32_t pair_count,
                       const uint32_t* body_pairs, uint32_t* world);
/* b3amd_world_create that waits up to timeout_ms for slots (FIFO among waiting admissions);
 * B3AMD_ERR_BUDGET when the time passes first. */
int b3amd_world_admit(b3amd_server* server, uint32_t body_count, const uint32_t* body_shapes, uint32_t pair_count,
                      const uint32_t* body_pairs, double timeout_ms, uint32_t* world);
/* Refuses further steps of the world, waits until its queued and in-flight steps are done,
 * then frees its slots and wakes waiting admissions. */
int b3amd_world_destroy(b3amd_server* server, uint32_t world);
/* Queues one step (inputs are copied before return); ticket identifies it. */
int b3amd_world_submit(b3amd_server* server, uint32_t world, const b3amd_transform* body_xfs,
                       const b3amd_quat* body_sweeps, uint64_t* ticket);
/* Blocks until the step's manifolds are in the world's device slots; latency_ms (may be
 * NULL) = time from submission to completion.  Each ticket is collected once. */
int b3amd_server_wait(b3amd_server* server, uint64_t ticket, double* latency_ms);
/* The world's manifolds / feature caches (b3ConvexCache per contact, the state each
 * b3ConvexContact::m_cache holds), in contact-list order, after every step submitted so far
 * has run (waits for them). */
int b3amd_world_read_manifolds(b3amd_server* server, uint32_t world, b3amd_manifold* out);
int b3amd_world_read_caches(b3amd_server* server, uint32_t world, b3amd_convex_cache* out);
/* Load generator: submits request i (world worlds[i], transforms body_xfs[i], which must
 * stay valid until the call returns) at times_s[i] seconds after the call starts, waits
 * for all, and writes each request's latency from its scheduled arrival. */
int b3amd_server_replay(b3amd_server* server, uint32_t count, const uint32_t* worlds, const double* times_s,
                        const b3amd_transform* const* body_xfs, double* latency_ms);
int b3amd_server_get_stats(b3amd_server* server, b3amd_server_stats* out);
/* Batch policy: at most batch_pairs contacts per batch (0 or more than the server's
 * max_batch_pairs: that maximum; a world larger than the cap still goes alone) and at most
 * max_inflight (1..3) batches in flight on the device, one per lane.  Defaults:
 * max_batch_pairs, 2. */
int b3amd_server_configure(b3amd_server* server, uint32_t batch_pairs, uint32_t max_inflight);

/* ---- statistics ----
 * Device equivalents of the reference's global counters: b3_convexCalls and
 * b3_convexCacheHits (collide_hulls.cpp:291-292,304; collide_hulls_cache.cpp:300,334),
 * b3_gjkCalls, b3_gjkIters, b3_gjkMaxIters and b3_gjkCacheHits (gjk.cpp:27,583,673,697,
 * 740,774), accumulated by every batch (convex pairs, mesh contacts and sensor overlap
 * tests) while counting is on.  Counting is off by default (it adds one atomic per wave and
 * counter).  b3World::Step resets the reference's counters every step
 * (world.cpp:122-127); call b3amd_reset_counters for the same.  Only the sensor path reads
 * simplex caches (b3TestOverlap); the narrowphase runs GJK without one (gjk.cpp:845-851),
 * as the reference's does. */
typedef struct b3amd_counters {
    uint64_t convex_calls;
    uint64_t convex_cache_hits;
    uint64_t gjk_calls;
    uint64_t gjk_iters;
    uint32_t gjk_max_iters;
    uint32_t gjk_cache_hits;
} b3amd_counters;
int b3amd_set_counting(b3amd_context* ctx, int on);
/* Synchronizes the context stream and reads the totals. */
int b3amd_get_counters(b3amd_context* ctx, b3amd_counters* out);
int b3amd_reset_counters(b3amd_context* ctx);

/* ---- triangle meshes ----
 * b3Mesh (mesh.h:41-99): vertices [V][3], triangles [T][3] (b3Triangle v1, v2, v3).  The
 * static tree is built on the host with the node numbering of b3Mesh::BuildTree /
 * b3StaticTree::Build (static_tree.cpp:42-180) and uploaded with the mesh. */
int b3amd_upload_mesh(b3amd_context* ctx, uint32_t vertexCount, const float* vertices, uint32_t triangleCount,
                      const uint32_t* triangles, uint32_t* meshId);
/* The same tree built on the host only (no context, no GPU): nodesOut holds 2T - 1 nodes. */
int b3amd_build_static_tree(uint32_t vertexCount, const float* vertices, uint32_t triangleCount,
                            const uint32_t* triangles, float* nodesOut);
/* Reads back the mesh's tree: nodeCount = 2T - 1 nodes of 8 words (lower xyz, upper xyz,
 * child1, child2 or triangle; child1 = 0xFFFFFFFF for a leaf), nodesOut may be NULL. */
int b3amd_mesh_tree(b3amd_context* ctx, uint32_t meshId, uint32_t* nodeCount, float* nodesOut);

/* Triangle lists hold B3AMD_MESH_TRIANGLE_CAP entries per contact.  A fresh contact
 * (b3amd_collide_mesh*) over more triangles is collided whole by a second pass (up to
 * B3AMD_MESH_WIDE_TRIANGLES triangles, as the reference's growing m_triangles): its
 * manifolds are the reference's and unflagged; its list returns the first
 * B3AMD_MESH_TRIANGLE_CAP triangles and its count all of them.  B3AMD_FLAG_TRIANGLE_OVERFLOW
 * marks the manifolds of a fresh contact beyond B3AMD_MESH_WIDE_TRIANGLES and of a
 * persistent contact (b3amd_mesh_contacts_*) beyond B3AMD_MESH_TRIANGLE_CAP, whose lists and
 * caches are cut there. */
#define B3AMD_MESH_TRIANGLE_CAP 64u
#define B3AMD_MESH_WIDE_TRIANGLES 2048u
#define B3AMD_FLAG_INVALID_PAIR 2u
#define B3AMD_FLAG_TRIANGLE_OVERFLOW 4u

/* n freshly created mesh contacts between convex shapes[shapeA[i]] (body transform
 * xfs[2i]) and mesh meshId (body transform xfs[2i+1]): what b3MeshContact does from
 * construction to Collide (mesh_contact.cpp:30-275) - fat AABB in the mesh frame, tree
 * query in DFS order, one narrowphase per triangle with an empty cache, clustering into at
 * most 3 manifolds (contact_cluster.cpp:475-590).  out[3i..3i+2] receive the manifolds
 * (manifoldCounts[i] of them); triangleCounts (may be NULL) the midphase hits and
 * triangleLists (may be NULL, [n][B3AMD_MESH_TRIANGLE_CAP]) the triangle indices. */
int b3amd_collide_mesh(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* shapeA,
                       const b3amd_transform* xfs, uint32_t* triangleCounts, uint32_t* triangleLists,
                       b3amd_manifold* out, uint32_t* manifoldCounts);
/* Same on device-resident buffers (b3amd_device_alloc), enqueued on the context stream
 * without synchronizing: the work-item count stays on the device, so a batch needs no
 * host round trip between midphase and narrowphase.  d_triangleCounts and d_triangleLists
 * may be NULL.  A shape index out of range or a non-convex shape A gives
 * B3AMD_FLAG_INVALID_PAIR in that contact's manifold flags and no manifolds. */
int b3amd_collide_mesh_device(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* d_shapeA,
                              const b3amd_transform* d_xfs, uint32_t* d_triangleCounts, uint32_t* d_triangleLists,
                              b3amd_manifold* d_out, uint32_t* d_manifoldCounts);

/* ---- persistent mesh contacts: b3MeshContact across steps (mesh_contact.cpp) ----
 * A set of n contacts between convex shapes[shapeA[i]] and mesh meshId that keeps, per
 * contact, the fat AABB of shape A in the mesh frame (m_aabbA), the triangles the tree
 * query found (at most B3AMD_MESH_TRIANGLE_CAP, as in b3amd_collide_mesh) and one
 * b3ConvexCache per triangle (b3TriangleCache, mesh_contact.h:30-34), on the device.
 *   create       the constructor and FindNewPairs (mesh_contact.cpp:30-55, 141-160),
 *                xfs[2i] = body A, xfs[2i+1] = mesh body;
 *   collide      b3MeshContact::Collide (:217-275) with every triangle's cache: per
 *                triangle b3CollideShapeAndShape, triangle keys, clustering into at most 3
 *                manifolds per contact in out[3i..]; sweeps (may be NULL: identity
 *                orientations) gives per contact body A's orientation0/orientation then
 *                the mesh body's (4 quats), read by the cached face rebuild;
 *   synchronize  SynchronizeShapes, MoveAABB and FindNewPairs (:62-160) for new transforms;
 *                centers0[3i..] is body A's sweep worldCenter0 (the displacement's origin).
 *                A contact whose shape still fits its fat AABB keeps its triangles and
 *                caches; otherwise the tree is queried again and the caches restart empty.
 *   read         triangle counts, lists [n][B3AMD_MESH_TRIANGLE_CAP], caches (same
 *                layout) and fat AABBs [n][6] (lower xyz, upper xyz); any may be NULL. */
typedef struct b3amd_mesh_contacts b3amd_mesh_contacts;
int b3amd_mesh_contacts_create(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* shapeA,
                               const b3amd_transform* xfs, b3amd_mesh_contacts** out);
int b3amd_mesh_contacts_destroy(b3amd_mesh_contacts* contacts);
int b3amd_mesh_contacts_collide(b3amd_mesh_contacts* contacts, const b3amd_transform* xfs, const b3amd_quat* sweeps,
                                b3amd_manifold* out, uint32_t* manifoldCounts);
/* b3MeshContact::TestOverlap (mesh_contact.cpp:191-215) for each contact whose shape pair
 * holds a sensor: b3TestOverlap of shape A against the listed triangles in list order, each
 * with its own simplex cache, stopping at the first overlap (the caches of later triangles
 * are left as they were).  overlap[i] = 1 / 0. */
int b3amd_mesh_contacts_test_overlap(b3amd_mesh_contacts* contacts, const b3amd_transform* xfs, uint32_t* overlap);
int b3amd_mesh_contacts_synchronize(b3amd_mesh_contacts* contacts, const b3amd_transform* xfs, const float* centers0);
int b3amd_mesh_contacts_read(b3amd_mesh_contacts* contacts, uint32_t* triangleCounts, uint32_t* triangleLists,
                             b3amd_convex_cache* caches, float* fatAabbs);
/* The same three steps on device-resident buffers (b3amd_device_alloc), enqueued on the
 * context stream without synchronizing, for a world whose body state stays on the device:
 * collide (mesh_contact.cpp:217-275; d_sweeps may be NULL), test_overlap (:191-215) and
 * synchronize (:62-160), with the layouts of the host forms above. */
int b3amd_mesh_contacts_collide_device(b3amd_mesh_contacts* contacts, const b3amd_transform* d_xfs,
                                       const b3amd_quat* d_sweeps, b3amd_manifold* d_out, uint32_t* d_manifoldCounts);
int b3amd_mesh_contacts_test_overlap_device(b3amd_mesh_contacts* contacts, const b3amd_transform* d_xfs,
                                            uint32_t* d_overlap);
int b3amd_mesh_contacts_synchronize_device(b3amd_mesh_contacts* contacts, const b3amd_transform* d_xfs,
                                           const float* d_centers0);

/* ---- device memory helpers (callers need no other GPU runtime) ---- */
int b3amd_device_alloc(b3amd_context* ctx, uint64_t bytes, void** ptr);
int b3amd_device_free(b3amd_context* ctx, void* ptr);

/* ---- contact caches ----
 * Device arrays of b3ConvexCache, one per contact, in the state a new contact's cache starts
 * in (b3ConvexContact's constructor, convex_contact.cpp:35-36, and each triangle cache of
 * b3MeshContact::FindNewPairs, mesh_contact.cpp:182-183: simplex count 0, feature cache
 * empty).  Reset re-empties caches [first, first + count) on the context stream with a fill
 * kernel (no memset node, so it may be captured), for contacts created in reused slots. */
int b3amd_cache_alloc(b3amd_context* ctx, uint32_t count, b3amd_convex_cache** d_caches);
int b3amd_cache_reset(b3amd_context* ctx, b3amd_convex_cache* d_caches, uint32_t first, uint32_t count);
int b3amd_cache_free(b3amd_context* ctx, b3amd_convex_cache* d_caches);
int b3amd_copy_to_device(b3amd_context* ctx, void* dst, const void* src, uint64_t bytes);
int b3amd_copy_to_host(b3amd_context* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy on the context stream (synchronous).  Also the plain copy of
 * known size that HBM traffic counters are calibrated against. */
int b3amd_copy_on_device(b3amd_context* ctx, void* dst, const void* src, uint64_t bytes);

/* ---- kernel timing (HIP events on the context stream) ----
 * While enabled, every kernel the context launches is bracketed by events; stats
 * accumulate per kernel name until reset.  Names: "sat_hull", "contact_hull", ... */
int b3amd_set_profiling(b3amd_context* ctx, int enabled);
int b3amd_kernel_stats(b3amd_context* ctx, const char* name, uint32_t* launches, double* total_ms);
int b3amd_reset_kernel_stats(b3amd_context* ctx);

/* ---- SAT edge-query filter data (host only; no reference counterpart) ----
 * For each unique edge k of hull `hull` (half-edges 2k, 2k+1), the bounding cap the SAT
 * kernel tests before the exact b3IsMinkowskiFace (sat.cpp:60-71): capsOut[5k..5k+4] =
 * centre m (3 floats), cos and sin of the widened radius, as the hull record stores them
 * at upload.  A whole-sphere cap is m = 0, cos 0, sin 1. */
int b3amd_hull_edge_caps(const b3amd_hull_pool* pool, uint32_t hull, float* capsOut);

/* ---- hull building and hull files (host only) ----
 * b3QHull::Set (qhull.cpp:104-354) on `count` points (b3Vec3, 12-byte stride) with the
 * reference's simplify flag: vertices, half-edges (origin, twin, face, next) and faces are
 * numbered exactly as the reference numbers them (qhull.cpp:282-347), planes and the
 * centroid (qhull.cpp:41-102) computed in its operation order.  counts = (V, HE, F), all 0
 * when the reference leaves the hull empty (fewer than 4 points after welding within
 * LINEAR_SLOP, or a degenerate point set).  Capacities: V <= 4 count + 16 (simplification
 * rebuilds the hull from its dual), HE <= 6 V, F <= 2 V.  A vertex quickhull keeps but no
 * face uses is written as the origin (the reference leaves that slot unwritten). */
int b3amd_quickhull(const float* points, uint32_t count, int simplify, uint32_t counts[3], float* vertices,
                    uint32_t* edges, uint32_t* faces, float* planes, float* centroid);
/* A hull pool as one binary structure-of-arrays file (header, the six pool arrays, FNV-1a
 * hash).  info reads the hull count and the vertex / half-edge / face totals, so the caller
 * can size the arrays read fills; read checks the header, the hash and every hull's ranges. */
int b3amd_hull_file_write(const char* path, const b3amd_hull_pool* pool);
int b3amd_hull_file_info(const char* path, uint32_t* count, uint64_t sizes[3]);
int b3amd_hull_file_read(const char* path, uint32_t* desc, float* centroids, float* vertices, uint32_t* edges,
                         uint32_t* faces, float* planes);

/* ---- workload generators (seeded, deterministic on every machine) ---- */

/* `count` hulls, each b3QHull::Set (no simplification) of N ~ U{nmin..nmax} points on an
 * ellipsoid with semi-axes ~ U[amin,amax] (redrawn if the hull would keep an unused vertex).
 * Output capacities: vertices >= count*nmax*3 floats, edges >= count*(6nmax-12)*4,
 * faces/planes >= count*(2nmax-4) (planes x4).  Writes totals to sizes[0..2]
 * (vertices, half-edges, faces). */
int b3amd_gen_hull_pool(uint64_t seed, uint32_t count, uint32_t nmin, uint32_t nmax,
                        float amin, float amax, uint32_t* desc, float* centroids,
                        float* vertices, uint32_t* edges, uint32_t* faces, float* planes,
                        uint64_t* sizes);

/* n hull-hull pairs over the pool: uniform random orientations; B is placed along a
 * random unit direction u so that the gap between the supports along u is
 * g ~ N(gap_mean, gap_sigma).  pairs[2i] < count, pairs[2i+1] < count (hull indices),
 * xfs[2i], xfs[2i+1] the transforms. */
int b3amd_gen_hull_pairs(uint64_t seed, const b3amd_hull_pool* pool, uint32_t n,
                         float gap_mean, float gap_sigma, uint32_t* pairs,
                         b3amd_transform* xfs);

/* For i < n: the distance t[i] along the unit direction dirs[3i..] at which hull
 * hb[i] (rotation columns rb[9i..]) has SAT signed separation gaps[i] from hull ha[i]
 * (rotation ra[9i..]) placed at the origin, both hull origins on the line. */
int b3amd_gen_place_along(const b3amd_hull_pool* pool, uint32_t n, const uint32_t* ha, const float* ra,
                          const uint32_t* hb, const float* rb, const double* dirs, const double* gaps,
                          double* t);

#ifdef __cplusplus
}
#endif
#endif
