// b3ContactManager for convex contacts on the device (contact_manager.cpp): the contact list
// that FindNewContacts grows from the broadphase's pairs (AddPair, :34-142) and
// UpdateContacts walks (:167-231), kept in HBM with each contact's b3ConvexCache and
// manifold, and b3Body::SynchronizeShapes (body.cpp:181-202) feeding the device broadphase.
//
// Layout.  A contact lives in a slot (record, cache, manifold) from its creation to its
// destruction; the list order of m_contactList is an array of slots.  AddPair pushes each
// new contact at the front in the order the pairs arrive, so after one FindNewContacts the
// list is the new contacts in reverse pair order followed by the old list; destroying
// contacts removes them in place.  Both are stable compactions over the list (one exclusive
// scan each).  The existence test of AddPair (a contact between the two shapes, :50-71)
// searches the live contacts' proxy-pair keys, kept sorted (the broadphase reports pairs
// sorted, so the new keys merge in).
//
// Each proxy carries the manager's shape index as user data, as the reference's carries the
// b3Shape*.  Bodies are flag words: b3BodyType (0 static, 1 kinematic, 2 dynamic) | awake << 2.
// Joints and contact filters are not modelled: ShouldCollide is the body-type test
// (body.cpp:204-211).
#include <cstring>
#include <hip/hip_runtime.h>
#include <rocprim/device/device_merge.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "broadphase.h"

namespace b3amd {

constexpr uint32_t kBodyType = 3u, kBodyAwake = 4u, kStatic = 0u, kDynamic = 2u;

struct CmShape {
    uint32_t table, body, proxy, type;
};
// One contact: shapes after Create's type swap (contact_manager.cpp:238-242), their bodies
// and proxies, the update count at creation and b3Contact's overlap flag (contact.cpp:94-114).
struct CmRec {
    uint32_t shapeA, shapeB, bodyA, bodyB, proxyA, proxyB, born, overlap;
};
static_assert(sizeof(CmRec) == 32, "contact record");

__device__ __forceinline__ bool has_key(const uint64_t* keys, uint32_t n, uint64_t k) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint64_t v = keys[mid];
        if (v == k) return true;
        if (v < k) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

// AddPair's filters (contact_manager.cpp:34-82) for every pair FindPairs reported.
__global__ __launch_bounds__(256) void cm_filter_kernel(uint32_t P, const uint4* pairs, const CmShape* shapes,
                                                        uint32_t shapeCount, const uint32_t* bodyFlags,
                                                        const uint64_t* keys, uint32_t L, uint32_t* keep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const uint4 p = pairs[i];
    uint32_t k = 0u;
    if (p.z < shapeCount && p.w < shapeCount) {
        const CmShape A = shapes[p.z], B = shapes[p.w];
        const uint64_t key = ((uint64_t)p.x << 32) | p.y;
        const uint32_t ta = bodyFlags[A.body] & kBodyType, tb = bodyFlags[B.body] & kBodyType;
        k = A.body != B.body && !has_key(keys, L, key) && (ta == kDynamic || tb == kDynamic) &&
                    A.type != B3AMD_MESH && B.type != B3AMD_MESH
                ? 1u
                : 0u;
    }
    keep[i] = k;
}

struct CmCreateArgs {
    uint32_t P, K;
    const uint4* pairs;
    const CmShape* shapes;
    const uint32_t* keep;
    const uint32_t* pos;
    uint32_t freeCount;
    const uint32_t* freeStack;
    uint32_t slotHigh;
    uint32_t born;
    CmRec* recs;
    ConvexCache* caches;
    Manifold* manifolds;
    uint32_t* order;    // new list: [0, K) the new contacts (reverse pair order), then the old list
    uint64_t* newKeys;  // [K] sorted
    uint32_t* newVals;
    uint32_t* bodyFlags;
};

// Create (contact_manager.cpp:233-268) and the list push (:84-130) for each kept pair.
__global__ __launch_bounds__(256) void cm_create_kernel(CmCreateArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.P || !a.keep[i]) return;
    const uint32_t j = a.pos[i];
    const uint32_t slot = j < a.freeCount ? a.freeStack[a.freeCount - 1u - j] : a.slotHigh + (j - a.freeCount);
    const uint4 p = a.pairs[i];
    uint32_t sa = p.z, sb = p.w;
    CmShape A = a.shapes[sa], B = a.shapes[sb];
    if (A.type > B.type) {
        const CmShape t = A;
        A = B, B = t;
        const uint32_t u = sa;
        sa = sb, sb = u;
    }
    a.recs[slot] = CmRec{sa, sb, A.body, B.body, A.proxy, B.proxy, a.born, 0u};  // m_flags = 0 (:88)
    ConvexCache c;  // b3ConvexContact's constructor (convex_contact.cpp:31-36)
    for (int t = 0; t < 5; ++t) c.simplex[t] = 0u;
    c.feature.state = kCacheEmpty;
    c.feature.type = c.feature.index1 = c.feature.index2 = 0u;
    a.caches[slot] = c;
    a.manifolds[slot].pointCount = 0u;
    a.manifolds[slot].flags = 0u;
    a.order[a.K - 1u - j] = slot;
    a.newKeys[j] = ((uint64_t)p.x << 32) | p.y;
    a.newVals[j] = slot;
    // Both bodies are woken (no sensor shapes here, :116-121).
    atomicOr(a.bodyFlags + A.body, kBodyAwake);
    atomicOr(a.bodyFlags + B.body, kBodyAwake);
}

__global__ __launch_bounds__(256) void cm_copy_kernel(uint32_t n, const uint32_t* src, uint32_t* dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

enum : uint32_t { kSkip = 0u, kCollide = 1u, kDestroy = 2u };

// UpdateContacts' decision per contact in list order (contact_manager.cpp:175-229).
__global__ __launch_bounds__(256) void cm_classify_kernel(uint32_t L, const uint32_t* order, const CmRec* recs,
                                                          const uint32_t* bodyFlags, BpState st, uint64_t* v,
                                                          uint32_t* dead) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L) return;
    const uint32_t slot = order[t];
    const CmRec r = recs[slot];
    const uint32_t fa = bodyFlags[r.bodyA], fb = bodyFlags[r.bodyB];
    const uint32_t ta = fa & kBodyType, tb = fb & kBodyType;
    uint32_t cls;
    if (ta != kDynamic && tb != kDynamic) {
        cls = kDestroy;
    } else {
        const bool activeA = (fa & kBodyAwake) && ta != kStatic, activeB = (fb & kBodyAwake) && tb != kStatic;
        if (!activeA && !activeB) cls = kSkip;
        else if (!overlap(st.lower[r.proxyA], st.upper[r.proxyA], st.lower[r.proxyB], st.upper[r.proxyB]))
            cls = kDestroy;
        else cls = kCollide;
    }
    v[t] = ((uint64_t)(cls != kDestroy ? 1u : 0u) << 32) | (cls == kCollide ? 1u : 0u);
    dead[slot] = cls == kDestroy ? 1u : 0u;
}

struct CmScatterArgs {
    uint32_t L;
    const uint32_t* order;
    const uint64_t* v;
    const uint64_t* pos;
    const CmRec* recs;
    const CmShape* shapes;
    uint32_t* newOrder;
    uint32_t* colSlots;
    uint32_t* colPairs;   // table shape pair
    uint32_t* colBodies;  // body pair
    uint32_t* freeStack;
    uint32_t freeCount;
};

__global__ __launch_bounds__(256) void cm_scatter_kernel(CmScatterArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.L) return;
    const uint32_t slot = a.order[t];
    const uint64_t v = a.v[t], p = a.pos[t];
    const uint32_t keepPos = (uint32_t)(p >> 32), colPos = (uint32_t)p;
    if (v >> 32) {
        a.newOrder[keepPos] = slot;
        if (v & 1u) {
            const CmRec r = a.recs[slot];
            a.colSlots[colPos] = slot;
            a.colPairs[2 * colPos] = a.shapes[r.shapeA].table;
            a.colPairs[2 * colPos + 1] = a.shapes[r.shapeB].table;
            a.colBodies[2 * colPos] = r.bodyA;
            a.colBodies[2 * colPos + 1] = r.bodyB;
        }
    } else {
        a.freeStack[a.freeCount + (t - keepPos)] = slot;
    }
}

// The end of b3Contact::Update (contact.cpp:94-114): the contact overlaps when its manifold
// has points; a change wakes both bodies.
__global__ __launch_bounds__(256) void cm_overlap_kernel(uint32_t n, const uint32_t* slots, const Manifold* manifolds,
                                                         CmRec* recs, uint32_t* bodyFlags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t slot = slots[i];
    const uint32_t now = manifolds[slot].pointCount > 0u ? 1u : 0u;
    CmRec& r = recs[slot];
    if (now != r.overlap) {
        atomicOr(bodyFlags + r.bodyA, kBodyAwake);
        atomicOr(bodyFlags + r.bodyB, kBodyAwake);
        r.overlap = now;
    }
}

__global__ __launch_bounds__(256) void cm_key_flags_kernel(uint32_t L, const uint32_t* vals, const uint32_t* dead,
                                                           uint32_t* live) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < L) live[k] = dead[vals[k]] ? 0u : 1u;
}

__global__ __launch_bounds__(256) void cm_key_scatter_kernel(uint32_t L, const uint64_t* keys, const uint32_t* vals,
                                                             const uint32_t* live, const uint32_t* pos,
                                                             uint64_t* outKeys, uint32_t* outVals) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= L || !live[k]) return;
    outKeys[pos[k]] = keys[k];
    outVals[pos[k]] = vals[k];
}

struct CmSyncArgs {
    uint32_t n;
    const uint32_t* shapeIds;
    const CmShape* shapes;
    uint32_t shapeCount, bodyCount;
    const ShapeRec* table;
    const float4* pool;
    const float* xfs;
    const float* sweeps;  // [B][18] b3Sweep
    BpState st;
    uint32_t* status;
};

// b3Body::SynchronizeShapes (body.cpp:181-202) for the listed shapes: the swept AABB and
// displacement, then MoveProxy.
__global__ __launch_bounds__(256) void cm_sync_kernel(CmSyncArgs a) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t s = a.shapeIds[i];
    if (s >= a.shapeCount) {
        atomicMax(a.status, i + 1u);
        return;
    }
    const CmShape sh = a.shapes[s];
    if (sh.body >= a.bodyCount) {
        atomicMax(a.status, i + 1u);
        return;
    }
    const ShapeRec rec = a.table[sh.table];
    const float* x = a.xfs + 12 * (size_t)sh.body;
    Xf xf2;
    xf2.R = M33{V3{x[0], x[1], x[2]}, V3{x[3], x[4], x[5]}, V3{x[6], x[7], x[8]}};
    xf2.p = V3{x[9], x[10], x[11]};
    const Xf xf1 = sweep_transform0(a.sweeps + 18 * (size_t)sh.body);
    V3 lo1, hi1, lo2, hi2;
    shape_aabb(rec, a.pool, xf1, lo1, hi1);
    shape_aabb(rec, a.pool, xf2, lo2, hi2);
    move_proxy(a.st, sh.proxy, vmin(lo1, lo2), vmax(hi1, hi2), xf2.p - xf1.p);
}

inline uint32_t cm_blocks(uint64_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace b3amd

struct b3amd_contact_manager {
    b3amd_context* ctx = nullptr;
    b3amd_broadphase* bp = nullptr;
    // bodies and shapes
    uint32_t bodyCount = 0, bodyCap = 0;
    uint32_t* dBodyFlags = nullptr;
    std::vector<CmShape> shapes;
    uint32_t shapeCap = 0;
    CmShape* dShapes = nullptr;
    // add_shapes scratch: table shapes, bodies, proxy records and AABBs of addCap shapes
    uint32_t addCap = 0;
    uint32_t* dAdd = nullptr;
    // contacts
    uint32_t L = 0, freeCount = 0, slotHigh = 0, cap = 0, updates = 0;
    CmRec* dRecs = nullptr;
    ConvexCache* dCaches = nullptr;
    Manifold* dManifolds = nullptr;
    uint32_t* dDead = nullptr;
    uint32_t* dOrder[2] = {nullptr, nullptr};
    int cur = 0;
    uint64_t* dKeys[2] = {nullptr, nullptr};
    uint32_t* dVals[2] = {nullptr, nullptr};
    uint32_t* dFree = nullptr;
    // per-call scratch, sized for cap
    uint64_t* dV = nullptr;
    uint64_t* dPos = nullptr;
    uint32_t* dColSlots = nullptr;
    uint32_t* dColPairs = nullptr;
    uint32_t* dColBodies = nullptr;
    // pair scratch, sized for pairCap
    uint32_t pairCap = 0;
    uint32_t* dKeep = nullptr;
    uint32_t* dKeepPos = nullptr;
    uint64_t* dNewKeys = nullptr;
    uint32_t* dNewVals = nullptr;
    uint32_t* dStatus = nullptr;
    void* dTmp = nullptr;
    size_t tmpBytes = 0;
};

namespace {

int cm_tmp(b3amd_contact_manager* m, size_t need) {
    if (need <= m->tmpBytes) return B3AMD_OK;
    m->ctx->release(m->dTmp);
    m->dTmp = nullptr;
    m->tmpBytes = 0;
    const size_t bytes = std::max<size_t>(need, 1 << 20);
    int rc = m->ctx->alloc(bytes, &m->dTmp);
    if (rc) return rc;
    m->tmpBytes = bytes;
    return B3AMD_OK;
}

template <class T>
int cm_regrow(b3amd_context* ctx, T** p, uint64_t oldCount, uint64_t newCount, bool keep) {
    T* q = nullptr;
    int rc = ctx->alloc(sizeof(T) * newCount, (void**)&q);
    if (rc) return rc;
    if (keep && *p && oldCount)
        HIP_TRY(hipMemcpyAsync(q, *p, sizeof(T) * oldCount, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->release(*p);
    *p = q;
    return B3AMD_OK;
}

// Slot and list capacity for `need` contacts.
int cm_reserve(b3amd_contact_manager* m, uint32_t need) {
    if (need <= m->cap) return B3AMD_OK;
    uint32_t cap = std::max<uint32_t>(m->cap ? m->cap : 4096u, 4096u);
    while (cap < need) cap *= 2u;
    b3amd_context* ctx = m->ctx;
    const uint32_t old = m->cap;
    int rc;
    if ((rc = cm_regrow(ctx, &m->dRecs, old, cap, true)) || (rc = cm_regrow(ctx, &m->dCaches, old, cap, true)) ||
        (rc = cm_regrow(ctx, &m->dManifolds, old, cap, true)) || (rc = cm_regrow(ctx, &m->dDead, old, cap, false)) ||
        (rc = cm_regrow(ctx, &m->dOrder[m->cur], old, cap, true)) ||
        (rc = cm_regrow(ctx, &m->dOrder[1 - m->cur], old, cap, false)) ||
        (rc = cm_regrow(ctx, &m->dKeys[m->cur], old, cap, true)) ||
        (rc = cm_regrow(ctx, &m->dKeys[1 - m->cur], old, cap, false)) ||
        (rc = cm_regrow(ctx, &m->dVals[m->cur], old, cap, true)) ||
        (rc = cm_regrow(ctx, &m->dVals[1 - m->cur], old, cap, false)) ||
        (rc = cm_regrow(ctx, &m->dFree, old, cap, true)) || (rc = cm_regrow(ctx, &m->dV, old, cap, false)) ||
        (rc = cm_regrow(ctx, &m->dPos, old, cap, false)) || (rc = cm_regrow(ctx, &m->dColSlots, old, cap, false)) ||
        (rc = cm_regrow(ctx, &m->dColPairs, 2ull * old, 2ull * cap, false)) ||
        (rc = cm_regrow(ctx, &m->dColBodies, 2ull * old, 2ull * cap, false)))
        return rc;
    m->cap = cap;
    return B3AMD_OK;
}

int cm_reserve_pairs(b3amd_contact_manager* m, uint32_t need) {
    if (need <= m->pairCap) return B3AMD_OK;
    uint32_t cap = std::max<uint32_t>(m->pairCap ? m->pairCap : 4096u, 4096u);
    while (cap < need) cap *= 2u;
    int rc;
    if ((rc = cm_regrow(m->ctx, &m->dKeep, 0, cap, false)) || (rc = cm_regrow(m->ctx, &m->dKeepPos, 0, cap, false)) ||
        (rc = cm_regrow(m->ctx, &m->dNewKeys, 0, cap, false)) || (rc = cm_regrow(m->ctx, &m->dNewVals, 0, cap, false)))
        return rc;
    m->pairCap = cap;
    return B3AMD_OK;
}

// add_shapes scratch for `need` shapes (12 words each).
int cm_reserve_add(b3amd_contact_manager* m, uint32_t need) {
    if (need <= m->addCap) return B3AMD_OK;
    uint32_t cap = std::max<uint32_t>(m->addCap ? m->addCap : 4096u, 4096u);
    while (cap < need) cap *= 2u;
    int rc = cm_regrow(m->ctx, &m->dAdd, 0, 12ull * cap, false);
    if (rc) return rc;
    m->addCap = cap;
    return B3AMD_OK;
}

template <class T>
int read1(b3amd_contact_manager* m, const T* d, T* h) {
    return m->ctx->read_back(d, h, sizeof(T));
}

// The last element of an exclusive scan and of its input (their sum is the total), read
// back together through the context's pinned words.
template <class T>
int read_scan_total(b3amd_contact_manager* m, const T* pos, const T* in, uint32_t n, T* total) {
    b3amd_context* ctx = m->ctx;
    T* h = reinterpret_cast<T*>(ctx->hWords);
    HIP_TRY(hipMemcpyAsync(h, pos + n - 1, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipMemcpyAsync(h + 1, in + n - 1, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *total = h[0] + h[1];
    return B3AMD_OK;
}

}  // namespace

extern "C" {

int b3amd_contact_manager_create(b3amd_context* ctx, b3amd_broadphase* bp, b3amd_contact_manager** out) {
    if (!ctx || !bp || !out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (bp->ctx != ctx) return fail(B3AMD_ERR_ARGUMENT, "the broadphase belongs to another context");
    b3amd_contact_manager* m = new b3amd_contact_manager();
    m->ctx = ctx;
    m->bp = bp;
    int rc;
    if ((rc = ctx->alloc(64, (void**)&m->dStatus)) || (rc = cm_reserve(m, 4096u)) || (rc = cm_reserve_pairs(m, 4096u)) ||
        (rc = cm_reserve_add(m, 4096u))) {
        b3amd_contact_manager_destroy(m);
        return rc;
    }
    HIP_TRY(hipMemsetAsync(m->dStatus, 0, 64, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    *out = m;
    return B3AMD_OK;
}

int b3amd_contact_manager_destroy(b3amd_contact_manager* m) {
    if (!m) return B3AMD_OK;
    b3amd_context* ctx = m->ctx;
    hipStreamSynchronize(ctx->stream);
    void* ptrs[] = {m->dBodyFlags, m->dShapes, m->dRecs, m->dCaches, m->dManifolds, m->dDead, m->dOrder[0],
                    m->dOrder[1], m->dKeys[0], m->dKeys[1], m->dVals[0], m->dVals[1], m->dFree, m->dV, m->dPos,
                    m->dColSlots, m->dColPairs, m->dColBodies, m->dKeep, m->dKeepPos, m->dNewKeys, m->dNewVals,
                    m->dStatus, m->dTmp, m->dAdd};
    for (void* p : ptrs) ctx->release(p);
    delete m;
    return B3AMD_OK;
}

int b3amd_contact_manager_set_bodies(b3amd_contact_manager* m, uint32_t body_count, const uint32_t* flags) {
    if (!m || (body_count && !flags)) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    for (const CmShape& s : m->shapes)
        if (s.body >= body_count) return fail(B3AMD_ERR_ARGUMENT, "a shape's body %u is beyond %u bodies", s.body, body_count);
    for (uint32_t b = 0; b < body_count; ++b)
        if ((flags[b] & kBodyType) > kDynamic || flags[b] > (kBodyType | kBodyAwake))
            return fail(B3AMD_ERR_ARGUMENT, "body %u: flags %u", b, flags[b]);
    int rc;
    if (body_count > m->bodyCap) {
        uint32_t cap = std::max<uint32_t>(1024u, m->bodyCap);
        while (cap < body_count) cap *= 2u;
        if ((rc = cm_regrow(m->ctx, &m->dBodyFlags, 0, cap, false))) return rc;
        m->bodyCap = cap;
    }
    if (body_count)
        HIP_TRY(hipMemcpyAsync(m->dBodyFlags, flags, 4ull * body_count, hipMemcpyHostToDevice, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    m->bodyCount = body_count;
    return B3AMD_OK;
}

int b3amd_contact_manager_get_bodies(b3amd_contact_manager* m, uint32_t* flags) {
    if (!m || !flags) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (m->bodyCount)
        HIP_TRY(hipMemcpyAsync(flags, m->dBodyFlags, 4ull * m->bodyCount, hipMemcpyDeviceToHost, m->ctx->stream));
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    return B3AMD_OK;
}

int b3amd_contact_manager_add_shapes(b3amd_contact_manager* m, uint32_t n, const uint32_t* table_shapes,
                                     const uint32_t* bodies, const b3amd_transform* d_body_xfs, uint32_t* proxy_ids) {
    if (!m) return fail(B3AMD_ERR_ARGUMENT, "null manager");
    if (n == 0) return B3AMD_OK;
    if (!table_shapes || !bodies || !d_body_xfs) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = m->ctx;
    for (uint32_t i = 0; i < n; ++i) {
        if (table_shapes[i] >= ctx->shapeCount) return fail(B3AMD_ERR_ARGUMENT, "entry %u: shape %u beyond the table", i, table_shapes[i]);
        if (ctx->shapes[table_shapes[i]].type == B3AMD_MESH)
            return fail(B3AMD_ERR_UNSUPPORTED, "entry %u: mesh shapes are not managed here", i);
        if (bodies[i] >= m->bodyCount) return fail(B3AMD_ERR_ARGUMENT, "entry %u: body %u beyond %u", i, bodies[i], m->bodyCount);
    }
    // b3Body::CreateShape: the shape's AABB at the body's m_xf, then CreateProxy with the
    // shape as user data (body.cpp:110-115); the AABBs stay on the device.
    int rc;
    if ((rc = cm_reserve_add(m, n))) return rc;
    uint32_t* dS = m->dAdd;                    // [n] table shapes
    uint32_t* dB = dS + m->addCap;             // [n] bodies
    uint32_t* dR = dB + m->addCap;             // [n][2] proxy records
    float* dA = reinterpret_cast<float*>(dR + 2ull * m->addCap);  // [n][6] AABBs
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(dS, table_shapes, 4ull * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dB, bodies, 4ull * n, hipMemcpyHostToDevice, s));
    if ((rc = launch_shape_aabbs(ctx, n, dS, dB, m->bodyCount, d_body_xfs, dA))) return rc;
    const uint32_t first = (uint32_t)m->shapes.size();
    std::vector<uint32_t> users(n), ids(n), records(2ull * n);  // kept until the synchronisation below
    for (uint32_t i = 0; i < n; ++i) users[i] = first + i;
    if ((rc = bp_create_proxies_device(m->bp, n, dA, users.data(), ids.data(), records.data(), dR))) {
        hipStreamSynchronize(s);
        return rc;
    }
    for (uint32_t i = 0; i < n; ++i) {
        m->shapes.push_back(CmShape{table_shapes[i], bodies[i], ids[i], ctx->shapes[table_shapes[i]].type});
        if (proxy_ids) proxy_ids[i] = ids[i];
    }
    if (m->shapes.size() > m->shapeCap) {
        uint32_t cap = std::max<uint32_t>(1024u, m->shapeCap);
        while (cap < m->shapes.size()) cap *= 2u;
        if ((rc = cm_regrow(ctx, &m->dShapes, 0, cap, false))) return rc;
        m->shapeCap = cap;
        HIP_TRY(hipMemcpyAsync(m->dShapes, m->shapes.data(), sizeof(CmShape) * m->shapes.size(), hipMemcpyHostToDevice, s));
    } else {
        HIP_TRY(hipMemcpyAsync(m->dShapes + first, m->shapes.data() + first, sizeof(CmShape) * n, hipMemcpyHostToDevice, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return B3AMD_OK;
}

int b3amd_contact_manager_synchronize(b3amd_contact_manager* m, uint32_t n, const uint32_t* d_shape_ids,
                                      const b3amd_transform* d_body_xfs, const b3amd_sweep* d_body_sweeps) {
    if (!m) return fail(B3AMD_ERR_ARGUMENT, "null manager");
    if (n == 0) return B3AMD_OK;
    if (!d_shape_ids || !d_body_xfs || !d_body_sweeps) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    CmSyncArgs a;
    a.n = n;
    a.shapeIds = d_shape_ids;
    a.shapes = m->dShapes;
    a.shapeCount = (uint32_t)m->shapes.size();
    a.bodyCount = m->bodyCount;
    a.table = ctx->dShapes;
    a.pool = ctx->dPool;
    a.xfs = reinterpret_cast<const float*>(d_body_xfs);
    a.sweeps = reinterpret_cast<const float*>(d_body_sweeps);
    a.st = m->bp->st;
    a.status = m->dStatus;
    // A bad shape id is counted in the status word and reported by the next find_new_contacts.
    return timed(ctx, s, "cm_sync", [&] {
        cm_sync_kernel<<<cm_blocks(n), 256, 0, s>>>(a);
        return hipGetLastError();
    });
}

int b3amd_contact_manager_find_new_contacts(b3amd_contact_manager* m, uint32_t* created) {
    if (!m) return fail(B3AMD_ERR_ARGUMENT, "null manager");
    if (created) *created = 0;
    b3amd_context* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    int rc;
    uint32_t status = 0;
    if ((rc = read1(m, m->dStatus, &status))) return rc;
    if (status) {
        HIP_TRY(hipMemsetAsync(m->dStatus, 0, 4, s));
        HIP_TRY(hipStreamSynchronize(s));
        return fail(B3AMD_ERR_ARGUMENT, "synchronize entry %u named a shape or body that does not exist", status - 1u);
    }
    uint32_t P = 0;
    const uint32_t* dPairs = nullptr;
    if ((rc = bp_find_pairs(m->bp, &P, &dPairs, false))) return rc;
    if (P == 0) {
        HIP_TRY(hipStreamSynchronize(s));  // the broadphase's move-buffer reset
        return B3AMD_OK;
    }
    if ((rc = cm_reserve_pairs(m, P))) return rc;
    const uint4* pairs = reinterpret_cast<const uint4*>(dPairs);
    rc = timed(ctx, s, "cm_filter", [&] {
        cm_filter_kernel<<<cm_blocks(P), 256, 0, s>>>(P, pairs, m->dShapes, (uint32_t)m->shapes.size(), m->dBodyFlags,
                                                      m->dKeys[m->cur], m->L, m->dKeep);
        return hipGetLastError();
    });
    if (rc) return rc;
    size_t need = 0;
    if (rocprim::exclusive_scan(nullptr, need, m->dKeep, m->dKeepPos, 0u, P, rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "scan sizing failed");
    if ((rc = cm_tmp(m, need))) return rc;
    size_t have = m->tmpBytes;
    if (rocprim::exclusive_scan(m->dTmp, have, m->dKeep, m->dKeepPos, 0u, P, rocprim::plus<uint32_t>(), s) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "scan failed");
    uint32_t K = 0;
    if ((rc = read_scan_total(m, m->dKeepPos, m->dKeep, P, &K))) return rc;
    if (K == 0) return B3AMD_OK;
    const uint32_t fresh = K > m->freeCount ? K - m->freeCount : 0u;
    if ((rc = cm_reserve(m, std::max(m->slotHigh + fresh, m->L + K)))) return rc;
    const int nxt = 1 - m->cur;
    CmCreateArgs a;
    a.P = P;
    a.K = K;
    a.pairs = pairs;
    a.shapes = m->dShapes;
    a.keep = m->dKeep;
    a.pos = m->dKeepPos;
    a.freeCount = m->freeCount;
    a.freeStack = m->dFree;
    a.slotHigh = m->slotHigh;
    a.born = m->updates;
    a.recs = m->dRecs;
    a.caches = m->dCaches;
    a.manifolds = m->dManifolds;
    a.order = m->dOrder[nxt];
    a.newKeys = m->dNewKeys;
    a.newVals = m->dNewVals;
    a.bodyFlags = m->dBodyFlags;
    rc = timed(ctx, s, "cm_create", [&] {
        cm_create_kernel<<<cm_blocks(P), 256, 0, s>>>(a);
        if (m->L) cm_copy_kernel<<<cm_blocks(m->L), 256, 0, s>>>(m->L, m->dOrder[m->cur], m->dOrder[nxt] + K);
        return hipGetLastError();
    });
    if (rc) return rc;
    // The new keys (sorted: FindPairs' order) merge into the live keys.
    if (m->L == 0) {
        HIP_TRY(hipMemcpyAsync(m->dKeys[nxt], m->dNewKeys, 8ull * K, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->dVals[nxt], m->dNewVals, 4ull * K, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        m->cur = nxt;
        m->L += K;
        m->slotHigh += fresh;
        m->freeCount -= K - fresh;
        if (created) *created = K;
        return B3AMD_OK;
    }
    need = 0;
    if (rocprim::merge(nullptr, need, m->dKeys[m->cur], m->dNewKeys, m->dKeys[nxt], m->dVals[m->cur], m->dNewVals,
                       m->dVals[nxt], m->L, K, rocprim::less<uint64_t>(), s) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "merge sizing failed");
    if ((rc = cm_tmp(m, need))) return rc;
    have = m->tmpBytes;
    if (rocprim::merge(m->dTmp, have, m->dKeys[m->cur], m->dNewKeys, m->dKeys[nxt], m->dVals[m->cur], m->dNewVals,
                       m->dVals[nxt], m->L, K, rocprim::less<uint64_t>(), s) != hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "merge failed");
    HIP_TRY(hipStreamSynchronize(s));
    m->cur = nxt;
    m->L += K;
    m->slotHigh += fresh;
    m->freeCount -= K - fresh;
    if (created) *created = K;
    return B3AMD_OK;
}

int b3amd_contact_manager_update_contacts(b3amd_contact_manager* m, const b3amd_transform* d_body_xfs,
                                          const b3amd_quat* d_body_sweeps, uint32_t stats[3]) {
    if (!m) return fail(B3AMD_ERR_ARGUMENT, "null manager");
    if (stats) stats[0] = stats[1] = stats[2] = 0u;
    b3amd_context* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    const uint32_t L = m->L;
    m->updates++;
    if (L == 0) return B3AMD_OK;
    if (!d_body_xfs) return fail(B3AMD_ERR_ARGUMENT, "null body transforms");
    if ((reinterpret_cast<uintptr_t>(d_body_xfs) | reinterpret_cast<uintptr_t>(d_body_sweeps)) & 15u)
        return fail(B3AMD_ERR_ARGUMENT, "body transforms and sweeps must be 16-byte aligned");
    int rc = timed(ctx, s, "cm_classify", [&] {
        cm_classify_kernel<<<cm_blocks(L), 256, 0, s>>>(L, m->dOrder[m->cur], m->dRecs, m->dBodyFlags, m->bp->st,
                                                        m->dV, m->dDead);
        return hipGetLastError();
    });
    if (rc) return rc;
    size_t need = 0;
    if (rocprim::exclusive_scan(nullptr, need, m->dV, m->dPos, (uint64_t)0, L, rocprim::plus<uint64_t>(), s) !=
        hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "scan sizing failed");
    if ((rc = cm_tmp(m, need))) return rc;
    size_t have = m->tmpBytes;
    if (rocprim::exclusive_scan(m->dTmp, have, m->dV, m->dPos, (uint64_t)0, L, rocprim::plus<uint64_t>(), s) !=
        hipSuccess)
        return fail(B3AMD_ERR_DEVICE, "scan failed");
    uint64_t tot = 0;
    if ((rc = read_scan_total(m, m->dPos, m->dV, L, &tot))) return rc;
    const uint32_t kept = (uint32_t)(tot >> 32), collided = (uint32_t)tot, destroyed = L - kept;
    const int nxt = 1 - m->cur;
    CmScatterArgs a;
    a.L = L;
    a.order = m->dOrder[m->cur];
    a.v = m->dV;
    a.pos = m->dPos;
    a.recs = m->dRecs;
    a.shapes = m->dShapes;
    a.newOrder = m->dOrder[nxt];
    a.colSlots = m->dColSlots;
    a.colPairs = m->dColPairs;
    a.colBodies = m->dColBodies;
    a.freeStack = m->dFree;
    a.freeCount = m->freeCount;
    rc = timed(ctx, s, "cm_scatter", [&] {
        cm_scatter_kernel<<<cm_blocks(L), 256, 0, s>>>(a);
        return hipGetLastError();
    });
    if (rc) return rc;
    if (destroyed) {
        // Drop the destroyed contacts' keys (a stable compaction of the sorted keys).
        uint32_t* live = reinterpret_cast<uint32_t*>(m->dV);            // dV and dPos are free again
        uint32_t* pos = reinterpret_cast<uint32_t*>(m->dPos);
        cm_key_flags_kernel<<<cm_blocks(L), 256, 0, s>>>(L, m->dVals[m->cur], m->dDead, live);
        HIP_TRY(hipGetLastError());
        need = 0;
        if (rocprim::exclusive_scan(nullptr, need, live, pos, 0u, L, rocprim::plus<uint32_t>(), s) != hipSuccess)
            return fail(B3AMD_ERR_DEVICE, "scan sizing failed");
        if ((rc = cm_tmp(m, need))) return rc;
        have = m->tmpBytes;
        if (rocprim::exclusive_scan(m->dTmp, have, live, pos, 0u, L, rocprim::plus<uint32_t>(), s) != hipSuccess)
            return fail(B3AMD_ERR_DEVICE, "scan failed");
        cm_key_scatter_kernel<<<cm_blocks(L), 256, 0, s>>>(L, m->dKeys[m->cur], m->dVals[m->cur], live, pos,
                                                           m->dKeys[nxt], m->dVals[nxt]);
        HIP_TRY(hipGetLastError());
    } else {
        HIP_TRY(hipMemcpyAsync(m->dKeys[nxt], m->dKeys[m->cur], 8ull * L, hipMemcpyDeviceToDevice, s));
        HIP_TRY(hipMemcpyAsync(m->dVals[nxt], m->dVals[m->cur], 4ull * L, hipMemcpyDeviceToDevice, s));
    }
    m->cur = nxt;
    m->L = kept;
    m->freeCount += destroyed;
    if (stats) stats[0] = destroyed, stats[1] = collided, stats[2] = kept - collided;
    if (collided == 0) return B3AMD_OK;
    // b3Contact::Update -> b3ConvexContact::Collide for the persisting active contacts
    // (contact.cpp:63-78, convex_contact.cpp:50-62), caches and manifolds in their slots.
    if (!ctx->dPool || !ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no hull pool or shape table uploaded");
    if ((rc = ensure_scratch(ctx, collided))) return rc;
    const BodyRefs bodies{m->dColBodies, reinterpret_cast<const float*>(d_body_xfs),
                          reinterpret_cast<const float*>(d_body_sweeps)};
    if ((rc = enqueue_batch(ctx, collided, m->dColPairs, nullptr, nullptr, m->dCaches, m->dManifolds,
                            BatchScratch{ctx->dStage, ctx->dSat, ctx->dCounter}, m->dColSlots, &bodies)))
        return rc;
    return timed(ctx, s, "cm_overlap", [&] {
        cm_overlap_kernel<<<cm_blocks(collided), 256, 0, s>>>(collided, m->dColSlots, m->dManifolds, m->dRecs,
                                                              m->dBodyFlags);
        return hipGetLastError();
    });
}

int b3amd_contact_manager_read(b3amd_contact_manager* m, uint32_t cap, uint32_t* count, uint32_t* shape_pairs,
                               uint32_t* born, b3amd_manifold* manifolds, b3amd_convex_cache* caches) {
    if (!m || !count) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = m->ctx;
    hipStream_t s = ctx->stream;
    *count = m->L;
    const uint32_t n = std::min(cap, m->L);
    if (n == 0) return B3AMD_OK;
    std::vector<uint32_t> order(n);
    std::vector<CmRec> recs(m->slotHigh);
    HIP_TRY(hipMemcpyAsync(order.data(), m->dOrder[m->cur], 4ull * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(recs.data(), m->dRecs, sizeof(CmRec) * m->slotHigh, hipMemcpyDeviceToHost, s));
    std::vector<Manifold> man(manifolds ? m->slotHigh : 0);
    std::vector<ConvexCache> cac(caches ? m->slotHigh : 0);
    if (manifolds) HIP_TRY(hipMemcpyAsync(man.data(), m->dManifolds, sizeof(Manifold) * m->slotHigh, hipMemcpyDeviceToHost, s));
    if (caches) HIP_TRY(hipMemcpyAsync(cac.data(), m->dCaches, sizeof(ConvexCache) * m->slotHigh, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t slot = order[i];
        if (slot >= m->slotHigh) return fail(B3AMD_ERR_DEVICE, "contact list holds slot %u beyond %u (internal)", slot, m->slotHigh);
        const CmRec& r = recs[slot];
        if (shape_pairs) shape_pairs[2 * i] = r.shapeA, shape_pairs[2 * i + 1] = r.shapeB;
        if (born) born[i] = r.born | (r.overlap << 31);
        if (manifolds) std::memcpy(manifolds + i, &man[slot], sizeof(Manifold));
        if (caches) std::memcpy(caches + i, &cac[slot], sizeof(ConvexCache));
    }
    return B3AMD_OK;
}

int b3amd_contact_manager_view(b3amd_contact_manager* m, uint32_t* count, const uint32_t** d_order,
                               const b3amd_manifold** d_manifolds, const b3amd_convex_cache** d_caches) {
    if (!m || !count) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    *count = m->L;
    if (d_order) *d_order = m->dOrder[m->cur];
    if (d_manifolds) *d_manifolds = reinterpret_cast<const b3amd_manifold*>(m->dManifolds);
    if (d_caches) *d_caches = reinterpret_cast<const b3amd_convex_cache*>(m->dCaches);
    return B3AMD_OK;
}

}  // extern "C"
