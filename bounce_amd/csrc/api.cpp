// C ABI of the narrowphase (include/bounce_amd.h): context, HBM-budgeted device memory,
// hull/shape registry and the batch entry points that replace the per-contact
// b3CollideShapeAndShape calls of b3ContactManager::UpdateContacts
// (contact_manager.cpp:167-231).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <string>
#include <unordered_map>
#include <mutex>
#include <vector>

#include "bounce_amd.h"
#include "context.h"
#include "device_hull.h"
#include "kernels.h"
#include "static_tree.h"

using namespace b3amd;

namespace b3amd {

namespace {
thread_local std::string g_error;
}

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

}  // namespace b3amd

// Scratch of b3amd_collide_device_noalloc: a file-scope device array, one copy per GPU,
// holding what ensure_scratch allocates for a batch of up to kNoAllocMaxPairs pairs: the
// stage / SAT-list / rebuild-list words, the SAT results and the counter words (zeroed by a
// kernel at the start of every batch).  Host code never names the array: its address on
// each device comes from hipGetSymbolAddress, taken once per device when the first context
// on that device is created.
static_assert(B3AMD_MESH_TRIANGLE_CAP == kMeshTriangleCap && B3AMD_MESH_WIDE_TRIANGLES == kWideTris,
              "mesh list capacities of the public header");
constexpr uint32_t kNoAllocMaxPairs = 262144u;
constexpr uint64_t kNoAllocStageWords =
    2ull * kNoAllocMaxPairs +
    (uint64_t)kRebuildLists * (((kNoAllocMaxPairs + kReadPairsPerBlock - 1u) / kReadPairsPerBlock + kRebuildLists - 1u) /
                               kRebuildLists * kReadPairsPerBlock);
static_assert(sizeof(SatResult) % 16 == 0 && kNoAllocStageWords % 64 == 0, "sections stay 256-byte aligned");
constexpr uint64_t kNoAllocSatWords = sizeof(SatResult) / 4 * (uint64_t)kNoAllocMaxPairs;
constexpr uint64_t kNoAllocWords = kNoAllocStageWords + kNoAllocSatWords + kCountWords;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
void* g_scratchPtr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current).
int scratch_address(int device, void** out) {
    if (device < 0 || device >= kMaxDevices) return fail(B3AMD_ERR_ARGUMENT, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(B3AMD_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratchPtr[device] = p;
    }
    *out = g_scratchPtr[device];
    return B3AMD_OK;
}
}  // namespace

int b3amd::ensure_scratch(b3amd_context* ctx, uint32_t n) {
    if (n <= ctx->scratchCap) return B3AMD_OK;
    uint32_t cap = n < 1024 ? 1024 : n;
    int rc;
    if ((rc = ctx->release(ctx->dStage)) || (rc = ctx->release(ctx->dSat)) || (rc = ctx->release(ctx->dCounter)))
        return rc;
    ctx->dStage = nullptr;
    ctx->dSat = nullptr;
    ctx->dCounter = nullptr;
    ctx->scratchCap = 0;
    if ((rc = ctx->alloc(sizeof(uint32_t) * scratch_words(cap), (void**)&ctx->dStage))) return rc;
    if ((rc = ctx->alloc(sizeof(SatResult) * (uint64_t)cap, (void**)&ctx->dSat))) return rc;
    if ((rc = ctx->alloc(4 * kCountWords, (void**)&ctx->dCounter))) return rc;
    ctx->scratchCap = cap;
    return B3AMD_OK;
}

namespace {

// Bounding cap of the Gauss-map arc of an edge with face normals U, V and direction E
// (origin to twin origin): a region containing every crossing that b3IsMinkowskiFace
// (sat.cpp:60-71), called with (U, V, -E) in place of (A, B, B x A), can accept for an edge
// pair that b3Project scores.  It is the shorter great-circle arc from U to V, centre
// m = (U + V) / |U + V|, radius angle(U, V) / 2 = atan2(|U - V|, |U + V|), widened by a
// margin; returned as m and cos, sin of the widened radius.
//
// Why the margin: the test follows the great circle normal to -E, not the arc's own circle
// (normal V x U).  With delta the angle between the two circles, a crossing it accepts lies
// within about 2 delta + delta / sin(phi) of the arc, phi being the angle between the two
// edges; b3Project returns -FLT_MAX below sin(phi) = 0.005 (sat.cpp:88-95), so those hits
// never count and delta / sin(phi) <= 200 delta.  delta is bounded by the angle between
// V x U and -E here plus what the SAT's FP32 arithmetic adds when the edge is hull 1's,
// transformed into hull 2's frame: the rotated normals (about 4e-7 / |V x U|) and the
// transformed end points (about 2.4e-6 posScale / |E|, posScale the distance of the
// hull's farthest vertex from its origin).  The margin is 0.001 rad (FP32 rounding of the
// test itself, below 1e-4 rad at sin(phi) >= 0.005) + 210 delta.
//
// The whole sphere is returned where no such cap exists or the shorter arc is not the
// tested one: normals nearly opposite (|U + V| < 0.02, as in flat two-face hulls), -E
// against or far from V x U (a reflex or degenerate edge, where the test follows the longer
// arc: delta > 1e-3 or |V x U| < 8 delta), a widened radius beyond pi/2 + 0.01, or zero or
// non-finite input.  It is encoded as m = 0, cos 0, sin 1, for which the kernel's test
// sin a1 sin a2 - cos a1 cos a2 + M1.M2 >= 0 reduces to sin a2 >= 0 and passes every pair,
// every term still at most 1 in magnitude (so the half-rounding bound of that test holds).
void arc_cap(const float* U, const float* V, const float* E, double posScale, float* m, float* cosA,
             float* sinA) {
    double u[3], v[3], e[3], nu = 0.0, nv = 0.0, ne = 0.0;
    for (int i = 0; i < 3; ++i) {
        u[i] = U[i], v[i] = V[i], e[i] = -(double)E[i];
        nu += u[i] * u[i], nv += v[i] * v[i], ne += e[i] * e[i];
    }
    nu = std::sqrt(nu), nv = std::sqrt(nv), ne = std::sqrt(ne);
    double s[3], d[3], ls = 0.0, ld = 0.0;
    for (int i = 0; i < 3; ++i) {
        u[i] /= nu, v[i] /= nv, e[i] /= ne;
        s[i] = u[i] + v[i], d[i] = u[i] - v[i];
        ls += s[i] * s[i], ld += d[i] * d[i];
    }
    ls = std::sqrt(ls), ld = std::sqrt(ld);
    const double c[3] = {v[1] * u[2] - v[2] * u[1], v[2] * u[0] - v[0] * u[2], v[0] * u[1] - v[1] * u[0]};
    const double lc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double x[3] = {c[1] * e[2] - c[2] * e[1], c[2] * e[0] - c[0] * e[2], c[0] * e[1] - c[1] * e[0]};
    const double delta = std::atan2(std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]),
                                    c[0] * e[0] + c[1] * e[1] + c[2] * e[2]) +
                         4e-7 / lc + 2.4e-6 * posScale / ne;
    const double a = std::atan2(ld, ls) + 0.001 + 210.0 * delta;
    const bool shortArc = ls >= 0.02 && delta <= 1e-3 && lc >= 8.0 * delta && a <= 1.5807963267948966;
    if (!shortArc) {  // false for NaN too
        m[0] = m[1] = m[2] = 0.0f;
        *cosA = 0.0f;
        *sinA = 1.0f;
        return;
    }
    for (int i = 0; i < 3; ++i) m[i] = (float)(s[i] / ls);
    *cosA = (float)std::cos(a);
    *sinA = (float)std::sin(a);
}

// Converts one b3Hull view into the device record (device_hull.h); returns an error
// string or nullptr.

const char* pack_hull(const b3amd_hull_pool* pool, uint32_t h, std::vector<uint32_t>& words,
                      HullLayout* layoutOut) {
    const uint32_t* d = pool->desc + 6 * (size_t)h;
    const uint32_t V = d[1], HE = d[3], F = d[5];
    if (V < 3 || V > kMaxHullVertices) return "vertexCount outside 3..255";
    if (F < 2 || F > kMaxHullFaces) return "faceCount outside 2..255";
    if (HE < 6 || HE > kMaxHullHalfEdges || (HE & 1u)) return "edgeCount must be even and within 6..4096";
    const float* verts = pool->vertices + 3 * (size_t)d[0];
    const uint32_t* edges = pool->edges + 4 * (size_t)d[2];
    const uint32_t* faces = pool->faces + d[4];
    const float* planes = pool->planes + 4 * (size_t)d[4];
    for (uint32_t e = 0; e < HE; ++e) {
        const uint32_t* he = edges + 4 * e;
        if (he[0] >= V || he[2] >= F || he[3] >= HE) return "half-edge index out of range";
        if (he[1] != (e ^ 1u)) return "half-edges 2k and 2k+1 must be twins";
    }
    uint32_t maxDeg = 0;
    for (uint32_t f = 0; f < F; ++f) {
        if (faces[f] >= HE) return "face edge out of range";
        uint32_t e = faces[f], deg = 0;
        do {
            e = edges[4 * e + 3];
            if (++deg > HE) return "face loop does not close";
        } while (e != faces[f]);
        if (deg > maxDeg) maxDeg = deg;
    }
    HullLayout L{V, F, HE};
    words.assign(4 * (size_t)L.units(), 0u);
    float* fw = reinterpret_cast<float*>(words.data());
    words[0] = V | (F << 16);
    words[1] = HE | (maxDeg << 16);
    words[2] = HE >> 1;
    std::memcpy(fw + 4, pool->centroids + 3 * (size_t)h, 12);
    for (uint32_t i = 0; i < V; ++i) std::memcpy(fw + 4 * (L.vertOff() + i), verts + 3 * i, 12);
    {
        // Bounding radius about the centroid, rounded up: no vertex lies farther than fw[7]
        // (non-finite coordinates give a non-finite radius, which no bound test passes).
        const float* c = pool->centroids + 3 * (size_t)h;
        double r2 = 0.0;
        for (uint32_t i = 0; i < V; ++i) {
            const double dx = (double)verts[3 * i] - c[0], dy = (double)verts[3 * i + 1] - c[1],
                         dz = (double)verts[3 * i + 2] - c[2];
            const double q = dx * dx + dy * dy + dz * dz;
            r2 = (q > r2 || q != q) ? q : r2;
        }
        fw[7] = std::nextafter((float)(std::sqrt(r2) * (1.0 + 1e-6)), std::numeric_limits<float>::infinity());
    }
    for (uint32_t i = 0; i < F; ++i) std::memcpy(fw + 4 * (L.planeOff() + i), planes + 4 * i, 16);
    uint32_t* ue = words.data() + 4 * L.uedgeOff();
    for (uint32_t k = 0; k < L.E(); ++k) {
        const uint32_t* a = edges + 8 * k;
        const uint32_t* b = a + 4;
        ue[k] = a[0] | (b[0] << 8) | (a[2] << 16) | (b[2] << 24);
    }
    uint32_t* hw = words.data() + 4 * L.halfOff();
    for (uint32_t e = 0; e < HE; ++e) {
        const uint32_t* he = edges + 4 * e;
        hw[e] = he[0] | (he[2] << 8) | (he[3] << 16);
    }
    uint16_t* fe = reinterpret_cast<uint16_t*>(words.data() + 4 * L.faceOff());
    for (uint32_t f = 0; f < F; ++f) fe[f] = (uint16_t)faces[f];
    // Negated Gauss map of the hull in its own frame (sat.cpp:131-156 for hull 2): -U, -V,
    // -E = -(Q - P) per unique edge and the arc's bounding cap (centre, widened radius).
    float* gm = fw + 4 * L.gaussOff();
    float* gc = fw + 4 * L.capOff();
    _Float16* gf = reinterpret_cast<_Float16*>(fw + 4 * L.filterOff());
    for (uint32_t k = L.E(); k < L.gaussEdges(); ++k) gc[4 * k + 3] = 1e30f;
    double posScale = 0.0;  // farthest vertex from the hull's origin
    for (uint32_t i = 0; i < V; ++i) {
        const double q = std::sqrt((double)verts[3 * i] * verts[3 * i] + (double)verts[3 * i + 1] * verts[3 * i + 1] +
                                   (double)verts[3 * i + 2] * verts[3 * i + 2]);
        posScale = q > posScale ? q : posScale;
    }
    for (uint32_t k = 0; k < L.E(); ++k) {
        const uint32_t* a = edges + 8 * k;
        const uint32_t* b = a + 4;
        const float* P = verts + 3 * a[0];
        const float* Q = verts + 3 * b[0];
        const float* U = planes + 4 * a[2];
        const float* Vn = planes + 4 * b[2];
        const float E[3] = {Q[0] - P[0], Q[1] - P[1], Q[2] - P[2]};
        float m[3], ca, sa;
        arc_cap(U, Vn, E, posScale, m, &ca, &sa);
        float* r = gm + 12 * k;
        r[0] = -U[0], r[1] = -U[1], r[2] = -U[2], r[3] = -Vn[0];
        r[4] = -Vn[1], r[5] = -Vn[2], r[6] = -E[0], r[7] = -E[1];
        r[8] = -E[2], r[9] = P[0], r[10] = P[1], r[11] = P[2];
        float* c = gc + 4 * k;
        c[0] = -m[0], c[1] = -m[1], c[2] = -m[2], c[3] = ca;
        // Halves: block k / 32 holds (-m, cos a) of edge k at half 4 (k % 32) and sin a at
        // half 128 + k % 32.
        _Float16* blk = gf + 160 * (k / 32);
        const float comp[4] = {-m[0], -m[1], -m[2], ca};
        for (uint32_t q = 0; q < 4; ++q) blk[4 * (k % 32) + q] = (_Float16)comp[q];
        blk[128 + k % 32] = (_Float16)sa;
    }
    *layoutOut = L;
    return nullptr;
}

// The record unit of each hull shape's hull in the current pool (0 for other shapes and for
// hull indices outside the pool, which the kernels reject before reading a record).
void set_hull_units(const b3amd_context* ctx, std::vector<ShapeRec>& shapes) {
    for (ShapeRec& r : shapes)
        r.hullUnit = (r.type == B3AMD_HULL && r.hull < ctx->hullOffsets.size() && ctx->hullOffsets[r.hull] != kFreeHull)
                         ? ctx->hullOffsets[r.hull]
                         : 0u;
}

int upload(b3amd_context* ctx, void** dptr, uint64_t* cap, const void* src, uint64_t bytes) {
    if (bytes > *cap) {
        int rc = ctx->release(*dptr);
        if (rc) return rc;
        *dptr = nullptr;
        *cap = 0;
        if ((rc = ctx->alloc(bytes, dptr))) return rc;
        *cap = bytes;
    }
    HIP_TRY(hipMemcpyAsync(*dptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return B3AMD_OK;
}

// Device records of every hull of a pool, validated before anything changes.
int pack_pool(const b3amd_hull_pool* pool, std::vector<std::vector<uint32_t>>& recs, std::vector<HullLayout>& layouts) {
    if (pool->count && (!pool->desc || !pool->centroids || !pool->vertices || !pool->edges || !pool->faces ||
                        !pool->planes))
        return fail(B3AMD_ERR_ARGUMENT, "null argument");
    recs.resize(pool->count);
    layouts.resize(pool->count);
    for (uint32_t h = 0; h < pool->count; ++h)
        if (const char* err = pack_hull(pool, h, recs[h], &layouts[h]))
            return fail(B3AMD_ERR_UNSUPPORTED, "hull %u: %s", h, err);
    return B3AMD_OK;
}

// Gives a record's units back: the range joins its neighbours, and a range at the arena's
// tail shortens the used part instead.
void release_range(b3amd_context* ctx, uint64_t first, uint64_t units) {
    auto& R = ctx->freeRanges;
    auto it = std::lower_bound(R.begin(), R.end(), std::make_pair(first, (uint64_t)0));
    it = R.insert(it, {first, units});
    if (it + 1 != R.end() && it->first + it->second == (it + 1)->first) {
        it->second += (it + 1)->second;
        R.erase(it + 1);
    }
    if (it != R.begin() && (it - 1)->first + (it - 1)->second == it->first) {
        (it - 1)->second += it->second;
        it = R.erase(it) - 1;
    }
    if (it->first + it->second == ctx->poolUsed) {
        ctx->poolUsed = it->first;
        R.erase(it);
    }
}

// Places packed records in the arena (first fit in freed ranges, else the tail, growing the
// arena once by at least double when the tail is short), uploads them and the slot offsets,
// and writes each record's slot to handles (when not null).  The plan is made before any
// device memory changes, so a failed allocation leaves the registry as it was.
int place_hulls(b3amd_context* ctx, const std::vector<std::vector<uint32_t>>& recs,
                const std::vector<HullLayout>& layouts, uint32_t* handles) {
    const size_t n = recs.size();
    std::vector<std::pair<uint64_t, uint64_t>> ranges = ctx->freeRanges;
    uint64_t used = ctx->poolUsed;
    std::vector<uint64_t> at(n);
    for (size_t i = 0; i < n; ++i) {
        const uint64_t units = recs[i].size() / 4;
        auto it = std::find_if(ranges.begin(), ranges.end(), [&](const std::pair<uint64_t, uint64_t>& r) {
            return r.second >= units;
        });
        if (it != ranges.end()) {
            at[i] = it->first;
            it->first += units;
            it->second -= units;
            if (it->second == 0) ranges.erase(it);
        } else {
            at[i] = used;
            used += units;
        }
    }
    if (used > 0xFFFFFFFFull) return fail(B3AMD_ERR_UNSUPPORTED, "hull pool larger than 2^32 units");
    size_t reuse = ctx->freeSlots.size() < n ? ctx->freeSlots.size() : n;
    const uint64_t slots = (uint64_t)ctx->hullCount + (n - reuse);
    if (slots > 0xFFFFFFFEull) return fail(B3AMD_ERR_UNSUPPORTED, "too many hulls");
    int rc;
    if (used > ctx->poolCap) {
        uint64_t cap = 2 * ctx->poolCap > used ? 2 * ctx->poolCap : used;
        if (cap > 0xFFFFFFFFull) cap = used;
        float4* p = nullptr;
        if ((rc = ctx->alloc(16 * cap, (void**)&p))) return rc;
        if (ctx->poolUsed && hipMemcpy(p, ctx->dPool, 16 * ctx->poolUsed, hipMemcpyDeviceToDevice) != hipSuccess) {
            ctx->release(p);
            return fail(B3AMD_ERR_DEVICE, "hull arena copy failed");
        }
        if ((rc = ctx->release(ctx->dPool))) return rc;
        ctx->dPool = p;
        ctx->poolCap = cap;
        ++ctx->generation;  // captured graphs hold the old arena
    }
    if (slots > ctx->offsetsCap) {
        const uint64_t cap = slots < 2ull * ctx->offsetsCap ? 2ull * ctx->offsetsCap : (slots < 64 ? 64 : slots);
        uint32_t* p = nullptr;
        if ((rc = ctx->alloc(4 * cap, (void**)&p))) return rc;
        if ((rc = ctx->release(ctx->dOffsets))) return rc;
        ctx->dOffsets = p;
        ctx->offsetsCap = (uint32_t)cap;
        ++ctx->generation;
    }
    for (size_t i = 0; i < n; ++i)
        HIP_TRY(hipMemcpy(ctx->dPool + at[i], recs[i].data(), 4 * recs[i].size(), hipMemcpyHostToDevice));
    ctx->freeRanges.swap(ranges);
    ctx->poolUsed = used;
    // Freed slots are reused lowest first.
    std::sort(ctx->freeSlots.begin(), ctx->freeSlots.end(), std::greater<uint32_t>());
    for (size_t i = 0; i < n; ++i) {
        uint32_t slot;
        if (!ctx->freeSlots.empty()) {
            slot = ctx->freeSlots.back();
            ctx->freeSlots.pop_back();
        } else {
            slot = ctx->hullCount++;
            ctx->hullOffsets.push_back(kFreeHull);
            ctx->hullUnits.push_back(0);
            ctx->layouts.push_back(HullLayout{});
        }
        ctx->hullOffsets[slot] = (uint32_t)at[i];
        ctx->hullUnits[slot] = (uint32_t)(recs[i].size() / 4);
        ctx->layouts[slot] = layouts[i];
        if (handles) handles[i] = slot;
    }
    if (ctx->hullCount)
        HIP_TRY(hipMemcpy(ctx->dOffsets, ctx->hullOffsets.data(), 4ull * ctx->hullCount, hipMemcpyHostToDevice));
    return B3AMD_OK;
}

int validate_pairs(const b3amd_context* ctx, uint32_t n, const uint32_t* pairs) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        if (a >= ctx->shapeCount || b >= ctx->shapeCount)
            return fail(B3AMD_ERR_ARGUMENT, "pair %u refers to shape %u/%u of %u", i, a, b, ctx->shapeCount);
        const ShapeRec& A = ctx->shapes[a];
        const ShapeRec& B = ctx->shapes[b];
        if (A.type > B.type) return fail(B3AMD_ERR_ARGUMENT, "pair %u: typeA > typeB (contact_manager.cpp:238-242)", i);
    }
    return B3AMD_OK;
}

}  // namespace

int b3amd::enqueue_batch(b3amd_context* ctx, uint32_t n, const uint32_t* pairs, const float* xfs,
                         const float* sweeps, ConvexCache* caches, Manifold* out, const BatchScratch& sc,
                         const uint32_t* slots, const BodyRefs* bodies, const BatchStreams* streams) {
    int rc = B3AMD_OK;
    PairInputs in;
    in.n = n;
    in.pairs = pairs;
    in.shapes = ctx->dShapes;
    in.xfs = xfs;
    in.pool = ctx->dPool;
    in.offsets = ctx->dOffsets;
    in.shapeCount = ctx->shapeCount;
    in.hullCount = ctx->hullCount;
    in.slots = slots;
    in.counters = ctx->counting ? ctx->dCounters : nullptr;
    in.bodies = bodies ? bodies->bodies : nullptr;
    in.bodyXfs = bodies ? bodies->xfs : nullptr;
    in.bodySweeps = bodies ? bodies->sweeps : nullptr;
    // sc.stage holds scratch_words(capacity) words for a capacity >= n.
    CacheArgs ca;
    static_cast<PairInputs&>(ca) = in;
    ca.counts = sc.counter;
    ca.sweeps = sweeps;
    ca.caches = caches;
    ca.stage = sc.stage;
    ca.satList = sc.stage + n;
    ca.rebuildList = sc.stage + 2 * (size_t)n;
    ca.out = out;
    ca.sat = sc.sat;
    SatArgs sa;
    static_cast<PairInputs&>(sa) = in;
    sa.list = ca.satList;
    sa.listCount = sc.counter + kCountSat;
    sa.out = sc.sat;
    sa.split = SatSplit{sc.counter + kCountSatTail, 0u};
    ContactArgs cta;
    static_cast<PairInputs&>(cta) = in;
    cta.sat = sc.sat;
    cta.caches = caches;
    cta.out = out;
    cta.list = sc.stage;  // the stage words are dead once the cache kernels are done
    cta.listCount = sc.counter + kCountContacts;
    cta.satList = ca.satList;
    cta.satCount = sc.counter + kCountSat;
    cta.wideList = ca.satList;  // free once contact_split_kernel has walked the SAT list
    cta.wideCount = sc.counter + kCountWide;
    const BatchStreams ss = streams ? *streams : BatchStreams{ctx->stream, ctx->side, ctx->forkEvent, ctx->joinEvent};
    hipStream_t s = ss.main;
    if (ctx->roundShapes) {
        // Sphere/capsule pairs depend on nothing the hull kernels write.
        ConvexArgs cv;
        static_cast<PairInputs&>(cv) = in;
        cv.out = out;
        HIP_TRY(hipEventRecord(ss.fork, s));
        HIP_TRY(hipStreamWaitEvent(ss.side, ss.fork, 0));
        if ((rc = timed(ctx, ss.side, "convex", [&] { return launch_convex(cv, ss.side); }))) return rc;
        HIP_TRY(hipEventRecord(ss.join, ss.side));
    }
    if ((rc = timed(ctx, s, "cache_hull", [&] { return launch_cache_hull(ca, s); }))) return rc;
    const SatTuning t = ctx->tuning;
    if ((rc = timed(ctx, s, "sat_hull", [&] { return launch_sat_persistent(sa, t, s); }))) return rc;
    if ((rc = timed(ctx, s, "contact_hull", [&] { return launch_contact_hull(cta, s); }))) return rc;
    if (ctx->roundShapes) HIP_TRY(hipStreamWaitEvent(s, ss.join, 0));
    return B3AMD_OK;
}

namespace {

int run_batch(b3amd_context* ctx, uint32_t n, const uint32_t* pairs, const float* xfs, const float* sweeps,
              ConvexCache* caches, Manifold* out) {
    if (n == 0) return B3AMD_OK;
    if (!ctx->dPool) return fail(B3AMD_ERR_ARGUMENT, "no hull pool uploaded");
    if (!ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    int rc = ensure_scratch(ctx, n);
    if (rc) return rc;
    return enqueue_batch(ctx, n, pairs, xfs, sweeps, caches, out, BatchScratch{ctx->dStage, ctx->dSat, ctx->dCounter},
                         nullptr);
}

}  // namespace

extern "C" {

const char* b3amd_last_error(void) { return g_error.c_str(); }

int b3amd_create(int device, uint64_t hbm_budget_bytes, b3amd_context** out) {
    if (!out) return fail(B3AMD_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(B3AMD_ERR_NO_DEVICE, "no GPU visible");
    if (device < 0 || device >= count) return fail(B3AMD_ERR_ARGUMENT, "device %d of %d", device, count);
    HIP_TRY(hipSetDevice(device));
    b3amd_context* ctx = new b3amd_context();
    ctx->device = device;
    ctx->budget = hbm_budget_bytes ? hbm_budget_bytes : (64ull << 30);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->forkEvent, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&ctx->joinEvent, hipEventDisableTiming) != hipSuccess) {
        b3amd_destroy(ctx);
        return fail(B3AMD_ERR_DEVICE, "stream or event creation failed");
    }
    if (hipHostMalloc((void**)&ctx->hWords, 4ull * b3amd_context::kHostWords, 0) != hipSuccess) {
        ctx->hWords = nullptr;
        b3amd_destroy(ctx);
        return fail(B3AMD_ERR_DEVICE, "pinned host allocation failed");
    }
    if (scratch_address(device, &ctx->noAllocScratch) != B3AMD_OK) {
        std::string msg = b3amd_last_error();
        b3amd_destroy(ctx);
        return fail(B3AMD_ERR_DEVICE, "%s", msg.c_str());
    }
    if (ctx->alloc(sizeof(Counters) * kCounterSlots, (void**)&ctx->dCounters) != B3AMD_OK ||
        hipMemset(ctx->dCounters, 0, sizeof(Counters) * kCounterSlots) != hipSuccess) {
        std::string msg = b3amd_last_error();
        b3amd_destroy(ctx);
        return fail(B3AMD_ERR_DEVICE, "counter allocation failed: %s", msg.c_str());
    }
    *out = ctx;
    return B3AMD_OK;
}

int b3amd_destroy(b3amd_context* ctx) {
    if (!ctx) return B3AMD_OK;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    if (ctx->side) hipStreamSynchronize(ctx->side);
    while (!ctx->graphs.empty()) b3amd_graph_destroy(ctx->graphs.back());
    for (auto& kv : ctx->allocs) hipFree(kv.first);
    for (PendingTiming& t : ctx->pending) {
        hipEventDestroy(t.start);
        hipEventDestroy(t.stop);
    }
    for (hipEvent_t e : ctx->freeEvents) hipEventDestroy(e);
    if (ctx->forkEvent) hipEventDestroy(ctx->forkEvent);
    if (ctx->joinEvent) hipEventDestroy(ctx->joinEvent);
    if (ctx->side) hipStreamDestroy(ctx->side);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    if (ctx->hWords) hipHostFree(ctx->hWords);
    delete ctx;
    return B3AMD_OK;
}

uint64_t b3amd_memory_in_use(const b3amd_context* ctx) { return ctx ? ctx->inUse : 0; }

int b3amd_upload_hulls(b3amd_context* ctx, const b3amd_hull_pool* pool) {
    if (!ctx || !pool) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::vector<std::vector<uint32_t>> recs;
    std::vector<HullLayout> layouts;
    int rc = pack_pool(pool, recs, layouts);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((rc = ctx->release(ctx->dPool)) || (rc = ctx->release(ctx->dOffsets))) return rc;
    ctx->dPool = nullptr;
    ctx->dOffsets = nullptr;
    ctx->poolCap = ctx->poolUsed = 0;
    ctx->offsetsCap = 0;
    ctx->hullCount = 0;
    ctx->layouts.clear();
    ctx->hullOffsets.clear();
    ctx->hullUnits.clear();
    ctx->freeSlots.clear();
    ctx->freeRanges.clear();
    ++ctx->generation;
    if ((rc = place_hulls(ctx, recs, layouts, nullptr))) return rc;
    // Shapes uploaded before this pool point into the old one: refresh their record units.
    if (ctx->shapeCount && ctx->dShapes) {
        set_hull_units(ctx, ctx->shapes);
        HIP_TRY(hipMemcpy(ctx->dShapes, ctx->shapes.data(), sizeof(ShapeRec) * (uint64_t)ctx->shapeCount,
                          hipMemcpyHostToDevice));
    }
    return B3AMD_OK;
}

int b3amd_hull_add(b3amd_context* ctx, const b3amd_hull_pool* hulls, uint32_t* handles) {
    if (!ctx || !hulls || (hulls->count && !handles)) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    std::vector<std::vector<uint32_t>> recs;
    std::vector<HullLayout> layouts;
    int rc = pack_pool(hulls, recs, layouts);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return place_hulls(ctx, recs, layouts, handles);
}

int b3amd_hull_free(b3amd_context* ctx, uint32_t handle) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    if (handle >= ctx->hullCount || ctx->hullOffsets[handle] == kFreeHull)
        return fail(B3AMD_ERR_ARGUMENT, "hull %u is not a live handle", handle);
    for (uint32_t i = 0; i < ctx->shapeCount; ++i)
        if (ctx->shapes[i].type == B3AMD_HULL && ctx->shapes[i].hull == handle)
            return fail(B3AMD_ERR_ARGUMENT, "hull %u is in use by shape %u", handle, i);
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    release_range(ctx, ctx->hullOffsets[handle], ctx->hullUnits[handle]);
    ctx->hullOffsets[handle] = kFreeHull;
    ctx->hullUnits[handle] = 0;
    ctx->freeSlots.push_back(handle);
    return B3AMD_OK;
}

int b3amd_hull_registry_info(const b3amd_context* ctx, uint32_t* slots, uint32_t* live, uint64_t* used_bytes,
                             uint64_t* capacity_bytes) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    uint64_t used = 0;
    uint32_t n = 0;
    for (uint32_t h = 0; h < ctx->hullCount; ++h)
        if (ctx->hullOffsets[h] != kFreeHull) ++n, used += ctx->hullUnits[h];
    if (slots) *slots = ctx->hullCount;
    if (live) *live = n;
    if (used_bytes) *used_bytes = 16 * used;
    if (capacity_bytes) *capacity_bytes = 16 * ctx->poolCap;
    return B3AMD_OK;
}

int b3amd_upload_shapes(b3amd_context* ctx, uint32_t count, const b3amd_shape* shapes) {
    if (!ctx || (!shapes && count)) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    static_assert(sizeof(b3amd_shape) == sizeof(ShapeRec), "shape layout");
    for (uint32_t i = 0; i < count; ++i) {
        if (shapes[i].type > B3AMD_HULL) return fail(B3AMD_ERR_UNSUPPORTED, "shape %u: type %u", i, shapes[i].type);
        if (shapes[i].type == B3AMD_HULL && shapes[i].hull >= ctx->hullCount)
            return fail(B3AMD_ERR_ARGUMENT, "shape %u: hull %u of %u", i, shapes[i].hull, ctx->hullCount);
        if (shapes[i].type == B3AMD_HULL && ctx->hullOffsets[shapes[i].hull] == kFreeHull)
            return fail(B3AMD_ERR_ARGUMENT, "shape %u: hull %u was freed", i, shapes[i].hull);
    }
    int rc;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((rc = ctx->release(ctx->dShapes))) return rc;
    ctx->dShapes = nullptr;
    ctx->shapeCount = 0;
    ctx->shapes.assign(reinterpret_cast<const ShapeRec*>(shapes), reinterpret_cast<const ShapeRec*>(shapes) + count);
    set_hull_units(ctx, ctx->shapes);
    if ((rc = ctx->alloc(sizeof(ShapeRec) * (uint64_t)count, (void**)&ctx->dShapes))) return rc;
    if (count)
        HIP_TRY(hipMemcpy(ctx->dShapes, ctx->shapes.data(), sizeof(ShapeRec) * (uint64_t)count, hipMemcpyHostToDevice));
    ctx->shapeCount = count;
    ++ctx->generation;
    ctx->roundShapes = false;
    for (uint32_t i = 0; i < count; ++i) ctx->roundShapes |= shapes[i].type != B3AMD_HULL;
    return B3AMD_OK;
}

int b3amd_collide(b3amd_context* ctx, uint32_t n, const uint32_t* pairs, const b3amd_transform* xfs,
                  const b3amd_quat* sweeps, b3amd_convex_cache* caches, b3amd_manifold* out) {
    if (!ctx || (n && (!pairs || !xfs || !out))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (n == 0) return B3AMD_OK;
    int rc = validate_pairs(ctx, n, pairs);
    if (rc) return rc;
    const uint64_t bPairs = 8ull * n, bXf = 96ull * n, bSw = sweeps ? 64ull * n : 0, bC = caches ? 36ull * n : 0,
                   bOut = 240ull * n;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t total = al(bPairs) + al(bXf) + al(bSw) + al(bC) + al(bOut);
    if (total > ctx->dInCap) {
        if ((rc = ctx->release(ctx->dIn))) return rc;
        ctx->dIn = nullptr;
        ctx->dInCap = 0;
        if ((rc = ctx->alloc(total, &ctx->dIn))) return rc;
        ctx->dInCap = total;
    }
    char* base = static_cast<char*>(ctx->dIn);
    char* dPairs = base;
    char* dXf = dPairs + al(bPairs);
    char* dSw = dXf + al(bXf);
    char* dC = dSw + al(bSw);
    char* dOut = dC + al(bC);
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(dPairs, pairs, bPairs, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dXf, xfs, bXf, hipMemcpyHostToDevice, s));
    if (sweeps) HIP_TRY(hipMemcpyAsync(dSw, sweeps, bSw, hipMemcpyHostToDevice, s));
    if (caches) HIP_TRY(hipMemcpyAsync(dC, caches, bC, hipMemcpyHostToDevice, s));
    rc = run_batch(ctx, n, reinterpret_cast<const uint32_t*>(dPairs), reinterpret_cast<const float*>(dXf),
                   sweeps ? reinterpret_cast<const float*>(dSw) : nullptr,
                   caches ? reinterpret_cast<ConvexCache*>(dC) : nullptr, reinterpret_cast<Manifold*>(dOut));
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, dOut, bOut, hipMemcpyDeviceToHost, s));
    if (caches) HIP_TRY(hipMemcpyAsync(caches, dC, bC, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ctx->resolve_timings();
}

int b3amd_contact_init(b3amd_context* ctx, uint32_t n, uint32_t M, const uint32_t* d_bodies,
                       const b3amd_solver_body* d_state, const b3amd_contact_material* d_material,
                       const uint32_t* d_counts, const b3amd_manifold* d_manifolds, const uint32_t* d_old_counts,
                       const b3amd_manifold* d_old_manifolds, const b3amd_manifold_impulses* d_old_impulses,
                       int warm_starting, b3amd_manifold_impulses* d_impulses, b3amd_velocity_manifold* d_constraints) {
    static_assert(sizeof(b3amd_solver_body) == 4 * kSolverBodyFloats, "solver body layout");
    static_assert(sizeof(b3amd_velocity_manifold) == sizeof(VelocityManifold), "velocity constraint layout");
    static_assert(sizeof(b3amd_manifold_impulses) == sizeof(ManifoldImpulses), "impulse layout");
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    if (n == 0) return B3AMD_OK;
    if (M == 0 || M > 3) return fail(B3AMD_ERR_ARGUMENT, "manifold slots per contact %u outside 1..3", M);
    if (!d_bodies || !d_state || !d_material || !d_counts || !d_manifolds || !d_impulses || !d_constraints ||
        (warm_starting && (!d_old_counts || !d_old_manifolds || !d_old_impulses)))
        return fail(B3AMD_ERR_ARGUMENT, "null argument");
    ContactInitArgs a;
    a.n = n;
    a.M = M;
    a.bodies = d_bodies;
    a.state = reinterpret_cast<const float*>(d_state);
    a.material = reinterpret_cast<const float*>(d_material);
    a.counts = d_counts;
    a.manifolds = reinterpret_cast<const Manifold*>(d_manifolds);
    a.oldCounts = d_old_counts;
    a.oldManifolds = reinterpret_cast<const Manifold*>(d_old_manifolds);
    a.oldImpulses = reinterpret_cast<const ManifoldImpulses*>(d_old_impulses);
    a.warmStarting = warm_starting ? 1u : 0u;
    a.impulses = reinterpret_cast<ManifoldImpulses*>(d_impulses);
    a.constraints = reinterpret_cast<VelocityManifold*>(d_constraints);
    return timed(ctx, ctx->stream, "contact_init", [&] { return launch_contact_init(a, ctx->stream); });
}

int b3amd_test_overlap_device(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                              b3amd_convex_cache* d_caches, uint32_t* d_overlap) {
    if (!ctx || (n && (!d_pairs || !d_xfs || !d_overlap))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (n == 0) return B3AMD_OK;
    if (!ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    OverlapArgs a;
    a.n = n;
    a.pairs = d_pairs;
    a.shapes = ctx->dShapes;
    a.xfs = reinterpret_cast<const float*>(d_xfs);
    a.pool = ctx->dPool;
    a.offsets = ctx->dOffsets;
    a.shapeCount = ctx->shapeCount;
    a.hullCount = ctx->hullCount;
    a.slots = nullptr;
    a.counters = ctx->counting ? ctx->dCounters : nullptr;
    a.caches = reinterpret_cast<ConvexCache*>(d_caches);
    a.overlap = d_overlap;
    return timed(ctx, ctx->stream, "overlap", [&] { return launch_overlap(a, ctx->stream); });
}

int b3amd_test_overlap(b3amd_context* ctx, uint32_t n, const uint32_t* pairs, const b3amd_transform* xfs,
                       b3amd_convex_cache* caches, uint32_t* overlap) {
    if (!ctx || (n && (!pairs || !xfs || !overlap))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (n == 0) return B3AMD_OK;
    int rc = validate_pairs(ctx, n, pairs);
    if (rc) return rc;
    const uint64_t bPairs = 8ull * n, bXf = 96ull * n, bC = caches ? 36ull * n : 0, bOut = 4ull * n;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t total = al(bPairs) + al(bXf) + al(bC) + al(bOut);
    if (total > ctx->dInCap) {
        if ((rc = ctx->release(ctx->dIn))) return rc;
        ctx->dIn = nullptr;
        ctx->dInCap = 0;
        if ((rc = ctx->alloc(total, &ctx->dIn))) return rc;
        ctx->dInCap = total;
    }
    char* dPairs = static_cast<char*>(ctx->dIn);
    char* dXf = dPairs + al(bPairs);
    char* dC = dXf + al(bXf);
    char* dOut = dC + al(bC);
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(dPairs, pairs, bPairs, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(dXf, xfs, bXf, hipMemcpyHostToDevice, s));
    if (caches) HIP_TRY(hipMemcpyAsync(dC, caches, bC, hipMemcpyHostToDevice, s));
    rc = b3amd_test_overlap_device(ctx, n, reinterpret_cast<const uint32_t*>(dPairs),
                                   reinterpret_cast<const b3amd_transform*>(dXf),
                                   caches ? reinterpret_cast<b3amd_convex_cache*>(dC) : nullptr,
                                   reinterpret_cast<uint32_t*>(dOut));
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(overlap, dOut, bOut, hipMemcpyDeviceToHost, s));
    if (caches) HIP_TRY(hipMemcpyAsync(caches, dC, bC, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ctx->resolve_timings();
}

int b3amd_collide_device(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                         const b3amd_quat* d_sweeps, b3amd_convex_cache* d_caches, b3amd_manifold* d_out) {
    if (!ctx || (n && (!d_pairs || !d_xfs || !d_out))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (d_sweeps && !d_caches) return fail(B3AMD_ERR_ARGUMENT, "sweeps without caches");
    return run_batch(ctx, n, d_pairs, reinterpret_cast<const float*>(d_xfs), reinterpret_cast<const float*>(d_sweeps),
                     reinterpret_cast<ConvexCache*>(d_caches), reinterpret_cast<Manifold*>(d_out));
}

int b3amd_collide_device_noalloc(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs,
                                 const b3amd_transform* d_xfs, const b3amd_quat* d_sweeps,
                                 b3amd_convex_cache* d_caches, b3amd_manifold* d_out) {
    if (!ctx || (n && (!d_pairs || !d_xfs || !d_out))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (d_sweeps && !d_caches) return fail(B3AMD_ERR_ARGUMENT, "sweeps without caches");
    if (n > kNoAllocMaxPairs)
        return fail(B3AMD_ERR_ARGUMENT, "%u pairs exceed the no-allocation maximum of %u per call", n, kNoAllocMaxPairs);
    if (n == 0) return B3AMD_OK;
    if (!ctx->dPool) return fail(B3AMD_ERR_ARGUMENT, "no hull pool uploaded");
    if (!ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    if (!ctx->noAllocScratch) return fail(B3AMD_ERR_DEVICE, "no scratch address for device %d", ctx->device);
    uint32_t* base = static_cast<uint32_t*>(ctx->noAllocScratch);
    const BatchScratch sc{base, reinterpret_cast<SatResult*>(base + kNoAllocStageWords),
                          base + kNoAllocStageWords + kNoAllocSatWords};
    // Kernel timing would create events on demand: off for this call.
    const bool profiling = ctx->profiling;
    ctx->profiling = false;
    const int rc = enqueue_batch(ctx, n, d_pairs, reinterpret_cast<const float*>(d_xfs),
                                 reinterpret_cast<const float*>(d_sweeps), reinterpret_cast<ConvexCache*>(d_caches),
                                 reinterpret_cast<Manifold*>(d_out), sc, nullptr);
    ctx->profiling = profiling;
    return rc;
}

int b3amd_update_contacts(b3amd_context* ctx, uint32_t body_count, const uint32_t* d_body_shapes,
                          const b3amd_transform* d_body_xfs, const b3amd_quat* d_body_sweeps,
                          uint32_t contact_count, const uint32_t* d_contacts, b3amd_convex_cache* d_caches,
                          b3amd_manifold* d_out) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    if (contact_count == 0) return B3AMD_OK;
    if (!d_body_shapes || !d_body_xfs || !d_contacts || !d_out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (d_body_sweeps && !d_caches) return fail(B3AMD_ERR_ARGUMENT, "sweeps without caches");
    if ((reinterpret_cast<uintptr_t>(d_body_xfs) | reinterpret_cast<uintptr_t>(d_body_sweeps)) & 15u)
        return fail(B3AMD_ERR_ARGUMENT, "body transforms and sweeps must be 16-byte aligned");
    if (!ctx->dPool) return fail(B3AMD_ERR_ARGUMENT, "no hull pool uploaded");
    if (!ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    int rc;
    // The contacts' shape pairs (8 B per contact); transforms and sweeps stay per body.
    if (contact_count > ctx->gatherCap) {
        const uint64_t cap = contact_count < 1024u ? 1024u : contact_count;
        if ((rc = ctx->release(ctx->dGather))) return rc;
        ctx->dGather = nullptr;
        ctx->gatherCap = 0;
        if ((rc = ctx->alloc(8 * cap, &ctx->dGather))) return rc;
        ctx->gatherCap = (uint32_t)cap;
    }
    if ((rc = ensure_scratch(ctx, contact_count))) return rc;
    BodyGatherArgs g;
    g.n = contact_count;
    g.bodyCount = body_count;
    g.bodyShapes = d_body_shapes;
    g.contacts = d_contacts;
    g.pairs = static_cast<uint32_t*>(ctx->dGather);
    hipStream_t s = ctx->stream;
    if ((rc = timed(ctx, s, "gather", [&] { return launch_body_gather(g, s); }))) return rc;
    const BodyRefs bodies{d_contacts, reinterpret_cast<const float*>(d_body_xfs),
                          reinterpret_cast<const float*>(d_body_sweeps)};
    return enqueue_batch(ctx, contact_count, g.pairs, nullptr, nullptr, reinterpret_cast<ConvexCache*>(d_caches),
                         reinterpret_cast<Manifold*>(d_out), BatchScratch{ctx->dStage, ctx->dSat, ctx->dCounter},
                         nullptr, &bodies);
}

int b3amd_set_sat_tuning(b3amd_context* ctx, uint32_t waves_per_simd, uint32_t tail_chunk) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    if (waves_per_simd > 32) return fail(B3AMD_ERR_ARGUMENT, "waves_per_simd %u > 32", waves_per_simd);
    if (tail_chunk > 64) return fail(B3AMD_ERR_ARGUMENT, "tail_chunk %u outside 0..64", tail_chunk);
    ctx->tuning.wavesPerSimd = waves_per_simd;
    ctx->tuning.tailChunk = tail_chunk;
    return B3AMD_OK;
}

int b3amd_get_sat_tuning(const b3amd_context* ctx, uint32_t* waves_per_simd, uint32_t* tail_chunk) {
    if (!ctx || !waves_per_simd || !tail_chunk) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    *waves_per_simd = ctx->tuning.wavesPerSimd;
    *tail_chunk = ctx->tuning.tailChunk;
    return B3AMD_OK;
}

int b3amd_autotune(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                   b3amd_manifold* d_out, uint32_t repeats, float* best_ms) {
    if (!ctx || !n || !d_pairs || !d_xfs || !d_out) return fail(B3AMD_ERR_ARGUMENT, "null or empty batch");
    if (repeats == 0) repeats = 1;
    static const uint32_t kWaves[] = {6, 5, 4, 3};
    const SatTuning saved = ctx->tuning;
    const bool savedProfiling = ctx->profiling;
    ctx->profiling = false;
    hipEvent_t e0 = ctx->event(), e1 = ctx->event();
    if (!e0 || !e1) return fail(B3AMD_ERR_DEVICE, "hipEventCreate failed");
    SatTuning best = saved;
    float bestMs = -1.0f;
    int rc = B3AMD_OK;
    static const uint32_t kChunks[] = {4, 8, 16, 32};
    for (uint32_t chunk : kChunks) {
        if (rc != B3AMD_OK) break;
        for (uint32_t w : kWaves) {
            ctx->tuning = SatTuning{w, chunk};
            // One untimed batch, then the timed ones; no feature caches, so every hull
            // pair runs the full SAT each time.
            if ((rc = b3amd_collide_device(ctx, n, d_pairs, d_xfs, nullptr, nullptr, d_out))) break;
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = fail(B3AMD_ERR_DEVICE, "hipEventRecord"); break; }
            for (uint32_t r = 0; r < repeats && rc == B3AMD_OK; ++r)
                rc = b3amd_collide_device(ctx, n, d_pairs, d_xfs, nullptr, nullptr, d_out);
            if (rc) break;
            float ms = 0.0f;
            if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fail(B3AMD_ERR_DEVICE, "event timing failed");
                break;
            }
            ms /= (float)repeats;
            if (bestMs < 0.0f || ms < bestMs) {
                bestMs = ms;
                best = ctx->tuning;
            }
        }
    }
    ctx->freeEvents.push_back(e0);
    ctx->freeEvents.push_back(e1);
    ctx->profiling = savedProfiling;
    ctx->tuning = rc == B3AMD_OK ? best : saved;
    if (rc == B3AMD_OK && best_ms) *best_ms = bestMs;
    return rc;
}

int b3amd_upload_mesh(b3amd_context* ctx, uint32_t vertexCount, const float* vertices, uint32_t triangleCount,
                      const uint32_t* triangles, uint32_t* meshId) {
    if (!ctx || !vertices || !triangles || !meshId || vertexCount == 0 || triangleCount == 0)
        return fail(B3AMD_ERR_ARGUMENT, "null or empty mesh");
    std::vector<TreeNode> nodes;
    if (!build_static_tree(vertices, vertexCount, triangles, triangleCount, nodes))
        return fail(B3AMD_ERR_ARGUMENT, "triangle refers to a vertex out of range");
    std::vector<float4> v4(vertexCount);
    for (uint32_t i = 0; i < vertexCount; ++i)
        v4[i] = make_float4(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2], 0.0f);
    std::vector<uint4> t4(triangleCount);
    for (uint32_t i = 0; i < triangleCount; ++i)
        t4[i] = make_uint4(triangles[3 * i], triangles[3 * i + 1], triangles[3 * i + 2], 0u);
    b3amd_context::MeshRecord m;
    int rc;
    if ((rc = ctx->alloc(sizeof(float4) * (uint64_t)vertexCount, (void**)&m.verts))) return rc;
    if ((rc = ctx->alloc(sizeof(uint4) * (uint64_t)triangleCount, (void**)&m.tris))) return rc;
    if ((rc = ctx->alloc(sizeof(TreeNode) * (uint64_t)nodes.size(), (void**)&m.nodes))) return rc;
    HIP_TRY(hipMemcpy(m.verts, v4.data(), sizeof(float4) * v4.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m.tris, t4.data(), sizeof(uint4) * t4.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(m.nodes, nodes.data(), sizeof(TreeNode) * nodes.size(), hipMemcpyHostToDevice));
    m.vertexCount = vertexCount;
    m.triangleCount = triangleCount;
    m.nodeCount = (uint32_t)nodes.size();
    ctx->meshes.push_back(m);
    *meshId = (uint32_t)ctx->meshes.size() - 1;
    return B3AMD_OK;
}

int b3amd_build_static_tree(uint32_t vertexCount, const float* vertices, uint32_t triangleCount,
                            const uint32_t* triangles, float* nodesOut) {
    if (!vertices || !triangles || !nodesOut || triangleCount == 0) return fail(B3AMD_ERR_ARGUMENT, "null or empty mesh");
    std::vector<TreeNode> nodes;
    if (!build_static_tree(vertices, vertexCount, triangles, triangleCount, nodes))
        return fail(B3AMD_ERR_ARGUMENT, "triangle refers to a vertex out of range");
    std::memcpy(nodesOut, nodes.data(), sizeof(TreeNode) * nodes.size());
    return B3AMD_OK;
}

int b3amd_hull_edge_caps(const b3amd_hull_pool* pool, uint32_t hull, float* capsOut) {
    if (!pool || !capsOut || hull >= pool->count) return fail(B3AMD_ERR_ARGUMENT, "null pool or hull out of range");
    std::vector<uint32_t> words;
    HullLayout L{};
    if (const char* err = pack_hull(pool, hull, words, &L)) return fail(B3AMD_ERR_ARGUMENT, err);
    const float* fw = reinterpret_cast<const float*>(words.data());
    const float* gc = fw + 4 * L.capOff();
    const uint32_t* d = pool->desc + 6 * (size_t)hull;
    const float* verts = pool->vertices + 3 * (size_t)d[0];
    const uint32_t* edges = pool->edges + 4 * (size_t)d[2];
    const float* planes = pool->planes + 4 * (size_t)d[4];
    double posScale = 0.0;
    for (uint32_t i = 0; i < L.V; ++i) {
        const double q = std::sqrt((double)verts[3 * i] * verts[3 * i] + (double)verts[3 * i + 1] * verts[3 * i + 1] +
                                   (double)verts[3 * i + 2] * verts[3 * i + 2]);
        posScale = q > posScale ? q : posScale;
    }
    for (uint32_t k = 0; k < L.E(); ++k) {
        // The record holds -m and cos (cap unit); sin is stored only as a half, so it is
        // recomputed here as the packing computed it.
        const uint32_t* a = edges + 8 * k;
        const uint32_t* b = a + 4;
        const float* P = verts + 3 * a[0];
        const float* Q = verts + 3 * b[0];
        const float E[3] = {Q[0] - P[0], Q[1] - P[1], Q[2] - P[2]};
        float m[3], ca, sa;
        arc_cap(planes + 4 * a[2], planes + 4 * b[2], E, posScale, m, &ca, &sa);
        capsOut[5 * k + 0] = -gc[4 * k + 0];
        capsOut[5 * k + 1] = -gc[4 * k + 1];
        capsOut[5 * k + 2] = -gc[4 * k + 2];
        capsOut[5 * k + 3] = gc[4 * k + 3];
        capsOut[5 * k + 4] = sa;
    }
    return B3AMD_OK;
}

int b3amd_mesh_tree(b3amd_context* ctx, uint32_t meshId, uint32_t* nodeCount, float* nodesOut) {
    if (!ctx || !nodeCount || meshId >= ctx->meshes.size()) return fail(B3AMD_ERR_ARGUMENT, "bad mesh id");
    const auto& m = ctx->meshes[meshId];
    *nodeCount = m.nodeCount;
    if (nodesOut) HIP_TRY(hipMemcpy(nodesOut, m.nodes, sizeof(TreeNode) * (uint64_t)m.nodeCount, hipMemcpyDeviceToHost));
    return B3AMD_OK;
}

}  // extern "C"

namespace {

// Grows the persistent mesh scratch to n contacts: staged inputs, per-contact midphase
// results, the segment work lists and the per-triangle manifolds (n x cap bound, so no
// step needs the work-item count on the host).
struct MeshScratch {
    uint32_t* shapeA;
    float* xfs;
    uint32_t* triCounts;
    uint32_t* triLists;
    uint32_t* flags;
    uint32_t* offsets;
    uint32_t* totals;
    uint2* pairContact;
    Manifold* out;
    uint32_t* manifoldCounts;
    Manifold* temp;
    float* sweeps;   // staged per-contact sweeps (persistent contacts)
    float* centers;  // staged body-A centres (persistent contacts)
    uint32_t* wideItems;
    uint32_t* wideContacts;
    void* wideSlabs;
};

int mesh_scratch(b3amd_context* ctx, uint32_t n, MeshScratch* m) {
    const uint64_t cap = kMeshTriangleCap;
    auto al = [](uint64_t x) { return (x + 255) & ~255ull; };
    const uint64_t sizes[16] = {4ull * n,  96ull * n, 4ull * n,   4ull * n * cap, 4ull * n,
                                4ull * n,  256,       24ull * n * cap, 720ull * n, 4ull * n,
                                sizeof(Manifold) * n * cap, 64ull * n, 12ull * n, 4ull * n * cap,
                                4ull * n, kWideSlabBytes * kWideContactBlocks};
    uint64_t total = 0;
    for (uint64_t b : sizes) total += al(b);
    if (total > ctx->meshScratchCap) {
        int rc;
        if ((rc = ctx->release(ctx->dMeshScratch))) return rc;
        ctx->dMeshScratch = nullptr;
        ctx->meshScratchCap = 0;
        if ((rc = ctx->alloc(total, &ctx->dMeshScratch))) return rc;
        ctx->meshScratchCap = total;
    }
    char* p = static_cast<char*>(ctx->dMeshScratch);
    void* parts[16];
    for (int i = 0; i < 16; ++i) {
        parts[i] = p;
        p += al(sizes[i]);
    }
    *m = MeshScratch{(uint32_t*)parts[0], (float*)parts[1], (uint32_t*)parts[2],  (uint32_t*)parts[3],
                     (uint32_t*)parts[4], (uint32_t*)parts[5], (uint32_t*)parts[6], (uint2*)parts[7],
                     (Manifold*)parts[8], (uint32_t*)parts[9], (Manifold*)parts[10], (float*)parts[11],
                     (float*)parts[12], (uint32_t*)parts[13], (uint32_t*)parts[14], parts[15]};
    return B3AMD_OK;
}

// Enqueues midphase, per-triangle narrowphase and clustering for n contacts whose inputs
// are on the device (no synchronisation).
int enqueue_mesh(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* shapeA, const float* xfs,
                 uint32_t* triCounts, uint32_t* triLists, Manifold* out, uint32_t* manifoldCounts,
                 const MeshScratch& sc) {
    MeshArgs a;
    a.counters = ctx->counting ? ctx->dCounters : nullptr;
    a.n = n;
    a.shapeA = shapeA;
    a.xfs = xfs;
    a.triCounts = triCounts ? triCounts : sc.triCounts;
    a.triLists = triLists ? triLists : sc.triLists;
    a.contactFlags = sc.flags;
    a.contactOffset = sc.offsets;
    a.totals = sc.totals;
    a.pairContact = sc.pairContact;
    a.wideItems = sc.wideItems;
    a.wideContacts = sc.wideContacts;
    a.wideSlabs = sc.wideSlabs;
    a.out = out;
    a.manifoldCounts = manifoldCounts;
    a.temp = sc.temp;
    a.triCaches = nullptr;
    a.sweeps = nullptr;
    a.fat = nullptr;
    a.centers0 = nullptr;
    a.shapes = ctx->dShapes;
    a.shapeCount = ctx->shapeCount;
    a.hullCount = ctx->hullCount;
    a.pool = ctx->dPool;
    a.offsets = ctx->dOffsets;
    const auto& mr = ctx->meshes[meshId];
    a.mesh = MeshDev{mr.verts, mr.tris, reinterpret_cast<const uint32_t*>(mr.nodes), mr.nodeCount};
    a.cap = kMeshTriangleCap;
    hipStream_t s = ctx->stream;
    int rc;
    if ((rc = timed(ctx, s, "zero", [&] { return launch_zero_words(a.totals, 8, s); }))) return rc;
    if ((rc = timed(ctx, s, "mesh_midphase", [&] { return launch_mesh_midphase(a, s); }))) return rc;
    if ((rc = timed(ctx, s, "mesh_narrow", [&] { return launch_mesh_narrow(a, s); }))) return rc;
    if ((rc = timed(ctx, s, "mesh_cluster", [&] { return launch_mesh_cluster(a, s); }))) return rc;
    return timed(ctx, s, "mesh_contact_wide", [&] { return launch_mesh_contact_wide(a, s); });
}

int check_mesh_call(b3amd_context* ctx, uint32_t meshId, uint32_t n) {
    if (meshId >= ctx->meshes.size()) return fail(B3AMD_ERR_ARGUMENT, "bad mesh id %u", meshId);
    if (n && !ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    if ((uint64_t)n * kMeshTriangleCap > 0xFFFFFFFFull)
        return fail(B3AMD_ERR_UNSUPPORTED, "%u mesh contacts exceed the work-item index range", n);
    return B3AMD_OK;
}

}  // namespace

extern "C" {

int b3amd_collide_mesh(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* shapeA,
                       const b3amd_transform* xfs, uint32_t* triangleCounts, uint32_t* triangleLists,
                       b3amd_manifold* out, uint32_t* manifoldCounts) {
    if (!ctx || (n && (!shapeA || !xfs || !out || !manifoldCounts))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    int rc;
    if ((rc = check_mesh_call(ctx, meshId, n))) return rc;
    if (n == 0) return B3AMD_OK;
    for (uint32_t i = 0; i < n; ++i) {
        if (shapeA[i] >= ctx->shapeCount) return fail(B3AMD_ERR_ARGUMENT, "contact %u: shape %u of %u", i, shapeA[i], ctx->shapeCount);
        if (ctx->shapes[shapeA[i]].type > B3AMD_HULL) return fail(B3AMD_ERR_ARGUMENT, "contact %u: shape A must be convex", i);
    }
    MeshScratch sc;
    if ((rc = mesh_scratch(ctx, n, &sc))) return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(sc.shapeA, shapeA, 4ull * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc.xfs, xfs, 96ull * n, hipMemcpyHostToDevice, s));
    if ((rc = enqueue_mesh(ctx, meshId, n, sc.shapeA, sc.xfs, nullptr, nullptr, sc.out, sc.manifoldCounts, sc)))
        return rc;
    hipError_t e = hipMemcpyAsync(out, sc.out, 720ull * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(manifoldCounts, sc.manifoldCounts, 4ull * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && triangleCounts)
        e = hipMemcpyAsync(triangleCounts, sc.triCounts, 4ull * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && triangleLists)
        e = hipMemcpyAsync(triangleLists, sc.triLists, 4ull * n * kMeshTriangleCap, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(B3AMD_ERR_DEVICE, "mesh readback: %s", hipGetErrorString(e));
    return ctx->resolve_timings();
}

int b3amd_collide_mesh_device(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* d_shapeA,
                              const b3amd_transform* d_xfs, uint32_t* d_triangleCounts, uint32_t* d_triangleLists,
                              b3amd_manifold* d_out, uint32_t* d_manifoldCounts) {
    if (!ctx || (n && (!d_shapeA || !d_xfs || !d_out || !d_manifoldCounts)))
        return fail(B3AMD_ERR_ARGUMENT, "null argument");
    int rc;
    if ((rc = check_mesh_call(ctx, meshId, n))) return rc;
    if (n == 0) return B3AMD_OK;
    MeshScratch sc;
    if ((rc = mesh_scratch(ctx, n, &sc))) return rc;
    return enqueue_mesh(ctx, meshId, n, d_shapeA, reinterpret_cast<const float*>(d_xfs), d_triangleCounts,
                        d_triangleLists, reinterpret_cast<Manifold*>(d_out), d_manifoldCounts, sc);
}

}  // extern "C"

struct b3amd_mesh_contacts {
    b3amd_context* ctx = nullptr;
    uint32_t meshId = 0, n = 0;
    uint32_t* dShapeA = nullptr;
    float* dFat = nullptr;            // [n][6]
    uint32_t* dTriCounts = nullptr;   // [n]
    uint32_t* dTriLists = nullptr;    // [n][cap]
    uint32_t* dFlags = nullptr;       // [n]
    ConvexCache* dCaches = nullptr;   // [n][cap]
};

namespace {

// MeshArgs over a persistent contact set and the per-call scratch.
MeshArgs mesh_set_args(b3amd_mesh_contacts* set, const MeshScratch& sc) {
    b3amd_context* ctx = set->ctx;
    MeshArgs a;
    a.counters = ctx->counting ? ctx->dCounters : nullptr;
    a.n = set->n;
    a.shapeA = set->dShapeA;
    a.shapes = ctx->dShapes;
    a.shapeCount = ctx->shapeCount;
    a.hullCount = ctx->hullCount;
    a.xfs = sc.xfs;
    a.pool = ctx->dPool;
    a.offsets = ctx->dOffsets;
    const auto& mr = ctx->meshes[set->meshId];
    a.mesh = MeshDev{mr.verts, mr.tris, reinterpret_cast<const uint32_t*>(mr.nodes), mr.nodeCount};
    a.cap = kMeshTriangleCap;
    a.triCounts = set->dTriCounts;
    a.triLists = set->dTriLists;
    a.contactFlags = set->dFlags;
    a.contactOffset = sc.offsets;
    a.totals = sc.totals;
    a.pairContact = sc.pairContact;
    a.wideItems = sc.wideItems;
    a.temp = sc.temp;
    a.out = sc.out;
    a.manifoldCounts = sc.manifoldCounts;
    a.triCaches = set->dCaches;
    a.sweeps = nullptr;
    a.fat = set->dFat;
    a.centers0 = nullptr;
    return a;
}

}  // namespace

extern "C" {

int b3amd_mesh_contacts_create(b3amd_context* ctx, uint32_t meshId, uint32_t n, const uint32_t* shapeA,
                               const b3amd_transform* xfs, b3amd_mesh_contacts** out) {
    if (!ctx || !out || (n && (!shapeA || !xfs))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    *out = nullptr;
    int rc;
    if ((rc = check_mesh_call(ctx, meshId, n))) return rc;
    if (!ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no shape table uploaded");
    for (uint32_t i = 0; i < n; ++i) {
        if (shapeA[i] >= ctx->shapeCount) return fail(B3AMD_ERR_ARGUMENT, "contact %u: shape %u of %u", i, shapeA[i], ctx->shapeCount);
        if (ctx->shapes[shapeA[i]].type > B3AMD_HULL) return fail(B3AMD_ERR_ARGUMENT, "contact %u: shape A must be convex", i);
    }
    b3amd_mesh_contacts* set = new b3amd_mesh_contacts();
    set->ctx = ctx;
    set->meshId = meshId;
    set->n = n;
    const uint64_t cap = kMeshTriangleCap, m = n ? n : 1;
    if ((rc = ctx->alloc(4 * m, (void**)&set->dShapeA)) || (rc = ctx->alloc(24 * m, (void**)&set->dFat)) ||
        (rc = ctx->alloc(4 * m, (void**)&set->dTriCounts)) || (rc = ctx->alloc(4 * m * cap, (void**)&set->dTriLists)) ||
        (rc = ctx->alloc(4 * m, (void**)&set->dFlags)) ||
        (rc = ctx->alloc(sizeof(ConvexCache) * m * cap, (void**)&set->dCaches))) {
        const std::string msg = b3amd_last_error();
        b3amd_mesh_contacts_destroy(set);
        return fail(rc, "%s", msg.c_str());
    }
    *out = set;
    if (n == 0) return B3AMD_OK;
    MeshScratch sc;
    if ((rc = mesh_scratch(ctx, n, &sc))) return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(set->dShapeA, shapeA, 4ull * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc.xfs, xfs, 96ull * n, hipMemcpyHostToDevice, s));
    const MeshArgs a = mesh_set_args(set, sc);
    if ((rc = timed(ctx, s, "mesh_sync", [&] { return launch_mesh_sync(a, kMeshCreate, s); }))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    return ctx->resolve_timings();
}

int b3amd_mesh_contacts_destroy(b3amd_mesh_contacts* set) {
    if (!set) return B3AMD_OK;
    b3amd_context* ctx = set->ctx;
    hipStreamSynchronize(ctx->stream);
    for (void* p : {(void*)set->dShapeA, (void*)set->dFat, (void*)set->dTriCounts, (void*)set->dTriLists,
                    (void*)set->dFlags, (void*)set->dCaches})
        ctx->release(p);
    delete set;
    return B3AMD_OK;
}

namespace {

// b3MeshContact::Collide of every contact of the set on device-resident transforms (and
// sweeps, or null), manifolds and counts into device buffers; enqueued, not synchronised.
int enqueue_mesh_contacts_collide(b3amd_mesh_contacts* set, const float* d_xfs, const float* d_sweeps, Manifold* d_out,
                                  uint32_t* d_counts, const MeshScratch& sc) {
    b3amd_context* ctx = set->ctx;
    hipStream_t s = ctx->stream;
    MeshArgs a = mesh_set_args(set, sc);
    a.xfs = d_xfs;
    a.sweeps = d_sweeps;
    a.out = d_out;
    a.manifoldCounts = d_counts;
    int rc;
    if ((rc = timed(ctx, s, "zero", [&] { return launch_zero_words(a.totals, 8, s); }))) return rc;
    if ((rc = timed(ctx, s, "mesh_list", [&] { return launch_mesh_list(a, s); }))) return rc;
    if ((rc = timed(ctx, s, "mesh_narrow", [&] { return launch_mesh_narrow(a, s); }))) return rc;
    return timed(ctx, s, "mesh_cluster", [&] { return launch_mesh_cluster(a, s); });
}

int enqueue_mesh_contacts_overlap(b3amd_mesh_contacts* set, const float* d_xfs, uint32_t* d_overlap,
                                  const MeshScratch& sc) {
    b3amd_context* ctx = set->ctx;
    hipStream_t s = ctx->stream;
    MeshArgs a = mesh_set_args(set, sc);
    a.xfs = d_xfs;
    a.overlap = d_overlap;
    return timed(ctx, s, "mesh_overlap", [&] { return launch_mesh_overlap(a, s); });
}

int enqueue_mesh_contacts_sync(b3amd_mesh_contacts* set, const float* d_xfs, const float* d_centers0,
                               const MeshScratch& sc) {
    b3amd_context* ctx = set->ctx;
    hipStream_t s = ctx->stream;
    MeshArgs a = mesh_set_args(set, sc);
    a.xfs = d_xfs;
    a.centers0 = d_centers0;
    return timed(ctx, s, "mesh_sync", [&] { return launch_mesh_sync(a, kMeshSync, s); });
}

}  // namespace

int b3amd_mesh_contacts_collide(b3amd_mesh_contacts* set, const b3amd_transform* xfs, const b3amd_quat* sweeps,
                                b3amd_manifold* out, uint32_t* manifoldCounts) {
    if (!set || (set->n && (!xfs || !out || !manifoldCounts))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = set->ctx;
    const uint32_t n = set->n;
    if (n == 0) return B3AMD_OK;
    int rc;
    MeshScratch sc;
    if ((rc = mesh_scratch(ctx, n, &sc))) return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(sc.xfs, xfs, 96ull * n, hipMemcpyHostToDevice, s));
    if (sweeps) HIP_TRY(hipMemcpyAsync(sc.sweeps, sweeps, 64ull * n, hipMemcpyHostToDevice, s));
    if ((rc = enqueue_mesh_contacts_collide(set, sc.xfs, sweeps ? sc.sweeps : nullptr, sc.out, sc.manifoldCounts, sc)))
        return rc;
    hipError_t e = hipMemcpyAsync(out, sc.out, 720ull * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(manifoldCounts, sc.manifoldCounts, 4ull * n, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail(B3AMD_ERR_DEVICE, "mesh readback: %s", hipGetErrorString(e));
    return ctx->resolve_timings();
}

int b3amd_mesh_contacts_collide_device(b3amd_mesh_contacts* set, const b3amd_transform* d_xfs,
                                       const b3amd_quat* d_sweeps, b3amd_manifold* d_out, uint32_t* d_manifoldCounts) {
    if (!set || (set->n && (!d_xfs || !d_out || !d_manifoldCounts))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (set->n == 0) return B3AMD_OK;
    int rc;
    MeshScratch sc;
    if ((rc = mesh_scratch(set->ctx, set->n, &sc))) return rc;
    return enqueue_mesh_contacts_collide(set, reinterpret_cast<const float*>(d_xfs),
                                         reinterpret_cast<const float*>(d_sweeps), reinterpret_cast<Manifold*>(d_out),
                                         d_manifoldCounts, sc);
}

int b3amd_mesh_contacts_test_overlap(b3amd_mesh_contacts* set, const b3amd_transform* xfs, uint32_t* overlap) {
    if (!set || (set->n && (!xfs || !overlap))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = set->ctx;
    const uint32_t n = set->n;
    if (n == 0) return B3AMD_OK;
    int rc;
    MeshScratch sc;
    if ((rc = mesh_scratch(ctx, n, &sc))) return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(sc.xfs, xfs, 96ull * n, hipMemcpyHostToDevice, s));
    // The result goes to [n] words of the scratch that only Collide uses.
    if ((rc = enqueue_mesh_contacts_overlap(set, sc.xfs, sc.manifoldCounts, sc))) return rc;
    HIP_TRY(hipMemcpyAsync(overlap, sc.manifoldCounts, 4ull * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return ctx->resolve_timings();
}

int b3amd_mesh_contacts_test_overlap_device(b3amd_mesh_contacts* set, const b3amd_transform* d_xfs,
                                            uint32_t* d_overlap) {
    if (!set || (set->n && (!d_xfs || !d_overlap))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (set->n == 0) return B3AMD_OK;
    int rc;
    MeshScratch sc;
    if ((rc = mesh_scratch(set->ctx, set->n, &sc))) return rc;
    return enqueue_mesh_contacts_overlap(set, reinterpret_cast<const float*>(d_xfs), d_overlap, sc);
}

int b3amd_mesh_contacts_synchronize(b3amd_mesh_contacts* set, const b3amd_transform* xfs, const float* centers0) {
    if (!set || (set->n && (!xfs || !centers0))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    b3amd_context* ctx = set->ctx;
    const uint32_t n = set->n;
    if (n == 0) return B3AMD_OK;
    int rc;
    MeshScratch sc;
    if ((rc = mesh_scratch(ctx, n, &sc))) return rc;
    hipStream_t s = ctx->stream;
    HIP_TRY(hipMemcpyAsync(sc.xfs, xfs, 96ull * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(sc.centers, centers0, 12ull * n, hipMemcpyHostToDevice, s));
    if ((rc = enqueue_mesh_contacts_sync(set, sc.xfs, sc.centers, sc))) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    return ctx->resolve_timings();
}

int b3amd_mesh_contacts_synchronize_device(b3amd_mesh_contacts* set, const b3amd_transform* d_xfs,
                                           const float* d_centers0) {
    if (!set || (set->n && (!d_xfs || !d_centers0))) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (set->n == 0) return B3AMD_OK;
    int rc;
    MeshScratch sc;
    if ((rc = mesh_scratch(set->ctx, set->n, &sc))) return rc;
    return enqueue_mesh_contacts_sync(set, reinterpret_cast<const float*>(d_xfs), d_centers0, sc);
}

int b3amd_mesh_contacts_read(b3amd_mesh_contacts* set, uint32_t* triangleCounts, uint32_t* triangleLists,
                             b3amd_convex_cache* caches, float* fatAabbs) {
    if (!set) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    const uint64_t n = set->n, cap = kMeshTriangleCap;
    if (n == 0) return B3AMD_OK;
    hipStream_t s = set->ctx->stream;
    if (triangleCounts) HIP_TRY(hipMemcpyAsync(triangleCounts, set->dTriCounts, 4 * n, hipMemcpyDeviceToHost, s));
    if (triangleLists) HIP_TRY(hipMemcpyAsync(triangleLists, set->dTriLists, 4 * n * cap, hipMemcpyDeviceToHost, s));
    if (caches) HIP_TRY(hipMemcpyAsync(caches, set->dCaches, sizeof(ConvexCache) * n * cap, hipMemcpyDeviceToHost, s));
    if (fatAabbs) HIP_TRY(hipMemcpyAsync(fatAabbs, set->dFat, 24 * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return B3AMD_OK;
}

int b3amd_synchronize(b3amd_context* ctx) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->resolve_timings();
}

int b3amd_device_alloc(b3amd_context* ctx, uint64_t bytes, void** ptr) {
    if (!ctx || !ptr) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    return ctx->alloc(bytes, ptr);
}

int b3amd_cache_alloc(b3amd_context* ctx, uint32_t count, b3amd_convex_cache** d_caches) {
    if (!ctx || !d_caches) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    void* p = nullptr;
    int rc = ctx->alloc(36ull * (count ? count : 1), &p);
    if (rc) return rc;
    if (launch_cache_reset(static_cast<ConvexCache*>(p), count, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
        ctx->release(p);
        return fail(B3AMD_ERR_DEVICE, "cache reset failed");
    }
    *d_caches = static_cast<b3amd_convex_cache*>(p);
    return B3AMD_OK;
}

int b3amd_cache_reset(b3amd_context* ctx, b3amd_convex_cache* d_caches, uint32_t first, uint32_t count) {
    if (!ctx || (count && !d_caches)) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if ((uint64_t)first + count > 0xFFFFFFFFull) return fail(B3AMD_ERR_ARGUMENT, "range beyond 2^32 caches");
    HIP_TRY(launch_cache_reset(reinterpret_cast<ConvexCache*>(d_caches) + first, count, ctx->stream));
    return B3AMD_OK;
}

int b3amd_cache_free(b3amd_context* ctx, b3amd_convex_cache* d_caches) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    if (!d_caches) return B3AMD_OK;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->release(d_caches);
}

int b3amd_collide_shape_and_shape(b3amd_context* ctx, b3amd_manifold* manifold, const b3amd_transform* xfA,
                                  uint32_t shapeA, const b3amd_transform* xfB, uint32_t shapeB,
                                  const b3amd_quat* sweeps, b3amd_convex_cache* cache) {
    if (!ctx || !manifold || !xfA || !xfB) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    if (sweeps && !cache) return fail(B3AMD_ERR_ARGUMENT, "sweeps without a cache");
    const uint32_t pair[2] = {shapeA, shapeB};
    b3amd_transform xf[2];
    std::memcpy(&xf[0], xfA, sizeof(b3amd_transform));
    std::memcpy(&xf[1], xfB, sizeof(b3amd_transform));
    return b3amd_collide(ctx, 1, pair, xf, sweeps, cache, manifold);
}

int b3amd_device_free(b3amd_context* ctx, void* ptr) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return ctx->release(ptr);
}

int b3amd_copy_to_device(b3amd_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || !dst || !src) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return B3AMD_OK;
}

int b3amd_copy_to_host(b3amd_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || !dst || !src) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return B3AMD_OK;
}

int b3amd_copy_on_device(b3amd_context* ctx, void* dst, const void* src, uint64_t bytes) {
    if (!ctx || !dst || !src) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return B3AMD_OK;
}

int b3amd_set_profiling(b3amd_context* ctx, int enabled) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    int rc = ctx->resolve_timings();
    ctx->profiling = enabled != 0;
    return rc;
}

int b3amd_kernel_stats(b3amd_context* ctx, const char* name, uint32_t* launches, double* total_ms) {
    if (!ctx || !name || !launches || !total_ms) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    int rc = ctx->resolve_timings();
    if (rc) return rc;
    auto it = ctx->stats.find(name);
    *launches = it == ctx->stats.end() ? 0 : it->second.launches;
    *total_ms = it == ctx->stats.end() ? 0.0 : it->second.total_ms;
    return B3AMD_OK;
}

int b3amd_reset_kernel_stats(b3amd_context* ctx) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    int rc = ctx->resolve_timings();
    ctx->stats.clear();
    return rc;
}

int b3amd_graph_create(b3amd_context* ctx, uint32_t n, const uint32_t* d_pairs, const b3amd_transform* d_xfs,
                       const b3amd_quat* d_sweeps, b3amd_convex_cache* d_caches, b3amd_manifold* d_out,
                       b3amd_graph** out) {
    if (!out) return fail(B3AMD_ERR_ARGUMENT, "out is null");
    *out = nullptr;
    if (!ctx || !n || !d_pairs || !d_xfs || !d_out) return fail(B3AMD_ERR_ARGUMENT, "null or empty batch");
    if (d_sweeps && !d_caches) return fail(B3AMD_ERR_ARGUMENT, "sweeps without caches");
    if (!ctx->dPool || !ctx->dShapes) return fail(B3AMD_ERR_ARGUMENT, "no hull pool or shape table uploaded");
    b3amd_graph* g = new b3amd_graph();
    g->ctx = ctx;
    g->generation = ctx->generation;
    int rc;
    if ((rc = ctx->alloc(sizeof(uint32_t) * scratch_words(n), (void**)&g->scratch.stage)) ||
        (rc = ctx->alloc(sizeof(SatResult) * (uint64_t)n, (void**)&g->scratch.sat)) ||
        (rc = ctx->alloc(4 * kCountWords, (void**)&g->scratch.counter))) {
        ctx->release(g->scratch.stage);
        ctx->release(g->scratch.sat);
        delete g;
        return rc;
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const bool profiling = ctx->profiling;
    ctx->profiling = false;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = enqueue_batch(ctx, n, d_pairs, reinterpret_cast<const float*>(d_xfs),
                           reinterpret_cast<const float*>(d_sweeps), reinterpret_cast<ConvexCache*>(d_caches),
                           reinterpret_cast<Manifold*>(d_out), g->scratch, nullptr);
        hipError_t e2 = hipStreamEndCapture(ctx->stream, &g->graph);
        if (rc == B3AMD_OK && e2 != hipSuccess) rc = fail(B3AMD_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e2));
        if (rc == B3AMD_OK) {
            e2 = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
            if (e2 != hipSuccess) rc = fail(B3AMD_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e2));
        }
    } else {
        rc = fail(B3AMD_ERR_DEVICE, "hipStreamBeginCapture: %s", hipGetErrorString(e));
    }
    ctx->profiling = profiling;
    if (rc != B3AMD_OK) {
        if (g->exec) hipGraphExecDestroy(g->exec);
        if (g->graph) hipGraphDestroy(g->graph);
        ctx->release(g->scratch.stage);
        ctx->release(g->scratch.sat);
        ctx->release(g->scratch.counter);
        delete g;
        return rc;
    }
    ctx->graphs.push_back(g);
    *out = g;
    return B3AMD_OK;
}

int b3amd_graph_launch(b3amd_graph* g) {
    if (!g) return fail(B3AMD_ERR_ARGUMENT, "null graph");
    if (g->generation != g->ctx->generation)
        return fail(B3AMD_ERR_ARGUMENT, "graph captured before the hull pool or shape table was replaced");
    HIP_TRY(hipGraphLaunch(g->exec, g->ctx->stream));
    return B3AMD_OK;
}

int b3amd_graph_destroy(b3amd_graph* g) {
    if (!g) return B3AMD_OK;
    b3amd_context* ctx = g->ctx;
    hipStreamSynchronize(ctx->stream);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    ctx->release(g->scratch.stage);
    ctx->release(g->scratch.sat);
    ctx->release(g->scratch.counter);
    for (size_t i = 0; i < ctx->graphs.size(); ++i)
        if (ctx->graphs[i] == g) {
            ctx->graphs.erase(ctx->graphs.begin() + (long)i);
            break;
        }
    delete g;
    return B3AMD_OK;
}

int b3amd_set_counting(b3amd_context* ctx, int on) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    ctx->counting = on != 0;
    return B3AMD_OK;
}

int b3amd_get_counters(b3amd_context* ctx, b3amd_counters* out) {
    if (!ctx || !out) return fail(B3AMD_ERR_ARGUMENT, "null argument");
    static_assert(sizeof(b3amd_counters) == sizeof(Counters), "counter layout");
    Counters slots[kCounterSlots];
    HIP_TRY(hipMemcpyAsync(slots, ctx->dCounters, sizeof(slots), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    Counters t{};
    for (const Counters& c : slots) {
        t.convexCalls += c.convexCalls;
        t.convexCacheHits += c.convexCacheHits;
        t.gjkCalls += c.gjkCalls;
        t.gjkIters += c.gjkIters;
        t.gjkMaxIters = c.gjkMaxIters > t.gjkMaxIters ? c.gjkMaxIters : t.gjkMaxIters;
        t.gjkCacheHits += c.gjkCacheHits;
    }
    std::memcpy(out, &t, sizeof(t));
    return B3AMD_OK;
}

int b3amd_reset_counters(b3amd_context* ctx) {
    if (!ctx) return fail(B3AMD_ERR_ARGUMENT, "null context");
    HIP_TRY(launch_zero_words(reinterpret_cast<uint32_t*>(ctx->dCounters), sizeof(Counters) * kCounterSlots / 4,
                              ctx->stream));
    return B3AMD_OK;
}

}  // extern "C"
