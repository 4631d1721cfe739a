// Persistent SAT kernel for hull-hull pairs on gfx950: the separation queries of
// b3CollideHulls (face query of hull 1, face query of hull 2, edge query; sat.cpp:33-173,
// collide_hulls.cpp:206-289 up to the contact build), one 64-lane wave per pair.
//
// A grid sized to the resident occupancy takes the pairs the cache kernel marked: each
// wave a static interleave of the batch, then chunks of the last sixteenth from a device
// counter (one atomic per chunk; per-pair atomics on one address serialise).  Per pair:
//   * face queries: the queried hull's planes are transformed in registers (face lane +
//     64k in slot k); the other hull's vertices are staged in LDS and read by every lane
//     at once (LDS broadcast); each lane keeps the maximum support value, from which the
//     reference's separation follows exactly;
//   * edge query: hull 1's unique edges, rotated into hull 2's frame, sit in up to three
//     register slots per lane as the bounding caps of their Gauss-map arcs;
//     hull 2's negated Gauss map and arc caps, stored per hull at upload (device_hull.h),
//     are staged in LDS 64 edges at a time and broadcast one edge per step.  A cheap
//     cap-overlap test selects the (i, j) worth the exact b3IsMinkowskiFace, which runs 64
//     candidates at a time from an LDS ring; the pairs that pass go to a second ring and
//     are projected 64 at a time by all lanes together, so b3Project's square roots and
//     division run converged.
// Ties resolve to the lowest index in the reference's scan order (sat.cpp:47,158), so the
// result is bit-identical to the sequential loops whatever the processing order.
#include <hip/hip_runtime.h>

#include "b3_math.h"
#include "contact_build.h"
#include "device_common.h"
#include "device_hull.h"
#include "kernels.h"

namespace b3amd {

// Read-only inputs through the constant address space: a load whose address is the same
// on every lane becomes a scalar load into SGPRs; per-lane addresses stay vector loads.
typedef const __attribute__((address_space(4))) uint32_t* CU32;
typedef const __attribute__((address_space(4))) float* CF32;

// float4 units through a constant-space float pointer (vector types have no copy from the
// constant address space; four adjacent loads merge into one dwordx4 load).
struct CF4 {
    CF32 p;
    __device__ __forceinline__ float4 operator[](uint32_t i) const {
        const CF32 q = p + 4u * i;
        return make_float4(q[0], q[1], q[2], q[3]);
    }
    __device__ __forceinline__ CF4 operator+(uint32_t i) const { return CF4{p + 4u * i}; }
};

__device__ __forceinline__ Xf load_xf_u(CF32 p) {
    Xf x;
    x.R.c0 = V3{p[0], p[1], p[2]};
    x.R.c1 = V3{p[3], p[4], p[5]};
    x.R.c2 = V3{p[6], p[7], p[8]};
    x.p = V3{p[9], p[10], p[11]};
    return x;
}

// A hull record (device_hull.h) addressed through a wave-uniform base.
struct UHull {
    CF4 base;
    uint32_t V, F, HE;
    __device__ static UHull at_unit(CF4 pool, uint32_t unit) {
        UHull r;
        r.base = pool + unit;
        const CU32 w = (CU32)r.base.p;
        const uint32_t x = w[0], y = w[1];
        r.V = x & 0xFFFFu;
        r.F = x >> 16;
        r.HE = y & 0xFFFFu;
        return r;
    }
    __device__ HullLayout layout() const { return HullLayout{V, F, HE}; }
};

constexpr uint32_t kStageEdges = 96;   // hull-2 edges staged in LDS at a time (a multiple of 32)
constexpr uint32_t kStageVerts = 64;   // hull vertices staged in LDS at a time by the face queries
constexpr uint32_t kLdsVerts = 64;     // hull 1 transformed into hull 2's frame in LDS when it has
constexpr uint32_t kLdsFaces = 128;    // at most this many vertices and faces (and unique edges,
constexpr uint32_t kLdsEdges = 192;    // whose records are then read from LDS too)
constexpr uint32_t kHitCap = 128;      // rings of pending pairs: fewer than 64 + 64 at any time
constexpr uint32_t kCandCap = 128;

// LDS of one wave (the block is one wave).
struct SatStage {
    union {
        float4 verts[kStageVerts];           // face queries: 64 of the other hull's vertices
        float4 exact[3 * kStageEdges];       // edge query: a chunk of hull 2's Gauss map
    } u;
    float4 tv1[kLdsVerts];  // edge query: hull 1's vertices, b3Mul(xf, v)
    float4 tn1[kLdsFaces];  // edge query: hull 1's face normals, b3Mul(xf.rotation, n)
    uint32_t ue1[kLdsEdges];  // edge query: hull 1's unique-edge records
    uint32_t hits[kHitCap];    // pending (i << 16 | j) pairs that passed the Gauss-map test
    uint32_t cands[kCandCap];  // pending pairs whose arc caps overlap (queue encoding)
};

// The block is a single wave: __syncthreads() orders its LDS writes and reads.
__device__ __forceinline__ void stage_vertices(const UHull& h, uint32_t v0, uint32_t vn, SatStage& st,
                                               uint32_t lane) {
    __syncthreads();
    if (lane < vn) st.u.verts[lane] = h.base[2 + v0 + lane];
    __syncthreads();
}

// One face query (sat.cpp:33-58) over faces fBase + lane + 64k of hull A: planes transformed
// by xf (T9, geometry.h:77-81) against hull B's vertices, staged in LDS 64 at a time.  The
// reference takes the first vertex of maximum dot(-n, v) (hull.inl:31-45) and returns
// b3Distance(v, plane) = dot(n, v) - d (geometry.h:90-93).  Negating n negates every FP32
// product and sum exactly, so dot(n, v) = -max and the separation is -max - d whichever
// tied vertex is first: the search keeps only the running maximum, started from vertex 0
// as the reference's is.  Each lane's best face and its separation fold into (best, key),
// earlier faces winning ties.
template <int KF>
__device__ __forceinline__ void face_query_l(const Xf& xf, const UHull& hA, uint32_t fBase, const UHull& hB,
                                             SatStage& st, uint32_t lane, float& best, uint32_t& key) {
    Plane pl[KF];
    V3 d[KF];
    float mp[KF];
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        const uint32_t f = fBase + lane + 64u * k;
        const float4 q = hA.base[2 + hA.V + (f < hA.F ? f : 0u)];
        pl[k] = mul(xf, Plane{V3{q.x, q.y, q.z}, q.w});
        d[k] = -pl[k].n;
    }
    for (uint32_t v0 = 0; v0 < hB.V; v0 += kStageVerts) {
        const uint32_t vn = hB.V - v0 < kStageVerts ? hB.V - v0 : kStageVerts;
        stage_vertices(hB, v0, vn, st, lane);
        uint32_t v = 0;
        if (v0 == 0u) {
            const V3 w0 = ld3(st.u.verts[0]);
#pragma unroll
            for (int k = 0; k < KF; ++k) mp[k] = dot(d[k], w0);
            v = 1;
        }
#pragma unroll 4
        for (; v < vn; ++v) {
            const V3 w = ld3(st.u.verts[v]);
#pragma unroll
            for (int k = 0; k < KF; ++k) mp[k] = fmaxf(mp[k], dot(d[k], w));
        }
    }
#pragma unroll
    for (int k = 0; k < KF; ++k) {
        const uint32_t f = fBase + lane + 64u * k;
        if (f < hA.F) {
            const float s = -mp[k] - pl[k].d;
            if (s > best) {
                best = s;
                key = f;
            }
        }
    }
}

__device__ __forceinline__ void face_query_any(const Xf& xf, const UHull& hA, const UHull& hB, SatStage& st,
                                               uint32_t lane, float& best, uint32_t& key) {
    best = -kMaxFloat;
    key = 0;
    for (uint32_t fBase = 0; fBase < hA.F; fBase += 128u) {
        if (hA.F - fBase <= 64u) face_query_l<1>(xf, hA, fBase, hB, st, lane, best, key);
        else face_query_l<2>(xf, hA, fBase, hB, st, lane, best, key);
    }
    wave_argmax(best, key);
}

// Projects `count` queued hits (lane l takes entry head + l) with b3Project (sat.cpp:73-108)
// and folds them into (best, key); j is a hull-2 edge of the staged chunk starting at c0.
__device__ __forceinline__ void drain_hits(const Xf& xf, V3 C1, const UHull& h1, CU32 ue1, const UHull& h2,
                                           CU32 ue2, const SatStage& st, bool lds1, uint32_t c0, uint32_t head,
                                           uint32_t count, uint32_t lane, float& best, uint32_t& key) {
    if (lane < count) {
        const uint32_t e = st.hits[(head + lane) & (kHitCap - 1u)];
        const uint32_t i = e >> 16, j = e & 0xFFFFu;
        const uint32_t u = lds1 ? st.ue1[i] : ue1[i];
        const V3 P1 = lds1 ? ld3(st.tv1[u & 0xFFu]) : mul(xf, ld3(h1.base[2 + (u & 0xFFu)]));
        const V3 Q1 = lds1 ? ld3(st.tv1[(u >> 8) & 0xFFu]) : mul(xf, ld3(h1.base[2 + ((u >> 8) & 0xFFu)]));
        const V3 E1 = Q1 - P1;
        const float4 r1 = st.u.exact[3u * (j - c0) + 1u], r2 = st.u.exact[3u * (j - c0) + 2u];
        const V3 E2 = -V3{r1.z, r1.w, r2.x};
        const V3 P2{r2.y, r2.z, r2.w};
        const float s = project_edges(P1, E1, length(E1), P1 - C1, P2, E2, length(E2));
        if (s > best || (s == best && e < key)) {
            best = s;
            key = e;
        }
    }
}

// Hull 1's operands of b3IsMinkowskiFace for unique edge i in hull 2's frame (sat.cpp:123-146):
// A = U1, B = V1, nE = -E1; from the transformed hull in LDS when it is there (the same
// FP32 operations, done once per pair).
__device__ __forceinline__ void edge1_operands(const Xf& xf, const UHull& h1, CU32 ue1, const SatStage& st, bool lds1,
                                               uint32_t i, V3& A, V3& B, V3& nE) {
    if (lds1) {
        const uint32_t u = st.ue1[i];
        const V3 P1 = ld3(st.tv1[u & 0xFFu]);
        const V3 Q1 = ld3(st.tv1[(u >> 8) & 0xFFu]);
        A = ld3(st.tn1[(u >> 16) & 0xFFu]);
        B = ld3(st.tn1[u >> 24]);
        nE = -(Q1 - P1);
    } else {
        const uint32_t u = ue1[i];
        const V3 P1 = mul(xf, ld3(h1.base[2 + (u & 0xFFu)]));
        const V3 Q1 = mul(xf, ld3(h1.base[2 + ((u >> 8) & 0xFFu)]));
        A = mul(xf.R, ld3(h1.base[2 + h1.V + ((u >> 16) & 0xFFu)]));
        B = mul(xf.R, ld3(h1.base[2 + h1.V + (u >> 24)]));
        nE = -(Q1 - P1);
    }
}

// Runs b3IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2) (sat.cpp:60-71) on `count` queued
// candidates (lane l takes entry head + l), hull 1's operands recomputed exactly as the
// reference computes them, and queues the pairs that pass for projection.
__device__ __forceinline__ void drain_candidates(const Xf& xf, V3 C1, const UHull& h1, CU32 ue1, const UHull& h2,
                                                 CU32 ue2, SatStage& st, bool lds1, uint32_t c0, uint32_t chead,
                                                 uint32_t count, uint32_t lane, uint32_t& head, uint32_t& tail,
                                                 float& best, uint32_t& key) {
    uint32_t e = 0;
    bool pass = false;
    if (lane < count) {
        // The queue's entry (j << 16 | pair-of-row-blocks << 6 | lane half << 5 | sign bit)
        // back to the edge pair (i << 16 | j), i.e. the reference's scan order.
        const uint32_t c = st.cands[(chead + lane) & (kCandCap - 1u)];
        const uint32_t q = 31u - (c & 31u), rr = q & 15u;
        const uint32_t i = 64u * ((c >> 6) & 1023u) + 32u * (q >> 4) + 8u * (rr >> 2) + 4u * ((c >> 5) & 1u) + (rr & 3u);
        const uint32_t j = c >> 16;
        e = (i << 16) | j;
        V3 A, B, nE;
        edge1_operands(xf, h1, ue1, st, lds1, i, A, B, nE);
        const float4* r = st.u.exact + 3u * (j - c0);
        const float4 r0 = r[0], r1 = r[1];
        const float r2x = r[2].x;
        const V3 C{r0.x, r0.y, r0.z}, Dn{r0.w, r1.x, r1.y}, DxC{r1.z, r1.w, r2x};
        const float CBA = dot(C, nE);
        const float DBA = dot(Dn, nE);
        const float ADC = dot(A, DxC);
        const float BDC = dot(B, DxC);
        pass = (CBA * DBA < 0.0f) && (ADC * BDC < 0.0f) && (CBA * BDC > 0.0f);
    }
    const uint64_t m = __ballot(pass);
    if (pass) st.hits[(tail + mbcnt(m)) & (kHitCap - 1u)] = e;
    tail += (uint32_t)__popcll(m);
    if (tail - head >= 64u) {
        __syncthreads();
        drain_hits(xf, C1, h1, ue1, h2, ue2, st, lds1, c0, head, 64u, lane, best, key);
        head += 64u;
    }
}

// Edge query (sat.cpp:110-173).  The bounding caps of the two hulls' Gauss-map arcs are
// tested on the matrix cores first: with hull 1's caps rotated into hull 2's frame as rows
// (M1, -cos a1, sin a1) and hull 2's stored caps as columns (-m2, cos a2, sin a2), one
// v_mfma_f32_32x32x8_f16 gives, for 32 x 32 edge pairs, sin a1 sin a2 - cos a1 cos a2 +
// M1.M2 + 2^-8: non-negative iff the caps may meet.  The operands are rounded to halves
// (relative 2^-11 each) and the products are exact in the FP32 accumulation, so with every
// term at most 1 in magnitude the sum is off by at most 3 * 2^-10 plus flushed half
// denormals and FP32 accumulation (under 4e-4 together) < 2^-8: a pair whose caps meet is
// never rejected.  The last two of the eight product terms mask padding: a row or column
// past the last edge adds -4, which no cap term can outweigh.  Caps contain the arcs with a
// margin, so a pair they reject cannot pass the exact test; the ~6 % that pass go through
// an LDS ring and are tested 64 at a time, converged, with the reference's operands and
// operation order, so the set of hits is unchanged.
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef float mfma16v __attribute__((ext_vector_type(16)));
typedef const __attribute__((address_space(4))) uint16_t* CU16;

constexpr uint32_t kRowBlocks = 4;  // hull-1 edge blocks of 32 held as operands per pass

__device__ __forceinline__ half4v half4_of(uint32_t lo, uint32_t hi) {
    return __builtin_bit_cast(half4v, make_uint2(lo, hi));
}

// Row block b of the cap test: lane l holds row i = 32b + l % 32, terms 4 (l / 32) .. + 3.
__device__ __forceinline__ half4v cap_rows(const Xf& xf, CF4 cap1, CU16 f1, uint32_t E1n, uint32_t b,
                                           uint32_t lane) {
    const uint32_t i = 32u * b + (lane & 31u);
    const bool live = i < E1n;
    if (lane < 32u) {
        if (!live) return half4_of(0u, 0u);
        const float4 cap = cap1[i];
        const V3 M = mul(xf.R, V3{-cap.x, -cap.y, -cap.z});
        return half4v{(_Float16)M.x, (_Float16)M.y, (_Float16)M.z, (_Float16)(-cap.w)};
    }
    const uint32_t s = live ? (uint32_t)f1[160u * (i >> 5) + 128u + (i & 31u)] : 0u;
    // (sin a1, 2^-8, 0 or -4 for a padding row, 1)
    return half4_of(s | (0x1C00u << 16), (live ? 0u : 0xC400u) | (0x3C00u << 16));
}

// Column block c of the cap test from hull 2's stored halves: lane l holds column
// j = 32c + l % 32, terms 4 (l / 32) .. + 3: (-m2, cos a2) or (sin a2, 1, 1, 0 or -4).
__device__ __forceinline__ half4v cap_cols(CU16 f2, uint32_t E2, uint32_t c, uint32_t lane) {
    const uint32_t j = 32u * c + (lane & 31u);
    const CU32 w = (CU32)(f2 + 160u * c);
    if (lane < 32u) return half4_of(w[2u * (lane & 31u)], w[2u * (lane & 31u) + 1u]);
    const uint32_t s = (uint32_t)f2[160u * c + 128u + (lane & 31u)];
    return half4_of(s | (0x3C00u << 16), 0x3C00u | ((j < E2 ? 0u : 0xC400u) << 16));
}

// Sign bits of a 32 x 32 block into acc: acc = acc << 16 | signs, register r at bit 15 - r.
__device__ __forceinline__ uint32_t shift_signs(uint32_t acc, const mfma16v& d) {
#pragma unroll
    for (int r = 0; r < 16; ++r) acc = __builtin_amdgcn_alignbit(acc, __float_as_uint(d[r]), 31);
    return acc;
}

__device__ __forceinline__ void edge_query_m(const Xf& xf, V3 C1, const UHull& h1, const UHull& h2, SatStage& st,
                                             uint32_t lane, float& best, uint32_t& key) {
    const uint32_t E1n = h1.HE >> 1;
    const HullLayout L1 = h1.layout(), L2 = h2.layout();
    const CU32 ue1 = (CU32)(h1.base + L1.uedgeOff()).p;
    const CU32 ue2 = (CU32)(h2.base + L2.uedgeOff()).p;
    const CF4 cap1 = h1.base + L1.capOff();
    const CU16 f1 = (CU16)(h1.base + L1.filterOff()).p;
    const CF4 g2 = h2.base + L2.gaussOff();
    const CU16 f2 = (CU16)(h2.base + L2.filterOff()).p;
    const uint32_t E2 = h2.HE >> 1;
    const uint32_t E2p = L2.gaussEdges();  // staged in chunks; edges past E2 are masked
    const uint32_t R1b = (E1n + 31u) / 32u;
    // Hull 1 in hull 2's frame, once per pair, for the exact tests and projections, and the
    // first chunk of hull 2's Gauss map: their loads are issued together, before one
    // barrier, so that the wave waits for memory once.
    const bool lds1 = h1.V <= kLdsVerts && h1.F <= kLdsFaces && E1n <= kLdsEdges;
    const uint32_t cn0 = E2p < kStageEdges ? E2p : kStageEdges;
    constexpr uint32_t kPreExact = (3u * kStageEdges + 63u) / 64u;
    float4 preExact[kPreExact];
    float4 preV = make_float4(0.f, 0.f, 0.f, 0.f), preN0 = preV, preN1 = preV;
    uint32_t preE[kLdsEdges / 64u];
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < kPreExact; ++q)
        if (lane + 64u * q < 3u * cn0) preExact[q] = g2[lane + 64u * q];
    if (lds1) {
#pragma unroll
        for (uint32_t q = 0; q < kLdsEdges / 64u; ++q)
            if (lane + 64u * q < E1n) preE[q] = ue1[lane + 64u * q];
        if (lane < h1.V) preV = h1.base[2 + lane];
        if (lane < h1.F) preN0 = h1.base[2 + h1.V + lane];
        if (lane + 64u < h1.F) preN1 = h1.base[2 + h1.V + lane + 64u];
        if (lane < h1.V) {
            const V3 p = mul(xf, ld3(preV));
            st.tv1[lane] = make_float4(p.x, p.y, p.z, 0.0f);
        }
        if (lane < h1.F) {
            const V3 n = mul(xf.R, ld3(preN0));
            st.tn1[lane] = make_float4(n.x, n.y, n.z, 0.0f);
        }
        if (lane + 64u < h1.F) {
            const V3 n = mul(xf.R, ld3(preN1));
            st.tn1[lane + 64u] = make_float4(n.x, n.y, n.z, 0.0f);
        }
#pragma unroll
        for (uint32_t q = 0; q < kLdsEdges / 64u; ++q)
            if (lane + 64u * q < E1n) st.ue1[lane + 64u * q] = preE[q];
    }
#pragma unroll
    for (uint32_t q = 0; q < kPreExact; ++q)
        if (lane + 64u * q < 3u * cn0) st.u.exact[lane + 64u * q] = preExact[q];
    __syncthreads();
    uint32_t head = 0, tail = 0, chead = 0, ctail = 0;
    for (uint32_t b0 = 0; b0 < R1b; b0 += kRowBlocks) {
        const uint32_t nb = R1b - b0 < kRowBlocks ? R1b - b0 : kRowBlocks;
        half4v A[kRowBlocks];
#pragma unroll
        for (uint32_t b = 0; b < kRowBlocks; ++b)
            if (b < nb) A[b] = cap_rows(xf, cap1, f1, E1n, b0 + b, lane);
        for (uint32_t c0 = 0; c0 < E2p; c0 += kStageEdges) {
            const uint32_t cn = E2p - c0 < kStageEdges ? E2p - c0 : kStageEdges;  // a multiple of 4
            if (b0 != 0u || c0 != 0u) {  // the first chunk was staged above
                __syncthreads();
                for (uint32_t u = lane; u < 3u * cn; u += 64u) st.u.exact[u] = g2[3u * c0 + u];
                __syncthreads();
            }
            half4v Bn = cap_cols(f2, E2, c0 / 32u, lane);
            for (uint32_t j0 = c0; j0 < c0 + cn; j0 += 32u) {
                const half4v B = Bn;
                if (j0 + 32u < c0 + cn) Bn = cap_cols(f2, E2, j0 / 32u + 1u, lane);
#pragma unroll
                for (uint32_t b = 0; b < kRowBlocks; b += 2u) {
                    if (b >= nb) break;
                    // Two row blocks per queue pass: bit 31 - r of acc is register r of the
                    // first block's result, bit 15 - r of the second's (all set, i.e.
                    // rejected, when there is no second block).
                    const mfma16v z = {};
                    uint32_t acc = shift_signs(0xFFFFFFFFu, __builtin_amdgcn_mfma_f32_32x32x8f16(A[b], B, z, 0, 0, 0));
                    if (b + 1u < nb)
                        acc = shift_signs(acc, __builtin_amdgcn_mfma_f32_32x32x8f16(A[b + 1u], B, z, 0, 0, 0));
                    else
                        acc = (acc << 16) | 0xFFFFu;
                    uint32_t cm = ~acc;
                    // Queue the candidates: each round appends one per lane that has any, as
                    // its column, the pass and the bit; the drain decodes the row.
                    const uint32_t base = ((j0 + (lane & 31u)) << 16) | (((b0 + b) >> 1) << 6) | ((lane >> 5) << 5);
                    for (;;) {
                        const uint64_t m = __ballot(cm != 0u);
                        if (m == 0) break;
                        if (cm) {
                            const uint32_t e = base | (uint32_t)__builtin_ctz(cm);
                            cm &= cm - 1u;
                            st.cands[(ctail + mbcnt(m)) & (kCandCap - 1u)] = e;
                        }
                        ctail += (uint32_t)__popcll(m);
                        if (ctail - chead >= 64u) {
                            __syncthreads();
                            drain_candidates(xf, C1, h1, ue1, h2, ue2, st, lds1, c0, chead, 64u, lane, head, tail, best,
                                             key);
                            chead += 64u;
                        }
                    }
                }
            }
            // Test and project what is queued before the staged chunk is replaced.
            __syncthreads();
            while (ctail != chead) {
                const uint32_t cnt = ctail - chead < 64u ? ctail - chead : 64u;
                drain_candidates(xf, C1, h1, ue1, h2, ue2, st, lds1, c0, chead, cnt, lane, head, tail, best, key);
                chead += cnt;
            }
            __syncthreads();
            while (tail != head) {
                const uint32_t cnt = tail - head < 64u ? tail - head : 64u;
                drain_hits(xf, C1, h1, ue1, h2, ue2, st, lds1, c0, head, cnt, lane, best, key);
                head += cnt;
            }
        }
    }
}

// b3CollideHulls' queries for one pair with the early exits on totalRadius
// (collide_hulls.cpp:206-289).
__device__ __forceinline__ void sat_pair(const SatArgs& a, uint32_t p, SatStage& st, uint32_t lane) {
    // The list kernel stored the record units and the radii in the result slot.
    const CU32 pre = (const CU32)(a.out + p);
    const uint32_t ua = pre[0], ub = pre[2];
    const float totalRadius = __uint_as_float(pre[1]) + __uint_as_float(pre[3]);
    const Xf xf1 = load_xf_u((const CF32)a.xf_ptr(p, 0)), xf2 = load_xf_u((const CF32)a.xf_ptr(p, 1));
    const CF4 pool{(CF32)a.pool};
    const UHull h1 = UHull::at_unit(pool, ua);
    const UHull h2 = UHull::at_unit(pool, ub);

    SatResult r;
    r.f2Index = 0;
    r.f2Sep = -kMaxFloat;
    r.eIndex1 = 0;
    r.eIndex2 = 0;
    r.eSep = -kMaxFloat;
    const Xf xf = mulT(xf2, xf1);
    face_query_any(xf, h1, h2, st, lane, r.f1Sep, r.f1Index);
    if (r.f1Sep > totalRadius) {
        r.stage = kStageSepFace1;
        if (lane == 0) a.out[p] = r;
        return;
    }
    face_query_any(mulT(xf1, xf2), h2, h1, st, lane, r.f2Sep, r.f2Index);
    if (r.f2Sep > totalRadius) {
        r.stage = kStageSepFace2;
        if (lane == 0) a.out[p] = r;
        return;
    }
    const V3 C1 = mul(xf, ld3(h1.base[1]));
    float best = -kMaxFloat;
    uint32_t key = 0;
    edge_query_m(xf, C1, h1, h2, st, lane, best, key);
    wave_argmax(best, key);
    r.eIndex1 = 2 * (key >> 16);
    r.eIndex2 = 2 * (key & 0xFFFFu);
    r.eSep = best;
    r.stage = r.eSep > totalRadius ? kStageSepEdge : kStageOverlap;
    if (lane == 0) a.out[p] = r;
}

// The pairs the cache kernels listed for a SAT: a static interleave up to staticEnd, then
// chunks from the tail counter (every wave leaves once the counter passes the list).
// W is the register budget: the compiler fits the kernel into 512 / W VGPRs so that W
// waves share a SIMD.
template <int W>
__global__ __launch_bounds__(64, W) void sat_persistent_kernel(SatArgs a) {
    __shared__ SatStage st;
    const uint32_t lane = threadIdx.x;
    const CU32 list = (const CU32)a.list;
    const uint32_t count = *a.listCount < a.n ? *a.listCount : a.n;
    // A short list (fewer than `chunk` pairs per wave) goes out one pair at a time.
    const uint32_t chunk = count >= a.split.chunk * gridDim.x ? a.split.chunk : 1u;
    uint32_t tail = count / 16u;
    if (tail < 2u * chunk * gridDim.x) tail = 2u * chunk * gridDim.x;
    if (tail > count) tail = count;
    const uint32_t staticEnd = count - tail;
    for (uint32_t i = blockIdx.x; i < staticEnd; i += gridDim.x) sat_pair(a, list[i], st, lane);
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(a.split.counter, chunk);
        const uint32_t i0 = staticEnd + (uint32_t)__builtin_amdgcn_readlane((int)t, 0);
        if (i0 >= count) break;
        const uint32_t i1 = count - i0 < chunk ? count : i0 + chunk;
        for (uint32_t i = i0; i < i1; ++i) sat_pair(a, list[i], st, lane);
    }
}

// Zeroes device words with a kernel (a stream operation that graph capture records as a
// kernel node, unlike a memset).
__global__ void zero_words_kernel(uint32_t* words, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) words[i] = 0u;
}

hipError_t launch_zero_words(uint32_t* words, uint32_t count, hipStream_t s) {
    if (count == 0) return hipSuccess;
    zero_words_kernel<<<(count + 255) / 256, 256, 0, s>>>(words, count);
    return hipGetLastError();
}

template <int W>
hipError_t launch_sat_w(SatArgs a, uint32_t wavesPerSimd, uint32_t chunk, hipStream_t s) {
    static int cus = 0, perCu = 0;
    if (cus == 0) {
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, sat_persistent_kernel<W>, 64, 0);
        if (e != hipSuccess) {
            cus = 0;
            return e;
        }
        if (perCu < 1) perCu = 1;
    }
    // A block is one wave; a CU has four SIMDs.
    uint32_t blocksPerCu = (uint32_t)perCu;
    if (4u * wavesPerSimd < blocksPerCu) blocksPerCu = 4u * wavesPerSimd;
    uint64_t blocks = (uint64_t)cus * blocksPerCu;
    if (a.n < blocks) blocks = a.n;
    // The last sixteenth of the list (at least two chunks per wave) goes out dynamically.
    a.split.chunk = chunk;
    sat_persistent_kernel<W><<<(uint32_t)blocks, 64, 0, s>>>(a);
    return hipGetLastError();
}

hipError_t launch_sat_persistent(SatArgs a, const SatTuning& t, hipStream_t s) {
    if (a.n == 0) return hipSuccess;
    const uint32_t chunk = t.tailChunk == 0 ? 8u : (t.tailChunk > 64u ? 64u : t.tailChunk);
    const uint32_t w = t.wavesPerSimd == 0 ? 4u : t.wavesPerSimd;
    // Register budget of the variant: the largest that still holds w waves per SIMD.
    if (w >= 6u) return launch_sat_w<6>(a, w, chunk, s);
    if (w == 5u) return launch_sat_w<5>(a, w, chunk, s);
    return launch_sat_w<4>(a, w, chunk, s);
}

}  // namespace b3amd
