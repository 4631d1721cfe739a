/*
 * TEST INFRASTRUCTURE ONLY - plain-C restatement of the reference contact narrowphase.
 *
 * This file is the oracle: tests, the smoke check and bench.py's CPU-baseline leg use it
 * as the checker.  The product (bounce_amd) never links, loads or calls it.
 *
 * Every function restates one reference function in the same FP32 operation order
 * (compiled with -ffp-contract=off, no fast-math), citing the file:line it follows.
 * Association order is written out explicitly: b3Dot = (x*x' + y*y') + z*z'
 * (vec3.h:158-161); matrix*vector is the column combination (mat33.h:143-146).
 *
 * With -DB3O_COUNT_FLOPS every FP32 + - * / sqrt executed is counted (FL(n)), giving the
 * algorithmic op count per pair used for the roofline.
 *
 * Pinning (what checks each function against something other than itself):
 *   - all functions: bit-exact comparison with the compiled reference (oracle/_ref) on
 *     seeded random hull pairs, boxes and degenerate placements (tests/test_oracle.py);
 *   - golden fixtures written by the compiled reference (tests/golden/);
 *   - face/edge queries: brute-force FP64 SAT separation (the separating-axis theorem);
 *   - face contact: the Box Face Contact known answer (box_face_contact.h:25-49).
 */
#include "b3oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- op counting */
/* Ops are counted per phase: 0 = feature-cache read and rebuilds, 1 = separation queries
 * (face/edge SAT), 2 = contact building (clip, reduce, manifold). */
#ifdef B3O_COUNT_FLOPS
static uint64_t g_flops[3];
static int g_phase;
#define FL(n) (g_flops[g_phase] += (uint64_t)(n))
#define PHASE(p) (g_phase = (p))
uint64_t b3o_flop_count(void) { return g_flops[0] + g_flops[1] + g_flops[2]; }
uint64_t b3o_flop_count_phase(int p) { return (p >= 0 && p < 3) ? g_flops[p] : 0; }
void b3o_flop_reset(void) { g_flops[0] = g_flops[1] = g_flops[2] = 0; g_phase = 0; }
#else
#define FL(n) ((void)0)
#define PHASE(p) ((void)0)
uint64_t b3o_flop_count(void) { return 0; }
uint64_t b3o_flop_count_phase(int p) { (void)p; return 0; }
void b3o_flop_reset(void) {}
#endif

/* ---------------------------------------------------------------- settings.h */
#define LINEAR_SLOP 0.005f          /* settings.h:62 */
#define MAX_FLOAT FLT_MAX           /* settings.h:41 */
#define EPSILON FLT_EPSILON         /* settings.h:42 */
#define NULL_EDGE 0xFFFFFFFFu       /* clip.h:25 */
#define NULL_TRIANGLE 0xFFFFFFFFu   /* manifold.h:25 */
#define MAX_POINTS 4                /* settings.h:81 */
#define POLY_CAP 65536              /* b3StackArray<...,32> grows on the heap; 65536 bounds it here */

typedef b3o_vec3 V3;
typedef b3o_xf XF;
typedef b3o_plane PL;
typedef b3o_quat QT;
typedef struct { V3 c[3]; } M33;

/* ---------------------------------------------------------------- vec3.h */
static V3 v3(float x, float y, float z) { V3 r = {x, y, z}; return r; }
static V3 add(V3 a, V3 b) { FL(3); return v3(a.x + b.x, a.y + b.y, a.z + b.z); }        /* vec3.h:140-143 */
static V3 sub(V3 a, V3 b) { FL(3); return v3(a.x - b.x, a.y - b.y, a.z - b.z); }        /* vec3.h:146-149 */
static V3 neg(V3 a) { return v3(-a.x, -a.y, -a.z); }                                  /* vec3.h:134-137 */
static V3 scale(float s, V3 v) { FL(3); return v3(s * v.x, s * v.y, s * v.z); }         /* vec3.h:128-131 */
static float dot(V3 a, V3 b) { FL(5); return a.x * b.x + a.y * b.y + a.z * b.z; }      /* vec3.h:158-161 */
static V3 cross(V3 a, V3 b) {                                                          /* vec3.h:164-167 */
    FL(9);
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static float length(V3 v) { FL(6); return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }  /* vec3.h:177-180 */
static float length_sq(V3 v) { FL(5); return v.x * v.x + v.y * v.y + v.z * v.z; }      /* vec3.h:183-186 */
static V3 normalize(V3 v) {                                                             /* vec3.h:189-198 */
    float len = length(v);
    if (len > EPSILON) { FL(1); float s = 1.0f / len; return scale(s, v); }
    return v;
}
static float normalize_in_place(V3* v) {                                                /* vec3.h:93-104 */
    FL(6);
    float len = sqrtf(v->x * v->x + v->y * v->y + v->z * v->z);
    if (len > EPSILON) {
        FL(4);
        float s = 1.0f / len;
        v->x *= s; v->y *= s; v->z *= s;
    }
    return len;
}

/* ---------------------------------------------------------------- mat33.h / transform.h */
static V3 mul_mv(const M33* A, V3 v) {                                                  /* mat33.h:143-146 */
    return add(add(scale(v.x, A->c[0]), scale(v.y, A->c[1])), scale(v.z, A->c[2]));
}
static V3 mulT_mv(const M33* A, V3 v) {                                                 /* mat33.h:157-160 */
    return v3(dot(A->c[0], v), dot(A->c[1], v), dot(A->c[2], v));
}
static M33 mulT_mm(const M33* A, const M33* B) {                                        /* mat33.h:163-169 */
    M33 C;
    for (int j = 0; j < 3; ++j) C.c[j] = v3(dot(A->c[0], B->c[j]), dot(A->c[1], B->c[j]), dot(A->c[2], B->c[j]));
    return C;
}
static const M33* rot(const XF* T) { return (const M33*)&T->c[0]; }
static V3 xf_mul(const XF* T, V3 v) { return add(mul_mv(rot(T), v), T->p); }           /* transform.h:74-77 */
static V3 xf_mulT(const XF* T, V3 v) { return mulT_mv(rot(T), sub(v, T->p)); }        /* transform.h:106-110 */
static XF xf_mulT_xf(const XF* A, const XF* B) {                                        /* transform.h:88-94 */
    XF C;
    M33 R = mulT_mm(rot(A), rot(B));
    memcpy(&C.c[0], &R, sizeof(R));
    C.p = mulT_mv(rot(A), sub(B->p, A->p));
    return C;
}

/* ---------------------------------------------------------------- geometry.h */
static PL plane_mul(const XF* T, PL p) {                                                /* geometry.h:77-81 */
    PL r;
    r.n = mul_mv(rot(T), p.n);
    FL(1);
    r.d = p.d + dot(r.n, T->p);
    return r;
}
static float plane_dist(V3 P, PL p) { FL(1); return dot(p.n, P) - p.d; }               /* geometry.h:90-93 */
static V3 closest_on_plane(V3 P, PL p) {                                                /* geometry.h:96-100 */
    float f = plane_dist(P, p);
    return sub(P, scale(f, p.n));
}
static void barycentric2(float out[3], V3 A, V3 B, V3 Q) {                             /* geometry.h:104-116 */
    V3 AB = sub(B, A), QA = sub(A, Q), QB = sub(B, Q);
    float divisor = dot(AB, AB);
    out[0] = dot(QB, AB);
    out[1] = -dot(QA, AB);
    out[2] = divisor;
}

/* ---------------------------------------------------------------- quat.h */
static QT q_mul(QT a, QT b) {                                                           /* quat.h:172-179 */
    FL(28);
    QT r;
    r.x = a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y;
    r.y = a.w * b.y + a.y * b.w + a.z * b.x - a.x * b.z;
    r.z = a.w * b.z + a.z * b.w + a.x * b.y - a.y * b.x;
    r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
    return r;
}
static QT q_conj(QT q) { QT r = {-q.x, -q.y, -q.z, q.w}; return r; }                  /* quat.h:195-198 */

/* ---------------------------------------------------------------- hull.h / hull.inl */
typedef struct {
    V3 centroid;
    uint32_t vertexCount, edgeCount, faceCount;
    const V3* vertices;
    const b3o_halfedge* edges;
    const uint32_t* faces;
    const PL* planes;
} Hull;

static Hull hull_at(const b3o_hull_pool* pool, uint32_t h) {
    const uint32_t* d = pool->desc + 6 * (size_t)h;
    Hull r;
    r.centroid = v3(pool->centroids[3 * h], pool->centroids[3 * h + 1], pool->centroids[3 * h + 2]);
    r.vertexCount = d[1];
    r.vertices = (const V3*)(pool->vertices + 3 * (size_t)d[0]);
    r.edgeCount = d[3];
    r.edges = (const b3o_halfedge*)(pool->edges + 4 * (size_t)d[2]);
    r.faceCount = d[5];
    r.faces = pool->faces + d[4];
    r.planes = (const PL*)(pool->planes + 4 * (size_t)d[4]);
    return r;
}

static uint32_t support_vertex(const Hull* h, V3 d) {                                  /* hull.inl:31-45 */
    uint32_t maxIndex = 0;
    float maxProjection = dot(d, h->vertices[0]);
    for (uint32_t i = 1; i < h->vertexCount; ++i) {
        float p = dot(d, h->vertices[i]);
        if (p > maxProjection) { maxIndex = i; maxProjection = p; }
    }
    return maxIndex;
}
static uint32_t support_face(const Hull* h, V3 d) {                                    /* hull.inl:47-61 */
    uint32_t maxIndex = 0;
    float maxProjection = dot(d, h->planes[0].n);
    for (uint32_t i = 1; i < h->faceCount; ++i) {
        float p = dot(d, h->planes[i].n);
        if (p > maxProjection) { maxIndex = i; maxProjection = p; }
    }
    return maxIndex;
}
static PL edge_side_plane(const Hull* h, uint32_t index) {                             /* hull.inl:63-78 */
    const b3o_halfedge* e = h->edges + index;
    const b3o_halfedge* t = h->edges + e->twin;
    V3 P = h->vertices[e->origin], Q = h->vertices[t->origin];
    V3 E = sub(Q, P);
    V3 D = cross(E, h->planes[e->face].n);
    PL r;
    r.n = normalize(D);
    r.d = dot(r.n, P);
    return r;
}

/* ---------------------------------------------------------------- sat.cpp */
typedef struct { uint32_t index; float separation; } FaceQuery;
typedef struct { uint32_t index1, index2; float separation; } EdgeQuery;

static float project_hull(const Hull* h, PL plane) {                                   /* sat.cpp:26-30 */
    V3 s = h->vertices[support_vertex(h, neg(plane.n))];
    return plane_dist(s, plane);
}

static FaceQuery query_face(const XF* xf1, const Hull* h1, const XF* xf2, const Hull* h2) { /* sat.cpp:33-58 */
    XF xf = xf_mulT_xf(xf2, xf1);
    FaceQuery q = {0, -MAX_FLOAT};
    for (uint32_t i = 0; i < h1->faceCount; ++i) {
        PL plane = plane_mul(&xf, h1->planes[i]);
        float s = project_hull(h2, plane);
        if (s > q.separation) { q.index = i; q.separation = s; }
    }
    return q;
}

static int is_minkowski_face(V3 A, V3 B, V3 BxA, V3 C, V3 D, V3 DxC) {                 /* sat.cpp:60-71 */
    float ADC = dot(A, DxC);
    float BDC = dot(B, DxC);
    float CBA = dot(C, BxA);
    float DBA = dot(D, BxA);
    FL(1);
    if (!(CBA * DBA < 0.0f)) return 0;
    FL(1);
    if (!(ADC * BDC < 0.0f)) return 0;
    FL(1);
    return CBA * BDC > 0.0f;
}

static float project_edges(V3 P1, V3 E1, V3 P2, V3 E2, V3 C1) {                        /* sat.cpp:73-108 */
    float L1 = length(E1);
    if (L1 < LINEAR_SLOP) return -MAX_FLOAT;
    float L2 = length(E2);
    if (L2 < LINEAR_SLOP) return -MAX_FLOAT;
    const float kTol = 0.005f;
    V3 E1xE2 = cross(E1, E2);
    float L = length(E1xE2);
    FL(2);
    if (L < kTol * L1 * L2) return -MAX_FLOAT;
    FL(1);
    V3 N = scale(1.0f / L, E1xE2);
    if (dot(N, sub(P1, C1)) < 0.0f) N = neg(N);
    return dot(N, sub(P2, P1));
}

static EdgeQuery query_edge(const XF* xf1, const Hull* h1, const XF* xf2, const Hull* h2) { /* sat.cpp:110-173 */
    XF xf = xf_mulT_xf(xf2, xf1);
    V3 C1 = xf_mul(&xf, h1->centroid);
    EdgeQuery q = {0, 0, -MAX_FLOAT};
    for (uint32_t i = 0; i < h1->edgeCount; i += 2) {
        const b3o_halfedge* e1 = h1->edges + i;
        const b3o_halfedge* t1 = h1->edges + i + 1;
        V3 P1 = xf_mul(&xf, h1->vertices[e1->origin]);
        V3 Q1 = xf_mul(&xf, h1->vertices[t1->origin]);
        V3 E1 = sub(Q1, P1);
        V3 U1 = mul_mv(rot(&xf), h1->planes[e1->face].n);
        V3 V1 = mul_mv(rot(&xf), h1->planes[t1->face].n);
        for (uint32_t j = 0; j < h2->edgeCount; j += 2) {
            const b3o_halfedge* e2 = h2->edges + j;
            const b3o_halfedge* t2 = h2->edges + j + 1;
            V3 P2 = h2->vertices[e2->origin];
            V3 Q2 = h2->vertices[t2->origin];
            V3 E2 = sub(Q2, P2);
            V3 U2 = h2->planes[e2->face].n;
            V3 V2 = h2->planes[t2->face].n;
            if (is_minkowski_face(U1, V1, neg(E1), neg(U2), neg(V2), neg(E2))) {
                float s = project_edges(P1, E1, P2, E2, C1);
                if (s > q.separation) { q.separation = s; q.index1 = i; q.index2 = j; }
            }
        }
    }
    return q;
}

static int read_face(const b3o_feature_cache* c, const XF* xf1, const Hull* h1,       /* sat.cpp:209-222 */
                     const XF* xf2, const Hull* h2, float totalRadius) {
    XF xf = xf_mulT_xf(xf2, xf1);
    PL plane = plane_mul(&xf, h1->planes[c->index1]);
    float s = project_hull(h2, plane);
    return s > totalRadius ? B3O_SEPARATION : B3O_OVERLAP;
}

static int read_edge(const b3o_feature_cache* c, const XF* xf1, const Hull* h1,       /* sat.cpp:224-278 */
                     const XF* xf2, const Hull* h2, float totalRadius) {
    uint32_t i = c->index1, j = c->index2;
    XF xf = xf_mulT_xf(xf2, xf1);
    V3 C1 = xf_mul(&xf, h1->centroid);
    const b3o_halfedge* e1 = h1->edges + i;
    const b3o_halfedge* t1 = h1->edges + i + 1;
    V3 P1 = xf_mul(&xf, h1->vertices[e1->origin]);
    V3 Q1 = xf_mul(&xf, h1->vertices[t1->origin]);
    V3 E1 = sub(Q1, P1);
    V3 U1 = mul_mv(rot(&xf), h1->planes[e1->face].n);
    V3 V1 = mul_mv(rot(&xf), h1->planes[t1->face].n);
    const b3o_halfedge* e2 = h2->edges + j;
    const b3o_halfedge* t2 = h2->edges + j + 1;
    V3 P2 = h2->vertices[e2->origin];
    V3 Q2 = h2->vertices[t2->origin];
    V3 E2 = sub(Q2, P2);
    V3 U2 = h2->planes[e2->face].n;
    V3 V2 = h2->planes[t2->face].n;
    if (is_minkowski_face(U1, V1, neg(E1), neg(U2), neg(V2), neg(E2))) {
        float s = project_edges(P1, E1, P2, E2, C1);
        return s > totalRadius ? B3O_SEPARATION : B3O_OVERLAP;
    }
    return B3O_EMPTY;
}

static int read_state(b3o_feature_cache* c, const XF* xf1, const Hull* h1,            /* sat.cpp:177-206 */
                      const XF* xf2, const Hull* h2, float totalRadius) {
    if (c->state == B3O_EMPTY) {
        c->state = B3O_SEPARATION; c->type = B3O_FACE1; c->index1 = 0; c->index2 = 0;
    }
    switch (c->type) {
    case B3O_EDGE1: return read_edge(c, xf1, h1, xf2, h2, totalRadius);
    case B3O_FACE1: return read_face(c, xf1, h1, xf2, h2, totalRadius);
    case B3O_FACE2: return read_face(c, xf2, h2, xf1, h1, totalRadius);
    default: return B3O_EMPTY;
    }
}

/* ---------------------------------------------------------------- clip.h / clip.cpp */
typedef struct { uint32_t inEdge1, outEdge1, inEdge2, outEdge2; } FPair;
/* b3MakePair takes (inEdge1, inEdge2, outEdge1, outEdge2) (clip.h:36-44). */
static FPair make_pair(uint32_t in1, uint32_t in2, uint32_t out1, uint32_t out2) {
    FPair p; p.inEdge1 = in1; p.inEdge2 = in2; p.outEdge1 = out1; p.outEdge2 = out2; return p;
}
typedef struct { V3 position; FPair pair; } ClipVertex;
typedef struct { PL plane; uint32_t id; } ClipPlane;
typedef struct { uint32_t count; ClipVertex v[POLY_CAP]; } ClipPoly;
static void copy_clip(ClipPoly* dst, const ClipPoly* src) {
    dst->count = src->count;
    memcpy(dst->v, src->v, src->count * sizeof(ClipVertex));
}

static void build_polygon(ClipPoly* out, const XF* xf, uint32_t index, const Hull* h) { /* clip.cpp:33-57 */
    out->count = 0;
    uint32_t begin = h->faces[index];
    uint32_t e = begin;
    do {
        const b3o_halfedge* edge = h->edges + e;
        const b3o_halfedge* twin = h->edges + edge->twin;
        uint32_t edgeId = twin->twin;
        uint32_t prevId = twin->next;
        ClipVertex cv;
        cv.position = xf_mul(xf, h->vertices[edge->origin]);
        cv.pair = make_pair(prevId, NULL_EDGE, edgeId, NULL_EDGE);
        if (out->count < POLY_CAP) out->v[out->count++] = cv;
        e = edge->next;
    } while (e != begin);
}

static void clip_polygon_to_plane(ClipPoly* out, const ClipPoly* in, ClipPlane plane) { /* clip.cpp:99-149 */
    out->count = 0;
    ClipVertex v1 = in->v[in->count - 1];
    float d1 = plane_dist(v1.position, plane.plane);
    for (uint32_t i = 0; i < in->count; ++i) {
        ClipVertex v2 = in->v[i];
        float d2 = plane_dist(v2.position, plane.plane);
        if (d1 <= 0.0f && d2 <= 0.0f) {
            if (out->count < POLY_CAP) out->v[out->count++] = v2;
        } else if (d1 <= 0.0f && d2 > 0.0f) {
            FL(2);
            float f = d1 / (d1 - d2);
            ClipVertex cv;
            cv.position = add(v1.position, scale(f, sub(v2.position, v1.position)));
            cv.pair = make_pair(v1.pair.inEdge1, NULL_EDGE, NULL_EDGE, plane.id);
            if (out->count < POLY_CAP) out->v[out->count++] = cv;
        } else if (d1 > 0.0f && d2 <= 0.0f) {
            FL(2);
            float f = d1 / (d1 - d2);
            ClipVertex cv;
            cv.position = add(v1.position, scale(f, sub(v2.position, v1.position)));
            cv.pair = make_pair(v1.pair.inEdge1, plane.id, NULL_EDGE, NULL_EDGE);
            if (out->count < POLY_CAP) out->v[out->count++] = cv;
            if (out->count < POLY_CAP) out->v[out->count++] = v2;
        }
        v1 = v2;
        d1 = d2;
    }
}

static void clip_polygon_to_face(ClipPoly* out, const ClipPoly* in, const XF* xf,      /* clip.cpp:243-280 */
                                 float r, uint32_t index, const Hull* h) {
    static ClipPoly tmp;  /* single-threaded oracle */
    copy_clip(out, in);
    uint32_t begin = h->faces[index];
    uint32_t e = begin;
    do {
        const b3o_halfedge* edge = h->edges + e;
        const b3o_halfedge* twin = h->edges + edge->twin;
        uint32_t edgeId = twin->twin;
        PL plane = edge_side_plane(h, edgeId);
        FL(1);
        plane.d += r;
        ClipPlane cp;
        cp.id = edgeId;
        cp.plane = plane_mul(xf, plane);
        clip_polygon_to_plane(&tmp, out, cp);
        copy_clip(out, &tmp);
        if (out->count == 0) return;
        e = edge->next;
    } while (e != begin);
}

/* ---------------------------------------------------------------- contact_cluster.cpp */
typedef struct { V3 position; uint32_t clipIndex; } CPV;
typedef struct { uint32_t count; CPV v[POLY_CAP]; } CPoly;
static void copy_cpoly(CPoly* dst, const CPoly* src) {
    dst->count = src->count;
    memcpy(dst->v, src->v, src->count * sizeof(CPV));
}

static int is_ccw3(V3 A, V3 B, V3 C, V3 N) {                                            /* contact_cluster.cpp:22-26 */
    V3 n = cross(sub(B, A), sub(C, A));
    return dot(n, N) > 0.0f;
}

static void sort_polygon(CPoly* out, const CPoly* in, V3 N) {                          /* contact_cluster.cpp:33-56 */
    copy_cpoly(out, in);
    V3 A = out->v[0].position;
    for (uint32_t i = 1; i < out->count; ++i) {
        V3 B = out->v[i].position;  /* stale pivot: not refreshed after swaps */
        for (uint32_t j = i + 1; j < out->count; ++j) {
            V3 C = out->v[j].position;
            if (!is_ccw3(A, B, C, N)) {
                CPV t = out->v[i]; out->v[i] = out->v[j]; out->v[j] = t;
            }
        }
    }
}

static void reduce_polygon(CPoly* out, const CPoly* in, V3 N, uint32_t initial) {      /* contact_cluster.cpp:58-200 */
    out->count = 0;
    if (in->count <= MAX_POINTS) { sort_polygon(out, in, N); return; }
    static char chosen[POLY_CAP];
    memset(chosen, 0, in->count);
    out->v[out->count++] = in->v[initial];
    chosen[initial] = 1;
    {
        V3 A = out->v[0].position;
        uint32_t index = 0;
        float mx = -MAX_FLOAT;
        for (uint32_t i = 0; i < in->count; ++i) {
            if (chosen[i]) continue;
            V3 d = sub(in->v[i].position, A);
            float dd = dot(d, d);
            if (dd > mx) { mx = dd; index = i; }
        }
        FL(1);
        if (mx < EPSILON * EPSILON) return;
        out->v[out->count++] = in->v[index];
        chosen[index] = 1;
    }
    {
        V3 A = out->v[0].position, B = out->v[1].position;
        uint32_t index = 0;
        float mx = -MAX_FLOAT;
        for (uint32_t i = 0; i < in->count; ++i) {
            if (chosen[i]) continue;
            V3 C = in->v[i].position;
            V3 n = cross(sub(B, A), sub(C, A));
            float a2 = fabsf(dot(n, N));
            if (a2 > mx) { mx = a2; index = i; }
        }
        const float kAreaTol = 0.01f;
        FL(1);
        if (mx < 2.0f * kAreaTol) return;
        V3 C = in->v[index].position;
        if (!is_ccw3(A, B, C, N)) { CPV t = out->v[0]; out->v[0] = out->v[1]; out->v[1] = t; }
        out->v[out->count++] = in->v[index];
        chosen[index] = 1;
    }
    {
        V3 A = out->v[0].position, B = out->v[1].position;
        uint32_t index = 0;
        float mn = MAX_FLOAT;
        for (uint32_t i = 0; i < in->count; ++i) {
            if (chosen[i]) continue;
            V3 D = in->v[i].position;
            V3 n = cross(sub(B, A), sub(D, A));
            float sa2 = dot(n, N);
            if (sa2 < mn) { mn = sa2; index = i; }
        }
        const float kAreaTol = 0.01f;
        FL(1);
        if (fabsf(mn) < 2.0f * kAreaTol) return;
        out->v[out->count++] = in->v[index];
        chosen[index] = 1;
    }
    {
        static CPoly quad;
        sort_polygon(&quad, out, N);
        copy_cpoly(out, &quad);
    }
}

/* ---------------------------------------------------------------- manifold */
static void set_key(b3o_mpoint* mp, FPair p) {                                         /* manifold.h:41-63 */
    mp->triangleKey = NULL_TRIANGLE;
    mp->key[0] = p.inEdge1; mp->key[1] = p.outEdge1; mp->key[2] = p.inEdge2; mp->key[3] = p.outEdge2;
}

/* ---------------------------------------------------------------- collide_hulls.cpp */
static void build_edge_contact(b3o_manifold* m, const XF* xf1, uint32_t index1, const Hull* h1, /* collide_hulls.cpp:26-91 */
                               const XF* xf2, uint32_t index2, const Hull* h2, int rebuild) {
    const b3o_halfedge* e1 = h1->edges + index1;
    const b3o_halfedge* t1 = h1->edges + index1 + 1;
    V3 C1 = xf_mul(xf1, h1->centroid);
    V3 P1 = xf_mul(xf1, h1->vertices[e1->origin]);
    V3 Q1 = xf_mul(xf1, h1->vertices[t1->origin]);
    V3 E1 = sub(Q1, P1);
    V3 N1 = E1;
    normalize_in_place(&N1);
    const b3o_halfedge* e2 = h2->edges + index2;
    const b3o_halfedge* t2 = h2->edges + index2 + 1;
    V3 C2 = xf_mul(xf2, h2->centroid);
    (void)C2;
    V3 P2 = xf_mul(xf2, h2->vertices[e2->origin]);
    V3 Q2 = xf_mul(xf2, h2->vertices[t2->origin]);
    V3 E2 = sub(Q2, P2);
    V3 N2 = E2;
    normalize_in_place(&N2);
    float b = dot(N1, N2);
    FL(2);
    float den = 1.0f - b * b;
    if (den == 0.0f) return;
    FL(1);
    float inv_den = 1.0f / den;
    V3 E3 = sub(P1, P2);
    float d = dot(N1, E3);
    float e = dot(N2, E3);
    FL(6);
    float s = inv_den * (b * e - d);
    float t = inv_den * (e - b * d);
    V3 c1 = add(P1, scale(s, N1));
    V3 c2 = add(P2, scale(t, N2));
    if (rebuild) {                                                                      /* collide_hulls_cache.cpp:82-93 */
        float w2[3], w1[3];
        barycentric2(w2, P1, Q1, c2);
        barycentric2(w1, P2, Q2, c1);
        if (!(w2[1] > 0.0f && w2[1] <= w2[2] && w1[1] > 0.0f && w1[1] <= w1[2])) return;
    }
    V3 N = cross(E1, E2);
    normalize_in_place(&N);
    if (dot(N, sub(P1, C1)) < 0.0f) N = neg(N);
    FPair pair = make_pair(index1, index1 + 1, index2, index2 + 1);
    m->pointCount = 1;
    m->points[0].localNormal1 = mulT_mv(rot(xf1), N);
    m->points[0].localPoint1 = xf_mulT(xf1, c1);
    m->points[0].localPoint2 = xf_mulT(xf2, c2);
    set_key(&m->points[0], pair);
}

static void build_face_contact(b3o_manifold* m, const XF* xf1, uint32_t index1, const Hull* h1, float r1, /* collide_hulls.cpp:93-204 */
                               const XF* xf2, const Hull* h2, float r2, int flip) {
    static ClipPoly polygon2, clipPolygon2;
    static CPoly polygon1, reduced;
    FL(1);
    float totalRadius = r1 + r2;
    PL localPlane1 = h1->planes[index1];
    PL plane1 = plane_mul(xf1, localPlane1);
    V3 normal1 = mulT_mv(rot(xf2), plane1.n);
    uint32_t index2 = support_face(h2, neg(normal1));
    build_polygon(&polygon2, xf2, index2, h2);
    clip_polygon_to_face(&clipPolygon2, &polygon2, xf1, totalRadius, index1, h1);
    if (clipPolygon2.count == 0) return;
    polygon1.count = 0;
    uint32_t minIndex = 0;
    float minSep = MAX_FLOAT;
    for (uint32_t i = 0; i < clipPolygon2.count; ++i) {
        V3 p = clipPolygon2.v[i].position;
        float sep = plane_dist(p, plane1);
        if (sep <= totalRadius) {
            if (sep < minSep) { minIndex = polygon1.count; minSep = sep; }
            CPV v;
            v.position = closest_on_plane(p, plane1);
            v.clipIndex = i;
            if (polygon1.count < POLY_CAP) polygon1.v[polygon1.count++] = v;
        }
    }
    if (polygon1.count == 0) return;
    V3 normal = plane1.n;
    V3 s_normal = flip ? neg(normal) : normal;
    reduce_polygon(&reduced, &polygon1, s_normal, minIndex);
    uint32_t pointCount = reduced.count;
    for (uint32_t i = 0; i < pointCount; ++i) {
        ClipVertex v2 = clipPolygon2.v[reduced.v[i].clipIndex];
        V3 v1 = closest_on_plane(v2.position, plane1);
        b3o_mpoint* mp = m->points + i;
        if (flip) {
            FPair p = make_pair(v2.pair.inEdge2, v2.pair.inEdge1, v2.pair.outEdge2, v2.pair.outEdge1);
            mp->localNormal1 = mulT_mv(rot(xf2), s_normal);
            mp->localPoint1 = xf_mulT(xf2, v2.position);
            mp->localPoint2 = xf_mulT(xf1, v1);
            set_key(mp, p);
        } else {
            mp->localNormal1 = mulT_mv(rot(xf1), normal);
            mp->localPoint1 = xf_mulT(xf1, v1);
            mp->localPoint2 = xf_mulT(xf2, v2.position);
            set_key(mp, v2.pair);
        }
    }
    m->pointCount = pointCount;
}

/* ---------------------------------------------------------------- collide_hulls_cache.cpp */
static void rebuild_face_contact(b3o_manifold* m, const XF* xf1, uint32_t index1, const Hull* h1, float r1, /* :113-153 */
                                 const XF* xf2, const Hull* h2, float r2, int flip,
                                 QT q10, QT q1, QT q20, QT q2) {
    QT dq0 = q_mul(q_conj(q10), q20);
    QT dq = q_mul(q_conj(q1), q2);
    QT q = q_mul(q_conj(dq0), dq);
    const float kTol = 0.995f;
    if (fabsf(q.w) > kTol) build_face_contact(m, xf1, index1, h1, r1, xf2, h2, r2, flip);
}

static void make_fp(b3o_feature_cache* c, int state, int type, uint32_t i1, uint32_t i2) {
    c->state = (uint32_t)state; c->type = (uint32_t)type; c->index1 = i1; c->index2 = i2;
}

static void collide_cache(b3o_manifold* m, const XF* xf1, const Hull* h1, float r1,   /* collide_hulls_cache.cpp:155-275 */
                          const XF* xf2, const Hull* h2, float r2, b3o_feature_cache* c) {
    PHASE(1);
    FL(1);
    float totalRadius = r1 + r2;
    FaceQuery f1 = query_face(xf1, h1, xf2, h2);
    if (f1.separation > totalRadius) { make_fp(c, B3O_SEPARATION, B3O_FACE1, f1.index, f1.index); return; }
    FaceQuery f2 = query_face(xf2, h2, xf1, h1);
    if (f2.separation > totalRadius) { make_fp(c, B3O_SEPARATION, B3O_FACE2, f2.index, f2.index); return; }
    EdgeQuery eq = query_edge(xf1, h1, xf2, h2);
    if (eq.separation > totalRadius) { make_fp(c, B3O_SEPARATION, B3O_EDGE1, eq.index1, eq.index2); return; }
    PHASE(2);
    const float kTol = 0.1f * LINEAR_SLOP;
    FL(1);
    float fmax12 = f1.separation > f2.separation ? f1.separation : f2.separation;      /* b3Max: a > b ? a : b */
    FL(1);
    if (eq.separation > fmax12 + kTol) {
        build_edge_contact(m, xf1, eq.index1, h1, xf2, eq.index2, h2, 0);
        if (m->pointCount > 0) { make_fp(c, B3O_OVERLAP, B3O_EDGE1, eq.index1, eq.index2); return; }
    } else {
        FL(1);
        if (f1.separation + kTol > f2.separation) {
            build_face_contact(m, xf1, f1.index, h1, r1, xf2, h2, r2, 0);
            if (m->pointCount > 0) { make_fp(c, B3O_OVERLAP, B3O_FACE1, f1.index, f1.index); return; }
        } else {
            build_face_contact(m, xf2, f2.index, h2, r2, xf1, h1, r1, 1);
            if (m->pointCount > 0) { make_fp(c, B3O_OVERLAP, B3O_FACE2, f2.index, f2.index); return; }
        }
    }
    if (eq.separation > fmax12) {
        build_edge_contact(m, xf1, eq.index1, h1, xf2, eq.index2, h2, 0);
        if (m->pointCount > 0) { make_fp(c, B3O_OVERLAP, B3O_EDGE1, eq.index1, eq.index2); return; }
    } else {
        if (f1.separation > f2.separation) {
            build_face_contact(m, xf1, f1.index, h1, r1, xf2, h2, r2, 0);
            if (m->pointCount > 0) { make_fp(c, B3O_OVERLAP, B3O_FACE1, f1.index, f1.index); return; }
        } else {
            build_face_contact(m, xf2, f2.index, h2, r2, xf1, h1, r1, 1);
            if (m->pointCount > 0) { make_fp(c, B3O_OVERLAP, B3O_FACE2, f2.index, f2.index); return; }
        }
    }
    build_edge_contact(m, xf1, eq.index1, h1, xf2, eq.index2, h2, 0);
    make_fp(c, B3O_OVERLAP, B3O_EDGE1, eq.index1, eq.index2);
}

/* Statistics kept by the reference in globals: b3_convexCalls, b3_convexCacheHits
 * (collide_hulls.cpp:291-292), b3_gjkCalls, b3_gjkIters, b3_gjkMaxIters (gjk.cpp:27). */
static uint64_t g_stats[6];  /* ... and b3_gjkCacheHits (gjk.cpp:740) */
void b3o_get_counters(uint64_t out[6]) { memcpy(out, g_stats, sizeof(g_stats)); }
void b3o_reset_counters(void) { memset(g_stats, 0, sizeof(g_stats)); }

static void collide_hulls_cached(b3o_manifold* m, const XF* xf1, const Hull* h1, float r1, /* collide_hulls_cache.cpp:279-344 */
                                 const XF* xf2, const Hull* h2, float r2, b3o_feature_cache* c,
                                 QT q10, QT q1, QT q20, QT q2) {
    PHASE(0);
    FL(1);
    ++g_stats[0];  /* b3CollideHullAndHull (collide_hulls.cpp:304) */
    float totalRadius = r1 + r2;
    int state0 = (int)c->state;
    int state1 = read_state(c, xf1, h1, xf2, h2, totalRadius);
    if (state0 == B3O_SEPARATION && state1 == B3O_SEPARATION) {
        ++g_stats[1];  /* separation cache hit (collide_hulls_cache.cpp:300) */
        return;
    }
    if (state0 == B3O_OVERLAP && state1 == B3O_OVERLAP) {
        switch (c->type) {
        case B3O_EDGE1: build_edge_contact(m, xf1, c->index1, h1, xf2, c->index2, h2, 1); break;
        case B3O_FACE1: rebuild_face_contact(m, xf1, c->index1, h1, r1, xf2, h2, r2, 0, q10, q1, q20, q2); break;
        case B3O_FACE2: rebuild_face_contact(m, xf2, c->index1, h2, r2, xf1, h1, r1, 1, q20, q2, q10, q1); break;
        default: break;
        }
        if (m->pointCount > 0) {
            ++g_stats[1];  /* overlap cache hit (collide_hulls_cache.cpp:334) */
            return;
        }
    }
    c->state = B3O_EMPTY;
    collide_cache(m, xf1, h1, r1, xf2, h2, r2, c);
}

/* ---------------------------------------------------------------- gjk.cpp (distance) */
typedef struct { const V3* v; uint32_t count; } Proxy;  /* b3GJKProxy vertices (gjk_proxy.h:26-44) */
typedef struct { V3 point1, point2, point; float weight; uint32_t index1, index2; } SVert;
typedef struct { SVert v[4]; uint32_t count; } Simplex;
typedef struct { V3 point1, point2; float distance; uint32_t iterations; } GJKOut;

static uint32_t proxy_support(const Proxy* p, V3 d) {                                  /* gjk_proxy.h:57-71 */
    uint32_t maxIndex = 0;
    float maxProjection = dot(d, p->v[0]);
    for (uint32_t i = 1; i < p->count; ++i) {
        float q = dot(d, p->v[i]);
        if (q > maxProjection) { maxIndex = i; maxProjection = q; }
    }
    return maxIndex;
}

static void bary2(float out[3], V3 A, V3 B, V3 Q) {                                    /* gjk.cpp:32-45 */
    V3 AB = sub(B, A), QA = sub(A, Q), QB = sub(B, Q);
    out[0] = dot(QB, AB);
    out[1] = -dot(QA, AB);
    FL(1);
    out[2] = out[0] + out[1];
}

static void bary3(float out[4], V3 A, V3 B, V3 C, V3 Q) {                              /* gjk.cpp:47-72 */
    V3 AB = sub(B, A), AC = sub(C, A);
    V3 QA = sub(A, Q), QB = sub(B, Q), QC = sub(C, Q);
    V3 QB_x_QC = cross(QB, QC), QC_x_QA = cross(QC, QA), QA_x_QB = cross(QA, QB);
    V3 AB_x_AC = cross(AB, AC);
    out[0] = dot(QB_x_QC, AB_x_AC);
    out[1] = dot(QC_x_QA, AB_x_AC);
    out[2] = dot(QA_x_QB, AB_x_AC);
    FL(2);
    out[3] = out[0] + out[1] + out[2];
}

static float det3(V3 a, V3 b, V3 c) { return dot(a, cross(b, c)); }                   /* vec3.h:170-173 */
static float signf_(float x) { return (float)((0.0f < x) - (x < 0.0f)); }             /* math.h:88-92 */

static void bary4(float out[5], V3 A, V3 B, V3 C, V3 D, V3 Q) {                        /* gjk.cpp:74-99 */
    V3 AB = sub(B, A), AC = sub(C, A), AD = sub(D, A);
    V3 QA = sub(A, Q), QB = sub(B, Q), QC = sub(C, Q), QD = sub(D, Q);
    float divisor = det3(AB, AC, AD);
    float sign = signf_(divisor);
    FL(5);
    out[0] = sign * det3(QB, QC, QD);
    out[1] = sign * det3(QA, QD, QC);
    out[2] = sign * det3(QA, QB, QD);
    out[3] = sign * det3(QA, QC, QB);
    out[4] = sign * divisor;
}

static V3 search_direction(const Simplex* s, V3 Q) {                                   /* gjk.cpp:101-146 */
    switch (s->count) {
    case 1: return sub(Q, s->v[0].point);
    case 2: {
        V3 A = s->v[0].point, B = s->v[1].point;
        V3 AB = sub(B, A), AQ = sub(Q, A);
        V3 AB_x_AQ = cross(AB, AQ);
        return cross(AB_x_AQ, AB);
    }
    case 3: {
        V3 A = s->v[0].point, B = s->v[1].point, C = s->v[2].point;
        V3 AB = sub(B, A), AC = sub(C, A), AQ = sub(Q, A);
        V3 AB_x_AC = cross(AB, AC);
        float sign = dot(AB_x_AC, AQ);
        return sign > 0.0f ? AB_x_AC : neg(AB_x_AC);
    }
    default: return v3(0.0f, 0.0f, 0.0f);
    }
}

static V3 closest_point(const Simplex* s) {                                            /* gjk.cpp:148-166 */
    switch (s->count) {
    case 1: return s->v[0].point;
    case 2: return add(scale(s->v[0].weight, s->v[0].point), scale(s->v[1].weight, s->v[1].point));
    case 3: return add(add(scale(s->v[0].weight, s->v[0].point), scale(s->v[1].weight, s->v[1].point)),
                       scale(s->v[2].weight, s->v[2].point));
    default: return v3(0.0f, 0.0f, 0.0f);
    }
}

static void closest_points(const Simplex* s, V3* p1, V3* p2) {                         /* gjk.cpp:168-198 */
    const SVert* v = s->v;
    switch (s->count) {
    case 1: *p1 = v[0].point1; *p2 = v[0].point2; break;
    case 2:
        *p1 = add(scale(v[0].weight, v[0].point1), scale(v[1].weight, v[1].point1));
        *p2 = add(scale(v[0].weight, v[0].point2), scale(v[1].weight, v[1].point2));
        break;
    case 3:
        *p1 = add(add(scale(v[0].weight, v[0].point1), scale(v[1].weight, v[1].point1)), scale(v[2].weight, v[2].point1));
        *p2 = add(add(scale(v[0].weight, v[0].point2), scale(v[1].weight, v[1].point2)), scale(v[2].weight, v[2].point2));
        break;
    case 4:
        *p1 = add(add(add(scale(v[0].weight, v[0].point1), scale(v[1].weight, v[1].point1)),
                      scale(v[2].weight, v[2].point1)), scale(v[3].weight, v[3].point1));
        *p2 = *p1;
        break;
    default: break;
    }
}

static void set1(Simplex* s, SVert a) { s->count = 1; s->v[0] = a; s->v[0].weight = 1.0f; }
static void set2(Simplex* s, SVert a, SVert b, const float w[3]) {
    FL(3);
    float d = 1.0f / w[2];
    s->count = 2;
    s->v[0] = a; s->v[0].weight = d * w[0];
    s->v[1] = b; s->v[1].weight = d * w[1];
}
static void set3(Simplex* s, SVert a, SVert b, SVert c, float w0, float w1, float w2, float divisor) {
    FL(4);
    float d = 1.0f / divisor;
    s->count = 3;
    s->v[0] = a; s->v[0].weight = d * w0;
    s->v[1] = b; s->v[1].weight = d * w1;
    s->v[2] = c; s->v[2].weight = d * w2;
}

static void solve2(Simplex* s, V3 Q) {                                                 /* gjk.cpp:201-234 */
    SVert A = s->v[0], B = s->v[1];
    float wAB[3];
    bary2(wAB, A.point, B.point, Q);
    if (wAB[1] <= 0.0f) { set1(s, A); return; }
    if (wAB[0] <= 0.0f) { set1(s, B); return; }
    set2(s, A, B, wAB);
}

static void solve3(Simplex* s, V3 Q) {                                                 /* gjk.cpp:236-332 */
    SVert A = s->v[0], B = s->v[1], C = s->v[2];
    float wAB[3], wBC[3], wCA[3];
    bary2(wAB, A.point, B.point, Q);
    bary2(wBC, B.point, C.point, Q);
    bary2(wCA, C.point, A.point, Q);
    if (wAB[1] <= 0.0f && wCA[0] <= 0.0f) { set1(s, A); return; }
    if (wAB[0] <= 0.0f && wBC[1] <= 0.0f) { set1(s, B); return; }
    if (wBC[0] <= 0.0f && wCA[1] <= 0.0f) { set1(s, C); return; }
    float wABC[4];
    bary3(wABC, A.point, B.point, C.point, Q);
    float area = wABC[3];
    FL(1);
    if (wAB[0] > 0.0f && wAB[1] > 0.0f && area * wABC[2] <= 0.0f) { set2(s, A, B, wAB); return; }
    FL(1);
    if (wBC[0] > 0.0f && wBC[1] > 0.0f && area * wABC[0] <= 0.0f) { set2(s, B, C, wBC); return; }
    FL(1);
    if (wCA[0] > 0.0f && wCA[1] > 0.0f && area * wABC[1] <= 0.0f) { set2(s, C, A, wCA); return; }
    float divisor = wABC[3];
    if (divisor <= 0.0f) return;  /* give up */
    set3(s, A, B, C, wABC[0], wABC[1], wABC[2], divisor);
}

static void solve4(Simplex* s, V3 Q) {                                                 /* gjk.cpp:334-575 */
    SVert A = s->v[0], B = s->v[1], C = s->v[2], D = s->v[3];
    float wAB[3], wAC[3], wAD[3], wBC[3], wCD[3], wDB[3];
    bary2(wAB, A.point, B.point, Q);
    bary2(wBC, B.point, C.point, Q);
    bary2(wAC, A.point, C.point, Q);
    bary2(wAD, A.point, D.point, Q);
    bary2(wCD, C.point, D.point, Q);
    bary2(wDB, D.point, B.point, Q);
    if (wAB[1] <= 0.0f && wAC[1] <= 0.0f && wAD[1] <= 0.0f) { set1(s, A); return; }
    if (wAB[0] <= 0.0f && wDB[0] <= 0.0f && wBC[1] <= 0.0f) { set1(s, B); return; }
    if (wAC[0] <= 0.0f && wBC[0] <= 0.0f && wCD[1] <= 0.0f) { set1(s, C); return; }
    if (wAD[0] <= 0.0f && wCD[0] <= 0.0f && wDB[1] <= 0.0f) { set1(s, D); return; }
    float wACB[4], wABD[4], wADC[4], wBCD[4];
    bary3(wACB, A.point, C.point, B.point, Q);
    bary3(wABD, A.point, B.point, D.point, Q);
    bary3(wADC, A.point, D.point, C.point, Q);
    bary3(wBCD, B.point, C.point, D.point, Q);
    if (wABD[2] <= 0.0f && wACB[1] <= 0.0f && wAB[0] > 0.0f && wAB[1] > 0.0f) { set2(s, A, B, wAB); return; }
    if (wACB[2] <= 0.0f && wADC[1] <= 0.0f && wAC[0] > 0.0f && wAC[1] > 0.0f) { set2(s, A, C, wAC); return; }
    if (wADC[2] <= 0.0f && wABD[1] <= 0.0f && wAD[0] > 0.0f && wAD[1] > 0.0f) { set2(s, A, D, wAD); return; }
    if (wACB[0] <= 0.0f && wBCD[2] <= 0.0f && wBC[0] > 0.0f && wBC[1] > 0.0f) { set2(s, B, C, wBC); return; }
    if (wADC[0] <= 0.0f && wBCD[0] <= 0.0f && wCD[0] > 0.0f && wCD[1] > 0.0f) { set2(s, C, D, wCD); return; }
    if (wABD[0] <= 0.0f && wBCD[1] <= 0.0f && wDB[0] > 0.0f && wDB[1] > 0.0f) { set2(s, D, B, wDB); return; }
    float wABCD[5];
    bary4(wABCD, A.point, B.point, C.point, D.point, Q);
    if (wABCD[3] <= 0.0f && wACB[0] > 0.0f && wACB[1] > 0.0f && wACB[2] > 0.0f) {
        FL(2);
        set3(s, A, C, B, wACB[0], wACB[1], wACB[2], wACB[0] + wACB[1] + wACB[2]);
        return;
    }
    if (wABCD[2] <= 0.0f && wABD[0] > 0.0f && wABD[1] > 0.0f && wABD[2] > 0.0f) {
        FL(2);
        set3(s, A, B, D, wABD[0], wABD[1], wABD[2], wABD[0] + wABD[1] + wABD[2]);
        return;
    }
    if (wABCD[1] <= 0.0f && wADC[0] > 0.0f && wADC[1] > 0.0f && wADC[2] > 0.0f) {
        FL(2);
        set3(s, A, D, C, wADC[0], wADC[1], wADC[2], wADC[0] + wADC[1] + wADC[2]);
        return;
    }
    if (wABCD[0] <= 0.0f && wBCD[0] > 0.0f && wBCD[1] > 0.0f && wBCD[2] > 0.0f) {
        FL(2);
        set3(s, B, C, D, wBCD[0], wBCD[1], wBCD[2], wBCD[0] + wBCD[1] + wBCD[2]);
        return;
    }
    FL(3);
    float divisor = wABCD[0] + wABCD[1] + wABCD[2] + wABCD[3];
    if (divisor <= 0.0f) return;  /* give up */
    FL(5);
    float d = 1.0f / divisor;
    s->count = 4;
    s->v[0].weight = d * wABCD[0];
    s->v[1].weight = d * wABCD[1];
    s->v[2].weight = d * wABCD[2];
    s->v[3].weight = d * wABCD[3];
}

/* b3Simplex::GetMetric (gjk.cpp:806-841): 0 for one vertex, the segment length, the
 * triangle's doubled area, or the tetrahedron's |det| (b3Det = dot(a, cross(b, c)),
 * vec3.h:171-174; b3Sign, math.h:80-83). */
static float simplex_metric(const Simplex* s) {
    switch (s->count) {
    case 2:
        return length(sub(s->v[0].point, s->v[1].point));  /* b3Distance (vec3.h:201-204) */
    case 3: {
        V3 E1 = sub(s->v[1].point, s->v[0].point), E2 = sub(s->v[2].point, s->v[0].point);
        return length(cross(E1, E2));
    }
    case 4: {
        V3 E1 = sub(s->v[1].point, s->v[0].point), E2 = sub(s->v[2].point, s->v[0].point);
        V3 E3 = sub(s->v[3].point, s->v[0].point);
        float det = dot(E1, cross(E2, E3));
        float sign = (float)(0.0f < det) - (float)(det < 0.0f);
        return sign * det;
    }
    default:
        return 0.0f;
    }
}

/* b3GJK (gjk.cpp:579-737) with b3Simplex::ReadCache / WriteCache (gjk.cpp:743-804).  cache
 * NULL: the no-cache wrapper (gjk.cpp:845-851), an empty cache that is not written back.
 * A vertex added to the simplex by a support step starts with weight 0; the reference
 * leaves it unset, which is only observable after the degenerate give-up returns of
 * Solve3/Solve4 (ReadCache sets the cached vertices' weights to 0 itself).  A cache count
 * above 4 or a cached index beyond a proxy's vertices (a cache the reference could not have
 * written for these shapes; it would read out of bounds) is read as an empty cache. */
static GJKOut gjk_cached(const XF* xf1, const Proxy* p1, const XF* xf2, const Proxy* p2, b3o_simplex_cache* cache) {
    Simplex s;
    memset(&s, 0, sizeof(s));
    ++g_stats[2];  /* gjk.cpp:583 */
    s.count = cache ? (uint8_t)cache->count : 0u;
    if (s.count > 4) s.count = 0;
    for (uint32_t i = 0; i < s.count; ++i)
        if (cache->index1[i] >= p1->count || cache->index2[i] >= p2->count) s.count = 0;
    for (uint32_t i = 0; i < s.count; ++i) {
        SVert* v = s.v + i;
        v->index1 = cache->index1[i];
        v->index2 = cache->index2[i];
        v->point1 = xf_mul(xf1, p1->v[v->index1]);
        v->point2 = xf_mul(xf2, p2->v[v->index2]);
        v->point = sub(v->point2, v->point1);
        v->weight = 0.0f;
    }
    if (s.count > 1) {
        float metric1 = cache->metric;
        float metric2 = simplex_metric(&s);
        if (metric2 < 0.5f * metric1 || 2.0f * metric1 < metric2 || metric2 < EPSILON)
            s.count = 0;  /* flush */
        else
            ++g_stats[5]; /* b3_gjkCacheHits (gjk.cpp:774) */
    }
    if (s.count == 0) {
        s.v[0].point1 = xf_mul(xf1, p1->v[0]);
        s.v[0].point2 = xf_mul(xf2, p2->v[0]);
        s.v[0].point = sub(s.v[0].point2, s.v[0].point1);
        s.v[0].weight = 1.0f;
        s.v[0].index1 = 0;
        s.v[0].index2 = 0;
        s.count = 1;
    }
    uint32_t save1[4], save2[4], saveCount;
    const V3 origin = v3(0.0f, 0.0f, 0.0f);
    uint32_t iter = 0;
    while (iter < 20) {
        saveCount = s.count;
        for (uint32_t i = 0; i < saveCount; ++i) { save1[i] = s.v[i].index1; save2[i] = s.v[i].index2; }
        switch (s.count) {
        case 2: solve2(&s, origin); break;
        case 3: solve3(&s, origin); break;
        case 4: solve4(&s, origin); break;
        default: break;
        }
        if (s.count == 4) break;
        V3 p = closest_point(&s);
        (void)dot(p, p);  /* distSq2: the "getting closer" exit is disabled (gjk.cpp:648-651) */
        V3 d = search_direction(&s, origin);
        if (dot(d, d) < EPSILON * EPSILON) break;
        SVert* vx = s.v + s.count;
        vx->index1 = proxy_support(p1, mulT_mv(rot(xf1), neg(d)));
        vx->point1 = xf_mul(xf1, p1->v[vx->index1]);
        vx->index2 = proxy_support(p2, mulT_mv(rot(xf2), d));
        vx->point2 = xf_mul(xf2, p2->v[vx->index2]);
        vx->point = sub(vx->point2, vx->point1);
        vx->weight = 0.0f;
        ++iter;
        int dup = 0;
        for (uint32_t i = 0; i < saveCount; ++i)
            if (vx->index1 == save1[i] && vx->index2 == save2[i]) { dup = 1; break; }
        if (dup) break;
        ++s.count;
    }
    g_stats[3] += iter;                          /* gjk.cpp:673 */
    if (iter > g_stats[4]) g_stats[4] = iter;    /* gjk.cpp:697 */
    GJKOut o;
    closest_points(&s, &o.point1, &o.point2);
    o.distance = length(sub(o.point1, o.point2));
    o.iterations = iter;
    if (cache) {  /* b3Simplex::WriteCache (gjk.cpp:796-804): metric, count and the used indices */
        cache->metric = simplex_metric(&s);
        cache->count = (uint16_t)s.count;
        for (uint32_t i = 0; i < s.count; ++i) {
            cache->index1[i] = (uint8_t)s.v[i].index1;
            cache->index2[i] = (uint8_t)s.v[i].index2;
        }
    }
    return o;
}

static GJKOut gjk(const XF* xf1, const Proxy* p1, const XF* xf2, const Proxy* p2) {
    return gjk_cached(xf1, p1, xf2, p2, NULL);
}

/* ---------------------------------------------------------------- convex shape records */
typedef struct { V3 c[2]; uint32_t count; float radius; } Round;  /* sphere: 1 centre, capsule: 2 */

static Round round_of(const b3o_shape* s) {
    Round r;
    r.c[0] = v3(s->v[0], s->v[1], s->v[2]);
    r.c[1] = v3(s->v[3], s->v[4], s->v[5]);
    r.count = s->type == B3O_SPHERE ? 1u : 2u;
    r.radius = s->radius;
    return r;
}

/* b3ShapeGJKProxy::Set (collide.cpp:30-82) for a convex shape: its vertices and radius. */
static Proxy shape_proxy(const b3o_hull_pool* pool, const b3o_shape* s, Round* round, Hull* hull, float* radius) {
    *radius = s->radius;
    if (s->type == B3O_HULL) {
        *hull = hull_at(pool, s->hull);
        Proxy p = {hull->vertices, hull->vertexCount};
        return p;
    }
    *round = round_of(s);
    Proxy p = {round->c, round->count};
    return p;
}

/* b3GJK's applyRadius branch (gjk.cpp:709-733) reduced to the distance, and b3TestOverlap's
 * tolerance (collide.cpp:84-95): the witness points it moves do not enter the result. */
static int overlap_of(GJKOut g, float r1, float r2) {
    float distance = g.distance;
    if (g.distance > r1 + r2 && g.distance > EPSILON)
        distance -= r1 + r2;
    else
        distance = 0.0f;
    const float kTol = 10.0f * EPSILON;
    return distance <= kTol;
}

static void set_zero_key(b3o_mpoint* mp) {
    mp->triangleKey = NULL_TRIANGLE;
    mp->key[0] = mp->key[1] = mp->key[2] = mp->key[3] = 0;
}

/* collide_spheres.cpp:24-56 */
static void collide_spheres(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, const Round* s2) {
    V3 c1 = xf_mul(xf1, s1->c[0]);
    float r1 = s1->radius;
    V3 c2 = xf_mul(xf2, s2->c[0]);
    float r2 = s2->radius;
    V3 d = sub(c2, c1);
    float dd = dot(d, d);
    FL(2);
    float totalRadius = r1 + r2;
    if (dd > totalRadius * totalRadius) return;
    float distance = length(d);
    V3 normal = v3(0.0f, 1.0f, 0.0f);
    if (distance > EPSILON) { FL(3); normal = v3(d.x / distance, d.y / distance, d.z / distance); }
    m->pointCount = 1;
    m->points[0].localNormal1 = mulT_mv(rot(xf1), normal);
    m->points[0].localPoint1 = s1->c[0];
    m->points[0].localPoint2 = s2->c[0];
    set_zero_key(&m->points[0]);
}

/* collide_sphere_capsule.cpp:24-133 */
static void collide_sphere_capsule(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, const Round* s2) {
    V3 Q = xf_mul(xf1, s1->c[0]);
    V3 A = xf_mul(xf2, s2->c[0]);
    V3 B = xf_mul(xf2, s2->c[1]);
    V3 AB = sub(B, A);
    float u = dot(sub(B, Q), AB);
    float v = dot(sub(Q, A), AB);
    FL(1);
    float radius = s1->radius + s2->radius;
    if (v <= 0.0f || u <= 0.0f) {
        V3 P = v <= 0.0f ? A : B;
        V3 d = sub(P, Q);
        float dd = dot(d, d);
        FL(1);
        if (dd > radius * radius) return;
        V3 n = v3(0.0f, 1.0f, 0.0f);
        float len = length(d);
        if (len > EPSILON) { FL(3); n = v3(d.x / len, d.y / len, d.z / len); }
        m->pointCount = 1;
        m->points[0].localNormal1 = mulT_mv(rot(xf1), n);
        m->points[0].localPoint1 = s1->c[0];
        m->points[0].localPoint2 = v <= 0.0f ? s2->c[0] : s2->c[1];
        set_zero_key(&m->points[0]);
        return;
    }
    float s = dot(AB, AB);
    V3 P;
    FL(1);
    if (s < LINEAR_SLOP * LINEAR_SLOP) {
        P = A;
    } else {
        V3 w = add(scale(u, A), scale(v, B));
        FL(3);
        P = v3(w.x / s, w.y / s, w.z / s);
    }
    V3 d = sub(P, Q);
    float dd = dot(d, d);
    FL(1);
    if (dd > radius * radius) return;
    V3 QA = sub(A, Q);
    V3 e = cross(AB, QA);
    V3 n = cross(AB, e);
    if (dot(n, QA) < 0.0f) n = neg(n);
    normalize_in_place(&n);
    m->pointCount = 1;
    m->points[0].localNormal1 = mulT_mv(rot(xf1), n);
    m->points[0].localPoint1 = s1->c[0];
    m->points[0].localPoint2 = xf_mulT(xf2, P);
    set_zero_key(&m->points[0]);
}

/* collide_capsules.cpp:27-57 */
static V3 closest_on_segment(V3 Q, V3 A, V3 B) {
    V3 AB = sub(B, A);
    float u = dot(sub(B, Q), AB);
    float v = dot(sub(Q, A), AB);
    if (v <= 0.0f) return A;
    if (u <= 0.0f) return B;
    float w = dot(AB, AB);
    FL(1);
    if (w <= LINEAR_SLOP * LINEAR_SLOP) return A;
    FL(1);
    float den = 1.0f / w;
    return scale(den, add(scale(u, A), scale(v, B)));
}

/* collide_capsules.cpp:59-152 */
static void closest_points_segments(V3* C1, V3* C2, V3 P1, V3 Q1, V3 P2, V3 Q2) {
    V3 E1 = sub(Q1, P1);
    float L1 = length(E1);
    V3 E2 = sub(Q2, P2);
    float L2 = length(E2);
    if (L1 < LINEAR_SLOP && L2 < LINEAR_SLOP) { *C1 = P1; *C2 = P2; return; }
    if (L1 < LINEAR_SLOP) { *C1 = P1; *C2 = closest_on_segment(P1, P2, Q2); return; }
    if (L2 < LINEAR_SLOP) { *C1 = closest_on_segment(P2, P1, Q1); *C2 = P2; return; }
    FL(6);
    V3 N1 = v3(E1.x / L1, E1.y / L1, E1.z / L1);
    V3 N2 = v3(E2.x / L2, E2.y / L2, E2.z / L2);
    float a11 = 1.0f, a12 = -dot(N1, N2);
    float a21 = -a12, a22 = -1.0f;
    /* b3Mat22 A: columns x = (a11, a21), y = (a12, a22) */
    float Axx = a11, Axy = a21, Ayx = a12, Ayy = a22;
    FL(6);
    float n0 = fabsf(Axx) + fabsf(Axy), n1 = fabsf(Ayx) + fabsf(Ayy);
    float norm_A = n0 > n1 ? n0 : n1;
    float det_A = 1.0f / (a11 * a22 - a12 * a21);
    FL(4);
    float Ixx = det_A * a22, Ixy = -det_A * a21;
    float Iyx = -det_A * a12, Iyy = det_A * a11;
    FL(3);
    float m0 = fabsf(Ixx) + fabsf(Ixy), m1 = fabsf(Iyx) + fabsf(Iyy);
    float norm_inv_A = m0 > m1 ? m0 : m1;
    float k_A = norm_A * norm_inv_A;
    const float kMaxConditionNumber = 1000.0f;
    if (!isinf(k_A) && k_A < kMaxConditionNumber) {
        V3 E3 = sub(P1, P2);
        float bx = -dot(N1, E3);
        float by = -dot(N2, E3);
        FL(6);
        float xx = bx * Ixx + by * Iyx;  /* b3Mul(Mat22, Vec2) = v.x * A.x + v.y * A.y (mat22.h:56-59) */
        float xy = bx * Ixy + by * Iyy;
        *C1 = add(P1, scale(xx, N1));
        *C2 = add(P2, scale(xy, N2));
    } else {
        *C1 = P1;
        *C2 = P2;
    }
    *C1 = closest_on_segment(*C1, P1, Q1);
    *C2 = closest_on_segment(*C1, P2, Q2);
    *C1 = closest_on_segment(*C2, P1, Q1);
}

/* collide_capsules.cpp:154-174 */
static int are_parallel(V3 P1, V3 Q1, V3 P2, V3 Q2) {
    V3 E1 = sub(Q1, P1);
    float L1 = length(E1);
    if (L1 < LINEAR_SLOP) return 0;
    V3 E2 = sub(Q2, P2);
    float L2 = length(E2);
    if (L2 < LINEAR_SLOP) return 0;
    const float kTol = 0.005f;
    V3 N = cross(E1, E2);
    FL(2);
    return length(N) < kTol * L1 * L2;
}

/* clip.cpp:59-97 (b3ClipEdgeToPlane) */
static uint32_t clip_edge_to_plane(ClipVertex out[2], const ClipVertex in[2], ClipPlane plane) {
    uint32_t n = 0;
    float d1 = plane_dist(in[0].position, plane.plane);
    float d2 = plane_dist(in[1].position, plane.plane);
    if (d1 <= 0.0f) out[n++] = in[0];
    if (d2 <= 0.0f) out[n++] = in[1];
    if (d1 <= 0.0f && d2 > 0.0f) {
        FL(2);
        float f = d1 / (d1 - d2);
        out[n].position = add(in[0].position, scale(f, sub(in[1].position, in[0].position)));
        out[n].pair = make_pair(in[0].pair.inEdge1, NULL_EDGE, NULL_EDGE, plane.id);
        ++n;
    } else if (d1 > 0.0f && d2 <= 0.0f) {
        FL(2);
        float f = d1 / (d1 - d2);
        out[n].position = add(in[0].position, scale(f, sub(in[1].position, in[0].position)));
        out[n].pair = make_pair(in[1].pair.inEdge1, plane.id, NULL_EDGE, NULL_EDGE);
        ++n;
    }
    return n;
}

/* clip.cpp:22-31 (b3BuildEdge) */
static void build_edge(ClipVertex e[2], V3 A, V3 B) {
    e[0].position = A;
    e[0].pair = make_pair(0, NULL_EDGE, 0, NULL_EDGE);
    e[1].position = B;
    e[1].pair = make_pair(1, NULL_EDGE, 0, NULL_EDGE);
}

/* clip.cpp:152-198 (b3ClipEdgeToFace against a capsule segment) */
static uint32_t clip_edge_to_segment(ClipVertex out[2], const ClipVertex in[2], V3 P1, V3 Q1) {
    ClipVertex cur[2] = {in[0], in[1]};
    V3 E1 = sub(Q1, P1);
    ClipPlane cp1;
    cp1.plane.n = normalize(E1);
    cp1.plane.d = dot(cp1.plane.n, Q1);
    cp1.id = 0;
    ClipVertex tmp[2];
    memset(tmp, 0, sizeof(tmp));
    uint32_t n = clip_edge_to_plane(tmp, cur, cp1);
    cur[0] = tmp[0]; cur[1] = tmp[1];
    if (n < 2) { out[0] = cur[0]; out[1] = cur[1]; return n; }
    ClipPlane cp2;
    cp2.plane.n = neg(cp1.plane.n);
    cp2.plane.d = dot(cp2.plane.n, P1);
    cp2.id = 1;
    ClipVertex tmp2[2];
    memset(tmp2, 0, sizeof(tmp2));
    n = clip_edge_to_plane(tmp2, cur, cp2);
    out[0] = tmp2[0]; out[1] = tmp2[1];
    return n;
}

/* clip.cpp:201-240 (b3ClipEdgeToFace against hull face side planes) */
static uint32_t clip_edge_to_hull_face(ClipVertex out[2], const ClipVertex in[2], const XF* xf, float r,
                                       uint32_t index, const Hull* h) {
    ClipVertex cur[2] = {in[0], in[1]};
    uint32_t n = 0;
    uint32_t begin = h->faces[index];
    uint32_t e = begin;
    do {
        const b3o_halfedge* edge = h->edges + e;
        const b3o_halfedge* twin = h->edges + edge->twin;
        uint32_t edgeId = twin->twin;
        PL plane = edge_side_plane(h, edgeId);
        FL(1);
        plane.d += r;
        ClipPlane cp;
        cp.id = edgeId;
        cp.plane = plane_mul(xf, plane);
        ClipVertex tmp[2];
        memset(tmp, 0, sizeof(tmp));
        n = clip_edge_to_plane(tmp, cur, cp);
        cur[0] = tmp[0]; cur[1] = tmp[1];
        if (n == 0) break;
        e = edge->next;
    } while (e != begin);
    out[0] = cur[0]; out[1] = cur[1];
    return n;
}

static void set_pair_key(b3o_mpoint* mp, FPair p) { set_key(mp, p); }

/* collide_capsules.cpp:176-259 */
static void collide_capsules(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, const Round* s2) {
    V3 A1 = xf_mul(xf1, s1->c[0]), B1 = xf_mul(xf1, s1->c[1]);
    V3 A2 = xf_mul(xf2, s2->c[0]), B2 = xf_mul(xf2, s2->c[1]);
    V3 point1, point2;
    closest_points_segments(&point1, &point2, A1, B1, A2, B2);
    float distance = length(sub(point1, point2));
    float r1 = s1->radius, r2 = s2->radius;
    FL(1);
    float totalRadius = r1 + r2;
    if (distance > totalRadius) return;
    if (!(distance > EPSILON)) return;
    if (are_parallel(A1, B1, A2, B2)) {
        ClipVertex edge1[2], clip1[2];
        build_edge(edge1, A1, B1);
        uint32_t clipCount = clip_edge_to_segment(clip1, edge1, A2, B2);
        if (clipCount == 2) {
            V3 cp1 = closest_on_segment(clip1[0].position, A2, B2);
            V3 cp2 = closest_on_segment(clip1[1].position, A2, B2);
            float d1 = length(sub(clip1[0].position, cp1));
            float d2 = length(sub(clip1[1].position, cp2));
            if (d1 > EPSILON && d1 <= totalRadius && d2 > EPSILON && d2 <= totalRadius) {
                V3 g1 = sub(cp1, clip1[0].position), g2 = sub(cp2, clip1[1].position);
                FL(6);
                V3 n1 = v3(g1.x / d1, g1.y / d1, g1.z / d1);
                V3 n2 = v3(g2.x / d2, g2.y / d2, g2.z / d2);
                m->pointCount = 2;
                m->points[0].localNormal1 = mulT_mv(rot(xf1), n1);
                m->points[0].localPoint1 = xf_mulT(xf1, clip1[0].position);
                m->points[0].localPoint2 = xf_mulT(xf2, cp1);
                set_pair_key(&m->points[0], clip1[0].pair);
                m->points[1].localNormal1 = mulT_mv(rot(xf1), n2);
                m->points[1].localPoint1 = xf_mulT(xf1, clip1[1].position);
                m->points[1].localPoint2 = xf_mulT(xf2, cp2);
                set_pair_key(&m->points[1], clip1[1].pair);
                return;
            }
        }
    }
    V3 g = sub(point2, point1);
    FL(3);
    V3 normal = v3(g.x / distance, g.y / distance, g.z / distance);
    m->pointCount = 1;
    m->points[0].localNormal1 = mulT_mv(rot(xf1), normal);
    m->points[0].localPoint1 = xf_mulT(xf1, point1);
    m->points[0].localPoint2 = xf_mulT(xf2, point2);
    set_zero_key(&m->points[0]);
}

/* collide_sphere_hull.cpp:26-87 with b3QueryFaceSeparation(sphere, hull) (sat_vertex_and_hull.cpp:29-54) */
static void collide_sphere_hull(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, const Hull* h2, float r2) {
    Proxy p1 = {s1->c, 1}, p2 = {h2->vertices, h2->vertexCount};
    GJKOut g = gjk(xf1, &p1, xf2, &p2);
    float r1 = s1->radius;
    FL(1);
    float totalRadius = r1 + r2;
    if (g.distance > totalRadius) return;
    if (g.distance > 0.0f) {
        V3 c1 = g.point1, c2 = g.point2;
        V3 dd = sub(c2, c1);
        FL(3);
        V3 normal = v3(dd.x / g.distance, dd.y / g.distance, dd.z / g.distance);
        m->pointCount = 1;
        m->points[0].localNormal1 = mulT_mv(rot(xf1), normal);
        m->points[0].localPoint1 = s1->c[0];
        m->points[0].localPoint2 = xf_mulT(xf2, c2);
        set_zero_key(&m->points[0]);
        return;
    }
    V3 support = xf_mulT(xf2, xf_mul(xf1, s1->c[0]));
    uint32_t maxIndex = 0;
    float maxSep = -MAX_FLOAT;
    for (uint32_t i = 0; i < h2->faceCount; ++i) {
        float sep = plane_dist(support, h2->planes[i]);
        if (sep > maxSep) { maxIndex = i; maxSep = sep; }
    }
    if (maxSep > totalRadius) return;
    PL plane2 = plane_mul(xf2, h2->planes[maxIndex]);
    V3 c1 = xf_mul(xf1, s1->c[0]);
    V3 c2 = closest_on_plane(c1, plane2);
    V3 n1 = neg(plane2.n);
    m->pointCount = 1;
    m->points[0].localNormal1 = mulT_mv(rot(xf1), n1);
    m->points[0].localPoint1 = s1->c[0];
    m->points[0].localPoint2 = xf_mulT(xf2, c2);
    set_zero_key(&m->points[0]);
}

/* collide_capsule_hull.cpp:29-87 */
static void capsule_edge_contact(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, uint32_t index2, const Hull* h2) {
    V3 P1 = xf_mul(xf1, s1->c[0]);
    V3 Q1 = xf_mul(xf1, s1->c[1]);
    V3 E1 = sub(Q1, P1);
    V3 N1 = E1;
    normalize_in_place(&N1);
    const b3o_halfedge* e2 = h2->edges + index2;
    const b3o_halfedge* t2 = h2->edges + index2 + 1;
    V3 C2 = xf_mul(xf2, h2->centroid);
    V3 P2 = xf_mul(xf2, h2->vertices[e2->origin]);
    V3 Q2 = xf_mul(xf2, h2->vertices[t2->origin]);
    V3 E2 = sub(Q2, P2);
    V3 N2 = E2;
    normalize_in_place(&N2);
    float b = dot(N1, N2);
    FL(2);
    float den = 1.0f - b * b;
    if (den == 0.0f) return;
    FL(1);
    float inv_den = 1.0f / den;
    V3 E3 = sub(P1, P2);
    float d = dot(N1, E3);
    float e = dot(N2, E3);
    FL(6);
    float s = inv_den * (b * e - d);
    float t = inv_den * (e - b * d);
    V3 c1 = add(P1, scale(s, N1));
    V3 c2 = add(P2, scale(t, N2));
    V3 N = cross(E1, E2);
    normalize_in_place(&N);
    if (dot(N, sub(P2, C2)) > 0.0f) N = neg(N);
    FPair pair = make_pair(0, 1, index2, index2 + 1);
    m->pointCount = 1;
    m->points[0].localNormal1 = mulT_mv(rot(xf1), N);
    m->points[0].localPoint1 = xf_mulT(xf1, c1);
    m->points[0].localPoint2 = xf_mulT(xf2, c2);
    set_key(&m->points[0], pair);
}

/* collide_capsule_hull.cpp:89-140 */
static void capsule_face_contact(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, uint32_t index2,
                                 const Hull* h2, float r2) {
    V3 A = xf_mul(xf1, s1->c[0]), B = xf_mul(xf1, s1->c[1]);
    float r1 = s1->radius;
    ClipVertex edge1[2], clip1[2];
    build_edge(edge1, A, B);
    uint32_t clipCount = clip_edge_to_hull_face(clip1, edge1, xf2, r2, index2, h2);
    PL plane2 = plane_mul(xf2, h2->planes[index2]);
    V3 n1 = neg(plane2.n);
    FL(1);
    float totalRadius = r1 + r2;
    uint32_t pc = 0;
    for (uint32_t i = 0; i < clipCount; ++i) {
        V3 c1 = clip1[i].position;
        float s = plane_dist(c1, plane2);
        if (s <= totalRadius) {
            V3 c2 = closest_on_plane(c1, plane2);
            b3o_mpoint* mp = m->points + pc;
            mp->localNormal1 = mulT_mv(rot(xf1), n1);
            mp->localPoint1 = xf_mulT(xf1, c1);
            mp->localPoint2 = xf_mulT(xf2, c2);
            set_key(mp, clip1[i].pair);
            ++pc;
        }
    }
    m->pointCount = pc;
}

/* collide_capsule_hull.cpp:142-222 with the capsule SAT queries (sat_edge_and_hull.cpp:23-143) */
static void collide_capsule_hull(b3o_manifold* m, const XF* xf1, const Round* s1, const XF* xf2, const Hull* h2, float r2) {
    Proxy p1 = {s1->c, 2}, p2 = {h2->vertices, h2->vertexCount};
    GJKOut g = gjk(xf1, &p1, xf2, &p2);
    float r1 = s1->radius;
    FL(1);
    float totalRadius = r1 + r2;
    if (g.distance > totalRadius) return;
    if (g.distance > EPSILON) {
        V3 dd = sub(g.point2, g.point1);
        FL(3);
        V3 N1 = v3(dd.x / g.distance, dd.y / g.distance, dd.z / g.distance);
        V3 localN1 = mulT_mv(rot(xf2), N1);
        uint32_t index2 = support_face(h2, neg(localN1));
        V3 N2 = mul_mv(rot(xf2), h2->planes[index2].n);
        const float kTol = 0.005f;
        V3 N = cross(N1, N2);
        float L = dot(N, N);
        FL(1);
        if (L < kTol * kTol) {
            capsule_face_contact(m, xf1, s1, xf2, index2, h2, r2);
            if (m->pointCount == 2) return;
        }
        m->pointCount = 1;
        m->points[0].localNormal1 = mulT_mv(rot(xf1), N1);
        m->points[0].localPoint1 = xf_mulT(xf1, g.point1);
        m->points[0].localPoint2 = xf_mulT(xf2, g.point2);
        set_zero_key(&m->points[0]);
        return;
    }
    /* face query: hull 2 planes in the capsule frame (sat_edge_and_hull.cpp:29-52) */
    XF xf = xf_mulT_xf(xf1, xf2);
    uint32_t fIndex = 0;
    float fSep = -MAX_FLOAT;
    for (uint32_t i = 0; i < h2->faceCount; ++i) {
        PL plane = plane_mul(&xf, h2->planes[i]);
        V3 dn = neg(plane.n);
        /* b3Capsule::GetSupportVertex (capsule.h:50-57) */
        V3 sv = dot(dn, s1->c[0]) > dot(dn, s1->c[1]) ? s1->c[0] : s1->c[1];
        float sep = plane_dist(sv, plane);
        if (sep > fSep) { fIndex = i; fSep = sep; }
    }
    if (fSep > totalRadius) return;
    /* edge query (sat_edge_and_hull.cpp:99-143) */
    XF xe = xf_mulT_xf(xf2, xf1);
    V3 P1 = xf_mul(&xe, s1->c[0]);
    V3 Q1 = xf_mul(&xe, s1->c[1]);
    V3 E1 = sub(Q1, P1);
    V3 C2 = h2->centroid;
    uint32_t eIndex = 0;
    float eSep = -MAX_FLOAT;
    for (uint32_t i = 0; i < h2->edgeCount; i += 2) {
        const b3o_halfedge* e2 = h2->edges + i;
        const b3o_halfedge* t2 = h2->edges + i + 1;
        V3 P2 = h2->vertices[e2->origin];
        V3 Q2 = h2->vertices[t2->origin];
        V3 E2 = sub(Q2, P2);
        V3 U2 = h2->planes[e2->face].n;
        V3 V2 = h2->planes[t2->face].n;
        FL(1);
        if (dot(E1, U2) * dot(E1, V2) < 0.0f) {
            float L1 = length(E1);
            float sep;
            if (L1 < LINEAR_SLOP) {
                sep = -MAX_FLOAT;
            } else {
                float L2 = length(E2);
                if (L2 < LINEAR_SLOP) {
                    sep = -MAX_FLOAT;
                } else {
                    const float kTol = 0.005f;
                    V3 E1xE2 = cross(E1, E2);
                    float L = length(E1xE2);
                    FL(2);
                    if (L < kTol * L1 * L2) {
                        sep = -MAX_FLOAT;
                    } else {
                        FL(1);
                        V3 N = scale(1.0f / L, E1xE2);
                        if (dot(N, sub(P2, C2)) > 0.0f) N = neg(N);
                        sep = dot(N, sub(P2, P1));
                    }
                }
            }
            if (sep > eSep) { eSep = sep; eIndex = i; }
        }
    }
    if (eSep > totalRadius) return;
    const float kTol = 0.1f * LINEAR_SLOP;
    FL(1);
    if (eSep > fSep + kTol) capsule_edge_contact(m, xf1, s1, xf2, eIndex, h2);
    else capsule_face_contact(m, xf1, s1, xf2, fIndex, h2, r2);
}

/* ---------------------------------------------------------------- aabb3.h */
typedef struct { V3 lo, hi; } AABB;

static V3 vmin(V3 a, V3 b) { return v3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
static V3 vmax(V3 a, V3 b) { return v3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }
static V3 aabb_centroid(const AABB* a) { return scale(0.5f, add(a->lo, a->hi)); }          /* aabb3.h:100-103 */
static float comp(V3 v, uint32_t i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
static uint32_t aabb_longest_axis(const AABB* a) {                                        /* aabb3.h:136-154 */
    V3 c = aabb_centroid(a);
    V3 r = sub(a->hi, c);
    float mx = r.x;
    uint32_t i = 0;
    if (r.y > mx) { mx = r.y; i = 1; }
    if (r.z > mx) { mx = r.z; i = 2; }
    return i;
}
static int aabb_overlap(const AABB* a, const AABB* b) {                                   /* aabb3.h:247-251 */
    return (a->lo.x <= b->hi.x) && (a->lo.y <= b->hi.y) && (a->lo.z <= b->hi.z) &&
           (a->hi.x >= b->lo.x) && (a->hi.y >= b->lo.y) && (a->hi.z >= b->lo.z);
}
static void aabb_extend(AABB* a, float s) {                                               /* aabb3.h:85-90 */
    V3 r = v3(s, s, s);
    a->lo = sub(a->lo, r);
    a->hi = add(a->hi, r);
}

/* ---------------------------------------------------------------- static_tree.cpp */
typedef struct { AABB aabb; uint32_t child1, child2; } TNode;  /* child2 doubles as the leaf index */
#define NULL_NODE 0xFFFFFFFFu

static int sort_pred(const AABB* set, uint32_t axis, uint32_t a, uint32_t b) {             /* static_tree.cpp:35-40 */
    V3 c1 = aabb_centroid(set + a), c2 = aabb_centroid(set + b);
    return comp(c1, axis) < comp(c2, axis);
}

/* static_tree.cpp:42-67: Lomuto quicksort with the last element as pivot.  The recursion
 * is unrolled onto an explicit stack; the two sub-ranges are disjoint, so the order of
 * visiting them does not change the result. */
static void tree_sort(const AABB* set, uint32_t axis, uint32_t* ids, uint32_t count, uint32_t* stack) {
    uint32_t top = 0;
    stack[top++] = 0;
    stack[top++] = count;
    while (top) {
        uint32_t n = stack[--top];
        uint32_t* p = ids + stack[--top];
        if (n <= 1) continue;
        uint32_t pivot = p[n - 1];
        uint32_t low = 0;
        for (uint32_t i = 0; i < n - 1; ++i) {
            if (sort_pred(set, axis, p[i], pivot)) {
                uint32_t t = p[i]; p[i] = p[low]; p[low] = t;
                low++;
            }
        }
        p[n - 1] = p[low];
        p[low] = pivot;
        stack[top++] = (uint32_t)(p - ids); stack[top++] = low;
        stack[top++] = (uint32_t)(p - ids) + low + 1; stack[top++] = n - 1 - low;
    }
}

static uint32_t tree_partition(const AABB* setAABB, const AABB* set, uint32_t* ids, uint32_t count,
                               uint32_t* stack) {                                          /* static_tree.cpp:69-105 */
    uint32_t axis = aabb_longest_axis(setAABB);
    float splitPos = comp(aabb_centroid(setAABB), axis);
    tree_sort(set, axis, ids, count, stack);
    uint32_t left = 0, right = count - 1, middle = left;
    while (middle < right) {
        V3 c = aabb_centroid(set + ids[middle]);
        if (comp(c, axis) > splitPos) break;
        ++middle;
    }
    uint32_t c1 = middle, c2 = count - middle;
    if (c1 == 0 || c2 == 0) middle = (left + right) / 2;
    return middle;
}

/* b3StaticTree::Build (static_tree.cpp:107-180), one object per leaf; nodes must hold
 * 2*count - 1 entries.  Recursion unrolled onto an explicit work list that preserves the
 * node numbering (children are numbered when their parent is split). */
void b3o_tree_build(const float* aabbs, uint32_t count, float* nodesOut /* [2n-1][8] */) {
    const AABB* set = (const AABB*)aabbs;
    TNode* nodes = (TNode*)nodesOut;
    uint32_t* ids = (uint32_t*)malloc(sizeof(uint32_t) * count);
    uint32_t* sortStack = (uint32_t*)malloc(sizeof(uint32_t) * (4 * (size_t)count + 8));
    uint32_t* work = (uint32_t*)malloc(sizeof(uint32_t) * (6 * (size_t)count + 8));
    for (uint32_t i = 0; i < count; ++i) ids[i] = i;
    uint32_t nodeCount = 1, top = 0;
    /* work item: node, first id, id count */
    work[top++] = 0; work[top++] = 0; work[top++] = count;
    while (top) {
        uint32_t n = work[--top], first = work[--top], node = work[--top];
        uint32_t* p = ids + first;
        AABB setAABB = set[p[0]];
        for (uint32_t i = 1; i < n; ++i) {
            setAABB.lo = vmin(setAABB.lo, set[p[i]].lo);
            setAABB.hi = vmax(setAABB.hi, set[p[i]].hi);
        }
        nodes[node].aabb = setAABB;
        if (n <= 1) {
            nodes[node].child1 = NULL_NODE;
            nodes[node].child2 = p[0];
        } else {
            uint32_t middle = tree_partition(&setAABB, set, p, n, sortStack);
            uint32_t c1 = nodeCount++, c2 = nodeCount++;
            nodes[node].child1 = c1;
            nodes[node].child2 = c2;
            /* The reference builds child1's subtree (numbering its descendants) before
             * child2's: process child1 first, so push child2 below it. */
            work[top++] = c2; work[top++] = first + middle; work[top++] = n - middle;
            work[top++] = c1; work[top++] = first; work[top++] = middle;
        }
    }
    free(ids);
    free(sortStack);
    free(work);
}

/* b3StaticTree::QueryAABB (static_tree.h:100-142): DFS pushing child1 then child2. */
static uint32_t tree_query(const TNode* nodes, uint32_t nodeCount, const AABB* box, uint32_t* out, uint32_t cap) {
    uint32_t stack[512];
    uint32_t top = 0, found = 0;
    if (nodeCount == 0) return 0;
    stack[top++] = 0;
    while (top) {
        uint32_t ni = stack[--top];
        const TNode* nd = nodes + ni;
        if (aabb_overlap(&nd->aabb, box)) {
            if (nd->child1 == NULL_NODE) {
                if (found < cap) out[found] = nd->child2;
                ++found;
            } else if (top + 2 <= 512) {
                stack[top++] = nd->child1;
                stack[top++] = nd->child2;
            }
        }
    }
    return found;
}

/* ---------------------------------------------------------------- shape AABBs */
static AABB shape_aabb(const b3o_hull_pool* pool, const b3o_shape* s, const XF* xf) {
    AABB a;
    if (s->type == B3O_SPHERE) {                                                           /* sphere_shape.cpp:53-59 */
        V3 c = xf_mul(xf, v3(s->v[0], s->v[1], s->v[2]));
        V3 r = v3(s->radius, s->radius, s->radius);
        a.lo = sub(c, r);
        a.hi = add(c, r);
    } else if (s->type == B3O_CAPSULE) {                                                   /* capsule_shape.cpp:130-137 */
        V3 c1 = xf_mul(xf, v3(s->v[0], s->v[1], s->v[2]));
        V3 c2 = xf_mul(xf, v3(s->v[3], s->v[4], s->v[5]));
        V3 r = v3(s->radius, s->radius, s->radius);
        a.lo = sub(vmin(c1, c2), r);
        a.hi = add(vmax(c1, c2), r);
    } else {                                                                               /* hull_shape.cpp:162-166 */
        Hull h = hull_at(pool, s->hull);
        a.lo = a.hi = xf_mul(xf, h.vertices[0]);
        for (uint32_t i = 1; i < h.vertexCount; ++i) {
            V3 v = xf_mul(xf, h.vertices[i]);
            a.lo = vmin(a.lo, v);
            a.hi = vmax(a.hi, v);
        }
        aabb_extend(&a, s->radius);
    }
    return a;
}

/* ---------------------------------------------------------------- triangle_hull.h:38-71 */
typedef struct {
    V3 v[3];
    b3o_halfedge e[6];
    uint32_t f[2];
    PL p[2];
} TriRec;

static Hull triangle_hull(TriRec* t, V3 A, V3 B, V3 C, V3* centroid) {
    V3 s = add(add(A, B), C);
    FL(3);
    *centroid = v3(s.x / 3.0f, s.y / 3.0f, s.z / 3.0f);
    t->v[0] = A; t->v[1] = B; t->v[2] = C;
    static const uint32_t tab[6][4] = {{0, 1, 0, 2}, {1, 0, 1, 3}, {1, 3, 0, 4}, {2, 2, 1, 5}, {2, 5, 0, 0}, {0, 4, 1, 1}};
    for (int i = 0; i < 6; ++i) {
        t->e[i].origin = tab[i][0]; t->e[i].twin = tab[i][1]; t->e[i].face = tab[i][2]; t->e[i].next = tab[i][3];
    }
    t->f[0] = 0;
    t->f[1] = 1;
    V3 N = normalize(cross(sub(B, A), sub(C, A)));
    t->p[0].n = N;
    t->p[0].d = dot(N, *centroid);
    t->p[1].n = neg(N);
    t->p[1].d = dot(neg(N), *centroid);
    Hull h;
    h.centroid = *centroid;
    h.vertexCount = 3;
    h.vertices = t->v;
    h.edgeCount = 6;
    h.edges = t->e;
    h.faceCount = 2;
    h.faces = t->f;
    h.planes = t->p;
    return h;
}

/* ---------------------------------------------------------------- manifold.cpp:57-69 */
typedef struct { V3 point, normal; float separation; } WPoint;
static WPoint world_point(const b3o_mpoint* p, float rA, const XF* xfA, float rB, const XF* xfB) {
    WPoint w;
    V3 nA = mul_mv(rot(xfA), p->localNormal1);
    V3 cA = xf_mul(xfA, p->localPoint1);
    V3 cB = xf_mul(xfB, p->localPoint2);
    V3 pA = add(cA, scale(rA, nA));
    V3 pB = sub(cB, scale(rB, nA));
    w.point = scale(0.5f, add(pA, pB));
    w.normal = nA;
    FL(2);
    w.separation = dot(sub(cB, cA), nA) - rA - rB;
    return w;
}

/* ---------------------------------------------------------------- contact_cluster.cpp:202-590 */
#define OBS_CAP 65536
typedef struct { V3 point; uint32_t cluster, manifold, manifoldPoint; } Obs;
typedef struct { Obs obs[OBS_CAP]; uint32_t nObs; V3 clusters[OBS_CAP]; uint32_t nClusters; } Cluster;

static float dist_sq(V3 a, V3 b) { return length_sq(sub(a, b)); }                         /* vec3.h:207-211 */

static uint32_t find_cluster(const Cluster* c, V3 p) {                                    /* :361-377 */
    uint32_t best = 0;
    float bestValue = MAX_FLOAT;
    for (uint32_t i = 0; i < c->nClusters; ++i) {
        float m = dist_sq(p, c->clusters[i]);
        if (m < bestValue) { bestValue = m; best = i; }
    }
    return best;
}

static void add_cluster(Cluster* c, V3 centroid) {                                        /* :336-359 */
    if (c->nClusters == 0) { c->clusters[c->nClusters++] = centroid; return; }
    uint32_t b = find_cluster(c, centroid);
    const float kTol = 0.05f;
    FL(1);
    if (dist_sq(centroid, c->clusters[b]) <= kTol * kTol) {
        c->clusters[b] = add(c->clusters[b], centroid);
        normalize_in_place(&c->clusters[b]);
        return;
    }
    c->clusters[c->nClusters++] = centroid;
}

static void initialize_clusters(Cluster* c) {                                             /* :212-334 */
    if (c->nObs <= 3) {
        for (uint32_t i = 0; i < c->nObs; ++i) add_cluster(c, c->obs[i].point);
        return;
    }
    static char chosen[OBS_CAP];
    memset(chosen, 0, c->nObs);
    add_cluster(c, c->obs[0].point);
    chosen[0] = 1;
    {
        V3 A = c->clusters[0];
        uint32_t index = 0;
        float mx = -MAX_FLOAT;
        for (uint32_t i = 0; i < c->nObs; ++i) {
            if (chosen[i]) continue;
            float dd = dist_sq(A, c->obs[i].point);
            if (dd > mx) { mx = dd; index = i; }
        }
        add_cluster(c, c->obs[index].point);
        chosen[index] = 1;
    }
    if (c->nClusters == 1) {
        V3 A = c->clusters[0];
        uint32_t index = 0;
        float mx = -MAX_FLOAT;
        for (uint32_t i = 0; i < c->nObs; ++i) {
            if (chosen[i]) continue;
            float dd = dist_sq(A, c->obs[i].point);
            if (dd > mx) { mx = dd; index = i; }
        }
        add_cluster(c, c->obs[index].point);
        chosen[index] = 1;
        return;
    }
    {
        V3 A = c->clusters[0], B = c->clusters[1];
        uint32_t index = 0;
        float mx = -MAX_FLOAT;
        for (uint32_t i = 0; i < c->nObs; ++i) {
            if (chosen[i]) continue;
            V3 C = c->obs[i].point;
            /* b3ClosestPointOnSegment (geometry.h:174-193) */
            float w[3];
            barycentric2(w, A, B, C);
            V3 Q;
            if (w[1] <= 0.0f) Q = A;
            else if (w[0] <= 0.0f) Q = B;
            else {
                FL(1);
                float s = 1.0f / w[2];
                FL(2);
                float wA = s * w[0], wB = s * w[1];
                Q = add(scale(wA, A), scale(wB, B));
            }
            float dd = dist_sq(C, Q);
            if (dd > mx) { mx = dd; index = i; }
        }
        add_cluster(c, c->obs[index].point);
        chosen[index] = 1;
    }
}

static void cluster_solve(Cluster* c) {                                                   /* :392-473 */
    initialize_clusters(c);
    for (uint32_t iter = 0; iter < 10; ++iter) {
        for (uint32_t i = 0; i < c->nObs; ++i) c->obs[i].cluster = find_cluster(c, c->obs[i].point);
        for (uint32_t i = 0; i < c->nClusters; ++i) {
            V3 centroid = v3(0.0f, 0.0f, 0.0f);
            uint32_t count = 0;
            for (uint32_t j = 0; j < c->nObs; ++j)
                if (c->obs[j].cluster == i) { centroid = add(centroid, c->obs[j].point); ++count; }
            if (count > 0) {
                /* centroid /= float(count): multiply by the reciprocal (vec3.h:70-76) */
                FL(4);
                float s = 1.0f / (float)count;
                centroid = v3(centroid.x * s, centroid.y * s, centroid.z * s);
                c->clusters[i] = centroid;
            }
        }
    }
    uint32_t used = 0;
    for (uint32_t i = 0; i < c->nClusters; ++i) {
        uint32_t cnt = 0;
        for (uint32_t j = 0; j < c->nObs; ++j) if (c->obs[j].cluster == i) ++cnt;
        if (cnt > 0) {
            for (uint32_t j = 0; j < c->nObs; ++j) if (c->obs[j].cluster == i) c->obs[j].cluster = used;
            c->clusters[used++] = c->clusters[i];
        }
    }
    c->nClusters = used;
}

/* b3ClusterSolver::Run (contact_cluster.cpp:475-590).  Returns the manifold count. */
static uint32_t cluster_run(b3o_manifold out[3], const b3o_manifold* in, uint32_t numIn,
                            const XF* xfA, float rA, const XF* xfB, float rB) {
    static Cluster c;
    c.nObs = 0;
    c.nClusters = 0;
    for (uint32_t i = 0; i < numIn; ++i)
        for (uint32_t j = 0; j < in[i].pointCount; ++j) {
            WPoint w = world_point(in[i].points + j, rA, xfA, rB, xfB);
            if (c.nObs < OBS_CAP) {
                c.obs[c.nObs].point = w.normal;
                c.obs[c.nObs].cluster = 0xFFFFFFFFu;
                c.obs[c.nObs].manifold = i;
                c.obs[c.nObs].manifoldPoint = j;
                ++c.nObs;
            }
        }
    cluster_solve(&c);
    uint32_t numOut = 0;
    static CPoly poly, quad;
    for (uint32_t i = 0; i < c.nClusters; ++i) {
        V3 normal = v3(0.0f, 0.0f, 0.0f);
        poly.count = 0;
        for (uint32_t j = 0; j < c.nObs; ++j) {
            if (c.obs[j].cluster != i) continue;
            WPoint w = world_point(in[c.obs[j].manifold].points + c.obs[j].manifoldPoint, rA, xfA, rB, xfB);
            if (poly.count < POLY_CAP) {
                poly.v[poly.count].position = w.point;
                poly.v[poly.count].clipIndex = j;
                ++poly.count;
            }
            normal = add(normal, c.obs[j].point);
        }
        if (poly.count == 0) continue;
        normalize_in_place(&normal);
        b3o_manifold* m = out + numOut++;
        uint32_t minIndex = 0;
        float minSep = MAX_FLOAT;
        for (uint32_t j = 0; j < poly.count; ++j) {
            const Obs* o = c.obs + poly.v[j].clipIndex;
            WPoint w = world_point(in[o->manifold].points + o->manifoldPoint, rA, xfA, rB, xfB);
            if (w.separation < minSep) { minIndex = j; minSep = w.separation; }
            poly.v[j].position = add(poly.v[j].position, scale(w.separation, normal));
        }
        reduce_polygon(&quad, &poly, normal, minIndex);
        for (uint32_t j = 0; j < quad.count; ++j) {
            const Obs* o = c.obs + quad.v[j].clipIndex;
            m->points[j] = in[o->manifold].points[o->manifoldPoint];
        }
        m->pointCount = quad.count;
    }
    return numOut;
}

/* ---------------------------------------------------------------- mesh_contact.cpp:30-275 */
/* b3MeshContact::Collide (mesh_contact.cpp:217-275) over `count` triangles: per triangle a
 * b3TriangleHull, b3CollideShapeAndShape with the triangle's cache (caches may be NULL: fresh
 * empty caches), the triangle index stamped into the keys; then b3ClusterSolver::Run.  The
 * triangle hull belongs to the mesh body: the cached face rebuild reads body A's sweep
 * (q10, q1) and the mesh body's (q20, q2). */
static uint32_t mesh_collide_contact(const b3o_hull_pool* pool, const b3o_shape* sa, const b3o_mesh* mesh,
                                     const XF* xfA, const XF* xfB, const uint32_t* tris, uint32_t count,
                                     b3o_convex_cache* caches, QT q10, QT q1, QT q20, QT q2, b3o_manifold out[3]) {
    static b3o_manifold temp[65536];
    for (uint32_t k = 0; k < count; ++k) {
        uint32_t ti = tris[k];
        const uint32_t* t = mesh->triangles + 3 * (size_t)ti;
        V3 a = v3(mesh->vertices[3 * t[0]], mesh->vertices[3 * t[0] + 1], mesh->vertices[3 * t[0] + 2]);
        V3 b = v3(mesh->vertices[3 * t[1]], mesh->vertices[3 * t[1] + 1], mesh->vertices[3 * t[1] + 2]);
        V3 c = v3(mesh->vertices[3 * t[2]], mesh->vertices[3 * t[2] + 1], mesh->vertices[3 * t[2] + 2]);
        TriRec rec;
        V3 centroid;
        Hull th = triangle_hull(&rec, a, b, c, &centroid);
        b3o_manifold* m = temp + k;
        memset(m, 0, sizeof(*m));
        if (sa->type == B3O_HULL) {
            Hull h1 = hull_at(pool, sa->hull);
            b3o_feature_cache fresh;
            b3o_feature_cache* fc = caches ? &caches[k].feature : &fresh;
            if (!caches) {
                fresh.state = B3O_EMPTY;
                fresh.type = 0; fresh.index1 = 0; fresh.index2 = 0;
            }
            collide_hulls_cached(m, xfA, &h1, sa->radius, xfB, &th, 0.0f, fc, q10, q1, q20, q2);
        } else {
            Round r1 = round_of(sa);
            PHASE(1);
            if (sa->type == B3O_SPHERE) collide_sphere_hull(m, xfA, &r1, xfB, &th, 0.0f);
            else collide_capsule_hull(m, xfA, &r1, xfB, &th, 0.0f);
        }
        for (uint32_t j = 0; j < m->pointCount; ++j) m->points[j].triangleKey = ti;
    }
    PHASE(2);
    memset(out, 0, 3 * sizeof(b3o_manifold));
    return cluster_run(out, temp, count, xfA, sa->radius, xfB, 0.0f);
}

/* A freshly created b3MeshContact between shape A and the mesh: fat AABB of A in the mesh
 * frame extended by B3_AABB_EXTENSION (mesh_contact.cpp:30-60), FindNewPairs through the
 * static tree (:140-189, all caches empty), then Collide (:217-275). */
void b3o_collide_mesh(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* mesh, uint32_t n,
                      const uint32_t* shapeA, const b3o_xf* xfs, uint32_t cap, uint32_t* triCounts,
                      uint32_t* triLists, b3o_manifold* out, uint32_t* manifoldCounts) {
    const TNode* nodes = (const TNode*)mesh->nodes;
    static uint32_t tris[65536];
    static const QT qid = {0.0f, 0.0f, 0.0f, 1.0f};
    for (uint32_t i = 0; i < n; ++i) {
        const b3o_shape* sa = shapes + shapeA[i];
        const XF* xfA = xfs + 2 * (size_t)i;
        const XF* xfB = xfA + 1;
        XF xf = xf_mulT_xf(xfB, xfA);
        AABB box = shape_aabb(pool, sa, &xf);
        aabb_extend(&box, 0.2f);  /* B3_AABB_EXTENSION (settings.h:54) */
        uint32_t found = tree_query(nodes, mesh->nodeCount, &box, tris, 65536);
        if (found > 65536) found = 65536;
        triCounts[i] = found;
        for (uint32_t k = 0; k < found && k < cap; ++k) triLists[(size_t)i * cap + k] = tris[k];
        manifoldCounts[i] = mesh_collide_contact(pool, sa, mesh, xfA, xfB, tris, found, NULL, qid, qid, qid, qid,
                                                 out + 3 * (size_t)i);
    }
}

/* Persistent b3MeshContacts (mesh_contact.cpp:30-275) kept across steps: per contact the fat
 * AABB m_aabbA, the triangle list with one b3ConvexCache each (b3TriangleCache), grown
 * without bound as the reference's list is. */
typedef struct {
    AABB fat;
    uint32_t count, capacity;
    uint32_t* tris;
    b3o_convex_cache* caches;
} MeshContact;

struct b3o_mesh_contacts {
    const b3o_hull_pool* pool;
    const b3o_shape* shapes;
    const b3o_mesh* mesh;
    uint32_t n;
    uint32_t* shapeA;
    MeshContact* c;
};

/* FindNewPairs (:141-160) with Report (:162-189): the list restarts and every triangle's cache
 * starts empty (simplex count 0, feature state empty). */
static void mesh_find_new_pairs(b3o_mesh_contacts* h, MeshContact* mc) {
    static uint32_t tris[65536];
    uint32_t found = tree_query((const TNode*)h->mesh->nodes, h->mesh->nodeCount, &mc->fat, tris, 65536);
    if (found > 65536) found = 65536;
    if (found > mc->capacity) {
        mc->capacity = found;
        mc->tris = (uint32_t*)realloc(mc->tris, sizeof(uint32_t) * found);
        mc->caches = (b3o_convex_cache*)realloc(mc->caches, sizeof(b3o_convex_cache) * found);
    }
    mc->count = found;
    for (uint32_t k = 0; k < found; ++k) {
        mc->tris[k] = tris[k];
        memset(mc->caches + k, 0, sizeof(b3o_convex_cache));
        mc->caches[k].feature.state = B3O_EMPTY;
    }
}

b3o_mesh_contacts* b3o_mesh_contacts_create(const b3o_hull_pool* pool, const b3o_shape* shapes, const b3o_mesh* mesh,
                                            uint32_t n, const uint32_t* shapeA, const b3o_xf* xfs) {
    b3o_mesh_contacts* h = (b3o_mesh_contacts*)calloc(1, sizeof(b3o_mesh_contacts));
    h->pool = pool;
    h->shapes = shapes;
    h->mesh = mesh;
    h->n = n;
    h->shapeA = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    memcpy(h->shapeA, shapeA, sizeof(uint32_t) * n);
    h->c = (MeshContact*)calloc(n ? n : 1, sizeof(MeshContact));
    for (uint32_t i = 0; i < n; ++i) {
        /* Constructor (:30-55): fat AABB of A in the mesh frame; then FindNewPairs. */
        const XF* xfA = xfs + 2 * (size_t)i;
        XF xf = xf_mulT_xf(xfA + 1, xfA);
        h->c[i].fat = shape_aabb(pool, shapes + shapeA[i], &xf);
        aabb_extend(&h->c[i].fat, 0.2f);  /* B3_AABB_EXTENSION (settings.h:54) */
        mesh_find_new_pairs(h, h->c + i);
    }
    return h;
}

void b3o_mesh_contacts_collide(b3o_mesh_contacts* h, const b3o_xf* xfs, const b3o_quat* sweeps, b3o_manifold* out,
                               uint32_t* manifoldCounts) {
    static const QT qid = {0.0f, 0.0f, 0.0f, 1.0f};
    for (uint32_t i = 0; i < h->n; ++i) {
        const XF* xfA = xfs + 2 * (size_t)i;
        QT q10 = qid, q1 = qid, q20 = qid, q2 = qid;
        if (sweeps) {
            q10 = sweeps[4 * (size_t)i]; q1 = sweeps[4 * (size_t)i + 1];
            q20 = sweeps[4 * (size_t)i + 2]; q2 = sweeps[4 * (size_t)i + 3];
        }
        MeshContact* mc = h->c + i;
        manifoldCounts[i] = mesh_collide_contact(h->pool, h->shapes + h->shapeA[i], h->mesh, xfA, xfA + 1, mc->tris,
                                                 mc->count, mc->caches, q10, q1, q20, q2, out + 3 * (size_t)i);
    }
}

/* b3MeshContact::TestOverlap (mesh_contact.cpp:191-215): b3TestOverlap (collide.cpp:84-95)
 * of shape A against each listed triangle in list order with that triangle's cache (the
 * triangle proxy: its three vertices, radius B3_HULL_RADIUS = 0, collide.cpp:61-76), stopping
 * at the first overlap, so the caches of later triangles are left as they were. */
void b3o_mesh_contacts_test_overlap(b3o_mesh_contacts* h, const b3o_xf* xfs, uint32_t* overlap) {
    for (uint32_t i = 0; i < h->n; ++i) {
        const XF* xfA = xfs + 2 * (size_t)i;
        MeshContact* mc = h->c + i;
        Round ra;
        Hull ha;
        float rA;
        Proxy pA = shape_proxy(h->pool, h->shapes + h->shapeA[i], &ra, &ha, &rA);
        overlap[i] = 0;
        for (uint32_t k = 0; k < mc->count; ++k) {
            const uint32_t* t = h->mesh->triangles + 3 * (size_t)mc->tris[k];
            V3 tv[3];
            for (int j = 0; j < 3; ++j)
                tv[j] = v3(h->mesh->vertices[3 * t[j]], h->mesh->vertices[3 * t[j] + 1], h->mesh->vertices[3 * t[j] + 2]);
            Proxy pB = {tv, 3};
            GJKOut g = gjk_cached(xfA, &pA, xfA + 1, &pB, &mc->caches[k].simplex);
            if (overlap_of(g, rA, 0.0f)) {
                overlap[i] = 1;
                break;
            }
        }
    }
}

/* SynchronizeShapes + MoveAABB (:62-139) and FindNewPairs. */
void b3o_mesh_contacts_synchronize(b3o_mesh_contacts* h, const b3o_xf* xfs, const float* centers0) {
    for (uint32_t i = 0; i < h->n; ++i) {
        const XF* xfA = xfs + 2 * (size_t)i;
        MeshContact* mc = h->c + i;
        V3 d = sub(xfA->p, v3(centers0[3 * i], centers0[3 * i + 1], centers0[3 * i + 2]));
        XF xf = xf_mulT_xf(xfA + 1, xfA);
        AABB aabb = shape_aabb(h->pool, h->shapes + h->shapeA[i], &xf);
        /* b3AABB3::Contains (aabb3.h:156-167) */
        int inside = mc->fat.lo.x <= aabb.lo.x && aabb.lo.x <= mc->fat.hi.x && mc->fat.lo.y <= aabb.lo.y &&
                     aabb.lo.y <= mc->fat.hi.y && mc->fat.lo.z <= aabb.lo.z && aabb.lo.z <= mc->fat.hi.z &&
                     mc->fat.lo.x <= aabb.hi.x && aabb.hi.x <= mc->fat.hi.x && mc->fat.lo.y <= aabb.hi.y &&
                     aabb.hi.y <= mc->fat.hi.y && mc->fat.lo.z <= aabb.hi.z && aabb.hi.z <= mc->fat.hi.z;
        if (inside) continue;
        AABB fat = aabb;
        aabb_extend(&fat, 0.2f);  /* B3_AABB_EXTENSION */
        const float kMul = 2.0f;  /* B3_AABB_MULTIPLIER (settings.h:59) */
        if (d.x < 0.0f) fat.lo.x += kMul * d.x; else fat.hi.x += kMul * d.x;
        if (d.y < 0.0f) fat.lo.y += kMul * d.y; else fat.hi.y += kMul * d.y;
        if (d.z < 0.0f) fat.lo.z += kMul * d.z; else fat.hi.z += kMul * d.z;
        mc->fat = fat;
        mesh_find_new_pairs(h, mc);
    }
}

void b3o_mesh_contacts_read(b3o_mesh_contacts* h, uint32_t cap, uint32_t* triCounts, uint32_t* triLists,
                            b3o_convex_cache* caches, float* fat) {
    for (uint32_t i = 0; i < h->n; ++i) {
        const MeshContact* mc = h->c + i;
        triCounts[i] = mc->count;
        for (uint32_t k = 0; k < mc->count && k < cap; ++k) {
            triLists[(size_t)i * cap + k] = mc->tris[k];
            caches[(size_t)i * cap + k] = mc->caches[k];
        }
        memcpy(fat + 6 * (size_t)i, &mc->fat.lo, 12);
        memcpy(fat + 6 * (size_t)i + 3, &mc->fat.hi, 12);
    }
}

void b3o_mesh_contacts_destroy(b3o_mesh_contacts* h) {
    if (!h) return;
    for (uint32_t i = 0; i < h->n; ++i) {
        free(h->c[i].tris);
        free(h->c[i].caches);
    }
    free(h->c);
    free(h->shapeA);
    free(h);
}

/* Triangle AABBs (b3Mesh::GetTriangleAABB, mesh.h:80-86) and the static tree over them. */
void b3o_mesh_build_tree(const float* vertices, uint32_t triangleCount, const uint32_t* triangles, float* nodesOut) {
    AABB* boxes = (AABB*)malloc(sizeof(AABB) * (size_t)triangleCount);
    for (uint32_t i = 0; i < triangleCount; ++i) {
        const uint32_t* t = triangles + 3 * (size_t)i;
        V3 a = v3(vertices[3 * t[0]], vertices[3 * t[0] + 1], vertices[3 * t[0] + 2]);
        V3 b = v3(vertices[3 * t[1]], vertices[3 * t[1] + 1], vertices[3 * t[1] + 2]);
        V3 c = v3(vertices[3 * t[2]], vertices[3 * t[2] + 1], vertices[3 * t[2] + 2]);
        boxes[i].lo = vmin(a, vmin(b, c));                                                 /* aabb3.h:64-68 */
        boxes[i].hi = vmax(a, vmax(b, c));
    }
    b3o_tree_build((const float*)boxes, triangleCount, nodesOut);
    free(boxes);
}
/* ---------------------------------------------------------------- entry points */
static void fresh_cache(b3o_convex_cache* c) {
    memset(c, 0, sizeof(*c));
    c->feature.state = B3O_EMPTY;
}

void b3o_collide_pairs(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t n,
                       const uint32_t* pairs, const b3o_xf* xfs, const b3o_quat* sweeps,
                       b3o_convex_cache* caches, b3o_manifold* out) {
    static const QT qid = {0.0f, 0.0f, 0.0f, 1.0f};
    for (uint32_t i = 0; i < n; ++i) {
        const b3o_shape* sa = shapes + pairs[2 * i];
        const b3o_shape* sb = shapes + pairs[2 * i + 1];
        const XF* xa = xfs + 2 * (size_t)i;
        const XF* xb = xa + 1;
        b3o_convex_cache local;
        b3o_convex_cache* c = caches ? caches + i : &local;
        if (!caches) fresh_cache(&local);
        b3o_manifold* m = out + i;
        memset(m, 0, sizeof(*m));
        QT q10 = qid, q1 = qid, q20 = qid, q2 = qid;
        if (sweeps) {
            q10 = sweeps[4 * (size_t)i]; q1 = sweeps[4 * (size_t)i + 1];
            q20 = sweeps[4 * (size_t)i + 2]; q2 = sweeps[4 * (size_t)i + 3];
        }
        /* s_CollideMatrix (collide.cpp:172-177); typeA <= typeB */
        if (sa->type == B3O_HULL && sb->type == B3O_HULL) {
            Hull h1 = hull_at(pool, sa->hull), h2 = hull_at(pool, sb->hull);
            collide_hulls_cached(m, xa, &h1, sa->radius, xb, &h2, sb->radius, &c->feature, q10, q1, q20, q2);
        } else if (sb->type == B3O_HULL) {
            Round r1 = round_of(sa);
            Hull h2 = hull_at(pool, sb->hull);
            PHASE(1);
            if (sa->type == B3O_SPHERE) collide_sphere_hull(m, xa, &r1, xb, &h2, sb->radius);
            else collide_capsule_hull(m, xa, &r1, xb, &h2, sb->radius);
        } else {
            Round r1 = round_of(sa), r2 = round_of(sb);
            PHASE(1);
            if (sa->type == B3O_SPHERE && sb->type == B3O_SPHERE) collide_spheres(m, xa, &r1, xb, &r2);
            else if (sa->type == B3O_SPHERE) collide_sphere_capsule(m, xa, &r1, xb, &r2);
            else collide_capsules(m, xa, &r1, xb, &r2);
        }
    }
}

void b3o_test_overlap(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t n, const uint32_t* pairs,
                      const b3o_xf* xfs, b3o_convex_cache* caches, uint32_t* overlap) {
    for (uint32_t i = 0; i < n; ++i) {
        Round ra, rb;
        Hull ha, hb;
        float r1, r2;
        Proxy p1 = shape_proxy(pool, shapes + pairs[2 * i], &ra, &ha, &r1);
        Proxy p2 = shape_proxy(pool, shapes + pairs[2 * i + 1], &rb, &hb, &r2);
        const XF* xa = xfs + 2 * (size_t)i;
        b3o_convex_cache local;
        b3o_convex_cache* c = caches ? caches + i : &local;
        if (!caches) fresh_cache(&local);
        GJKOut g = gjk_cached(xa, &p1, xa + 1, &p2, &c->simplex);
        overlap[i] = (uint32_t)overlap_of(g, r1, r2);
    }
}

void b3o_hull_queries(const b3o_hull_pool* pool, uint32_t n, const uint32_t* hullPairs,
                      const b3o_xf* xfs, uint32_t* out) {
    for (uint32_t i = 0; i < n; ++i) {
        Hull h1 = hull_at(pool, hullPairs[2 * i]), h2 = hull_at(pool, hullPairs[2 * i + 1]);
        const XF* x1 = xfs + 2 * (size_t)i;
        const XF* x2 = x1 + 1;
        FaceQuery f1 = query_face(x1, &h1, x2, &h2);
        FaceQuery f2 = query_face(x2, &h2, x1, &h1);
        EdgeQuery e = query_edge(x1, &h1, x2, &h2);
        uint32_t* o = out + 8 * (size_t)i;
        o[0] = f1.index; memcpy(o + 1, &f1.separation, 4);
        o[2] = f2.index; memcpy(o + 3, &f2.separation, 4);
        o[4] = e.index1; o[5] = e.index2; memcpy(o + 6, &e.separation, 4);
        o[7] = 0;
    }
}

/* ---------------------------------------------------------------- the step after the narrowphase */
static int same_key(const b3o_mpoint* a, const b3o_mpoint* b) {  /* b3ManifoldPointKey::operator== (manifold.h:30-34) */
    return a->triangleKey == b->triangleKey && a->key[0] == b->key[0] && a->key[1] == b->key[1] &&
           a->key[2] == b->key[2] && a->key[3] == b->key[3];
}

static M33 quat_mat33(b3o_quat q) {  /* b3QuatMat33 (quat.h:306-318) */
    float x = q.x, y = q.y, z = q.z, w = q.w;
    float x2 = x + x, y2 = y + y, z2 = z + z;
    float xx = x * x2, xy = x * y2, xz = x * z2;
    float yy = y * y2, yz = y * z2, zz = z * z2;
    float wx = w * x2, wy = w * y2, wz = w * z2;
    M33 m;
    m.c[0] = v3(1.0f - (yy + zz), xy + wz, xz - wy);
    m.c[1] = v3(xy - wz, 1.0f - (xx + zz), yz + wx);
    m.c[2] = v3(xz + wy, yz - wx, 1.0f - (xx + yy));
    return m;
}

static M33 m33_of(const float a[9]) {
    M33 m;
    m.c[0] = v3(a[0], a[1], a[2]);
    m.c[1] = v3(a[3], a[4], a[5]);
    m.c[2] = v3(a[6], a[7], a[8]);
    return m;
}

static V3 v3o(b3o_vec3 a) { return v3(a.x, a.y, a.z); }
static b3o_vec3 o3(V3 a) { b3o_vec3 r = {a.x, a.y, a.z}; return r; }

void b3o_contact_init(uint32_t n, uint32_t M, const uint32_t* bodies, const b3o_solver_body* state,
                      const b3o_contact_material* material, const uint32_t* counts, const b3o_manifold* manifolds,
                      const uint32_t* oldCounts, const b3o_manifold* oldManifolds,
                      const b3o_manifold_impulses* oldImpulses, int warmStarting, b3o_manifold_impulses* impulses,
                      b3o_velocity_manifold* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t count = counts[i];
        /* b3Manifold::Initialize() of the new manifolds, then Initialize(old) of each old
         * manifold in turn: tangent and motor impulses end up from the last one (T21). */
        for (uint32_t j = 0; j < count; ++j) {
            b3o_manifold_impulses* imp = impulses + (size_t)i * M + j;
            memset(imp, 0, sizeof(*imp));
            if (!warmStarting) continue;
            const b3o_manifold* m2 = manifolds + (size_t)i * M + j;
            for (uint32_t k = 0; k < oldCounts[i]; ++k) {
                const b3o_manifold* m1 = oldManifolds + (size_t)i * M + k;
                const b3o_manifold_impulses* i1 = oldImpulses + (size_t)i * M + k;
                imp->tangentImpulse[0] = i1->tangentImpulse[0];
                imp->tangentImpulse[1] = i1->tangentImpulse[1];
                imp->motorImpulse = i1->motorImpulse;
                for (uint32_t a = 0; a < m2->pointCount; ++a)
                    for (uint32_t b = 0; b < m1->pointCount; ++b)
                        if (same_key(m2->points + a, m1->points + b)) {
                            imp->normalImpulse[a] = i1->normalImpulse[b];
                            imp->persisting[a] = 1;
                            break;
                        }
            }
        }
        const b3o_solver_body* A = state + bodies[2 * i];
        const b3o_solver_body* B = state + bodies[2 * i + 1];
        const b3o_contact_material* mt = material + i;
        const float mA = A->invMass, mB = B->invMass;
        const M33 iA = m33_of(A->invI), iB = m33_of(B->invI);
        const V3 xA = v3o(A->x), xB = v3o(B->x);
        const V3 vA = v3o(A->v), wA = v3o(A->w), vB = v3o(B->v), wB = v3o(B->w);
        XF xfA, xfB;
        M33 rAm = quat_mat33(A->q), rBm = quat_mat33(B->q);
        xfA.c[0] = rAm.c[0], xfA.c[1] = rAm.c[1], xfA.c[2] = rAm.c[2];
        xfA.p = sub(xA, mul_mv(&rAm, v3o(A->localCenter)));
        xfB.c[0] = rBm.c[0], xfB.c[1] = rBm.c[1], xfB.c[2] = rBm.c[2];
        xfB.p = sub(xB, mul_mv(&rBm, v3o(B->localCenter)));
        const float friction = sqrtf(mt->frictionA * mt->frictionB);                 /* b3MixFriction (contact_solver.h:93-96) */
        const float restitution = mt->restitutionA > mt->restitutionB ? mt->restitutionA : mt->restitutionB;  /* b3MixRestitution */
        for (uint32_t j = 0; j < count; ++j) {
            const b3o_manifold* m = manifolds + (size_t)i * M + j;
            const b3o_manifold_impulses* imp = impulses + (size_t)i * M + j;
            b3o_velocity_manifold* vcm = out + (size_t)i * M + j;
            memset(vcm, 0, sizeof(*vcm));
            vcm->pointCount = m->pointCount;
            vcm->tangentImpulse[0] = imp->tangentImpulse[0];
            vcm->tangentImpulse[1] = imp->tangentImpulse[1];
            vcm->motorImpulse = imp->motorImpulse;
            vcm->friction = friction;
            vcm->restitution = restitution;
            /* b3WorldManifold::Initialize (manifold.cpp:57-95) */
            V3 center = v3(0.0f, 0.0f, 0.0f), normal = v3(0.0f, 0.0f, 0.0f);
            V3 wpoint[4], wnormal[4];
            for (uint32_t k = 0; k < m->pointCount; ++k) {
                const b3o_mpoint* p = m->points + k;
                V3 nA = mul_mv(rot(&xfA), v3o(p->localNormal1));
                V3 cA = xf_mul(&xfA, v3o(p->localPoint1));
                V3 cB = xf_mul(&xfB, v3o(p->localPoint2));
                V3 pA = add(cA, scale(mt->radiusA, nA));
                V3 pB = sub(cB, scale(mt->radiusB, nA));
                wpoint[k] = scale(0.5f, add(pA, pB));
                wnormal[k] = nA;
                center = add(center, wpoint[k]);
                normal = add(normal, wnormal[k]);
            }
            if (m->pointCount == 0) continue;
            {
                float s = 1.0f / (float)m->pointCount;
                center = v3(center.x * s, center.y * s, center.z * s);
            }
            {
                float len = sqrtf(normal.x * normal.x + normal.y * normal.y + normal.z * normal.z);
                if (len > EPSILON) {
                    float s = 1.0f / len;
                    normal = v3(normal.x * s, normal.y * s, normal.z * s);
                }
            }
            V3 t1 = fabsf(normal.x) >= 0.57735027f ? v3(normal.y, -normal.x, 0.0f)
                                                  : v3(0.0f, normal.z, -normal.y);  /* b3Perp (vec3.h:255-267) */
            V3 t2 = cross(t1, normal);
            vcm->normal = o3(normal);
            vcm->tangent1 = o3(t1);
            vcm->tangent2 = o3(t2);
            for (uint32_t k = 0; k < m->pointCount; ++k) {
                b3o_velocity_point* vcp = vcm->points + k;
                V3 nrm = wnormal[k];
                V3 rA = sub(wpoint[k], xA), rB = sub(wpoint[k], xB);
                vcp->rA = o3(rA);
                vcp->rB = o3(rB);
                vcp->normal = o3(nrm);
                vcp->normalImpulse = imp->normalImpulse[k];
                V3 rnA = cross(rA, nrm), rnB = cross(rB, nrm);
                float K = mA + mB + dot(mul_mv(&iA, rnA), rnA) + dot(mul_mv(&iB, rnB), rnB);
                vcp->normalMass = K > 0.0f ? 1.0f / K : 0.0f;
                V3 dv = sub(sub(add(vB, cross(wB, rB)), vA), cross(wA, rA));
                float vn = dot(nrm, dv);
                vcp->velocityBias = 0.0f;
                if (vn < -1.0f) vcp->velocityBias = -restitution * vn;  /* B3_VELOCITY_THRESHOLD (settings.h:105) */
            }
            V3 rA = sub(center, xA), rB = sub(center, xB);
            vcm->rA = o3(rA);
            vcm->rB = o3(rB);
            V3 rn1A = cross(rA, t1), rn1B = cross(rB, t1), rn2A = cross(rA, t2), rn2B = cross(rB, t2);
            float k11 = mA + mB + dot(mul_mv(&iA, rn1A), rn1A) + dot(mul_mv(&iB, rn1B), rn1B);
            float k12 = dot(mul_mv(&iA, rn1A), rn2A) + dot(mul_mv(&iB, rn1B), rn2B);
            float k22 = mA + mB + dot(mul_mv(&iA, rn2A), rn2A) + dot(mul_mv(&iB, rn2B), rn2B);
            /* b3Inverse(b3Mat22) (mat22.cpp:77-93) of K = [x: (k11, k12), y: (k12, k22)] */
            float det = k11 * k22 - k12 * k12;
            if (det != 0.0f) det = 1.0f / det;
            vcm->tangentMass[0] = det * k22;
            vcm->tangentMass[1] = -det * k12;
            vcm->tangentMass[2] = -det * k12;
            vcm->tangentMass[3] = det * k11;
            M33 iAB;
            iAB.c[0] = add(iA.c[0], iB.c[0]), iAB.c[1] = add(iA.c[1], iB.c[1]), iAB.c[2] = add(iA.c[2], iB.c[2]);
            float mass = dot(normal, mul_mv(&iAB, normal));
            vcm->motorMass = mass > 0.0f ? 1.0f / mass : 0.0f;
        }
    }
}

/* ---------------------------------------------------------------- the step before: broadphase */
/* b3DynamicTree (dynamic_tree.cpp) as the reference builds it: node pool with its free list
 * (AllocateNode / FreeNode / AddToFreeList :41-100), surface-area insertion (FindBest
 * :128-197, InsertLeaf :199-249), removal (RemoveLeaf :251-292), no rebalancing; the AABB
 * query with an explicit stack (dynamic_tree.h:150-182).  b3BroadPhase (broad_phase.cpp,
 * broad_phase.h:160-208) on top of it. */
#define B3O_NULL_NODE 0xFFFFFFFFu
typedef struct {
    AABB aabb;
    uint32_t user;
    uint32_t parent; /* parent, or next free node while free (the reference's union) */
    uint32_t child1, child2;
    int32_t height;
} DNode;

struct b3o_broadphase {
    DNode* nodes;
    uint32_t root, nodeCount, nodeCapacity, freeList;
    uint32_t proxyCount;
    uint32_t* move;
    uint32_t moveCount, moveCapacity;
    uint32_t* pairs; /* [2 * pairCapacity] */
    uint32_t pairCount, pairCapacity;
    uint32_t queryId;
    uint32_t* unique; /* the last FindPairs' reported pairs, 4 words each */
    uint32_t uniqueCount, uniqueCapacity;
};

static float aabb_area(const AABB* a) {                                                   /* aabb3.h:106-133 */
    float w = a->hi.x - a->lo.x, h = a->hi.y - a->lo.y, d = a->hi.z - a->lo.z;
    return 2.0f * (w * d + w * h + d * h);
}
static AABB aabb_combine(const AABB* a, const AABB* b) {                                 /* aabb3.h:238-244 */
    AABB r;
    r.lo = vmin(a->lo, b->lo);
    r.hi = vmax(a->hi, b->hi);
    return r;
}
static int aabb_contains(const AABB* a, const AABB* b) {                                 /* aabb3.h:156-167 */
    return a->lo.x <= b->lo.x && b->lo.x <= a->hi.x && a->lo.y <= b->lo.y && b->lo.y <= a->hi.y &&
           a->lo.z <= b->lo.z && b->lo.z <= a->hi.z && a->lo.x <= b->hi.x && b->hi.x <= a->hi.x &&
           a->lo.y <= b->hi.y && b->hi.y <= a->hi.y && a->lo.z <= b->hi.z && b->hi.z <= a->hi.z;
}
static int dleaf(const DNode* n) { return n->child1 == B3O_NULL_NODE; }

static void dt_add_to_free_list(b3o_broadphase* t, uint32_t node) {                    /* dynamic_tree.cpp:87-100 */
    for (uint32_t i = node; i < t->nodeCapacity - 1; ++i) {
        t->nodes[i].parent = i + 1;
        t->nodes[i].height = -1;
    }
    t->nodes[t->nodeCapacity - 1].parent = B3O_NULL_NODE;
    t->nodes[t->nodeCapacity - 1].height = -1;
    t->freeList = node;
}

static uint32_t dt_allocate(b3o_broadphase* t) {                                       /* dynamic_tree.cpp:41-75 */
    if (t->freeList == B3O_NULL_NODE) {
        t->nodeCapacity *= 2;
        t->nodes = (DNode*)realloc(t->nodes, sizeof(DNode) * t->nodeCapacity);
        memset(t->nodes + t->nodeCount, 0, sizeof(DNode) * (t->nodeCapacity - t->nodeCount));
        dt_add_to_free_list(t, t->nodeCount);
    }
    uint32_t node = t->freeList;
    t->freeList = t->nodes[node].parent;
    t->nodes[node].parent = B3O_NULL_NODE;
    t->nodes[node].child1 = B3O_NULL_NODE;
    t->nodes[node].child2 = B3O_NULL_NODE;
    t->nodes[node].height = 0;
    t->nodes[node].user = 0;
    ++t->nodeCount;
    return node;
}

static void dt_free(b3o_broadphase* t, uint32_t node) {                                /* dynamic_tree.cpp:77-84 */
    t->nodes[node].parent = t->freeList;
    t->nodes[node].height = -1;
    t->freeList = node;
    --t->nodeCount;
}

static uint32_t dt_find_best(const b3o_broadphase* t, const AABB* leaf) {              /* dynamic_tree.cpp:128-197 */
    uint32_t index = t->root;
    while (!dleaf(t->nodes + index)) {
        const DNode* n = t->nodes + index;
        float branchArea = aabb_area(&n->aabb);
        AABB comb = aabb_combine(leaf, &n->aabb);
        float combinedArea = aabb_area(&comb);
        float branchCost = 2.0f * combinedArea;
        float inheritanceCost = 2.0f * (combinedArea - branchArea);
        uint32_t c1 = n->child1, c2 = n->child2;
        float cost1, cost2;
        AABB a1 = aabb_combine(leaf, &t->nodes[c1].aabb);
        if (dleaf(t->nodes + c1)) cost1 = aabb_area(&a1);
        else cost1 = (aabb_area(&a1) - aabb_area(&t->nodes[c1].aabb)) + inheritanceCost;
        AABB a2 = aabb_combine(leaf, &t->nodes[c2].aabb);
        if (dleaf(t->nodes + c2)) cost2 = aabb_area(&a2);
        else cost2 = (aabb_area(&a2) - aabb_area(&t->nodes[c2].aabb)) + inheritanceCost;
        if (branchCost < cost1 && branchCost < cost2) break;
        index = cost1 < cost2 ? c1 : c2;
    }
    return index;
}

static void dt_walk_back(b3o_broadphase* t, uint32_t node) {                           /* dynamic_tree.cpp:294-312 */
    while (node != B3O_NULL_NODE) {
        DNode* n = t->nodes + node;
        int32_t h1 = t->nodes[n->child1].height, h2 = t->nodes[n->child2].height;
        n->height = 1 + (h1 > h2 ? h1 : h2);
        n->aabb = aabb_combine(&t->nodes[n->child1].aabb, &t->nodes[n->child2].aabb);
        node = n->parent;
    }
}

static void dt_insert_leaf(b3o_broadphase* t, uint32_t leaf) {                         /* dynamic_tree.cpp:199-249 */
    if (t->root == B3O_NULL_NODE) {
        t->root = leaf;
        t->nodes[leaf].parent = B3O_NULL_NODE;
        return;
    }
    AABB leafAabb = t->nodes[leaf].aabb;
    uint32_t sibling = dt_find_best(t, &leafAabb);
    uint32_t oldParent = t->nodes[sibling].parent;
    uint32_t newParent = dt_allocate(t);
    DNode* p = t->nodes + newParent;
    p->parent = oldParent;
    p->child1 = sibling;
    t->nodes[sibling].parent = newParent;
    p->child2 = leaf;
    t->nodes[leaf].parent = newParent;
    p->user = 0;
    p->aabb = aabb_combine(&leafAabb, &t->nodes[sibling].aabb);
    p->height = t->nodes[sibling].height + 1;
    if (oldParent != B3O_NULL_NODE) {
        if (t->nodes[oldParent].child1 == sibling) t->nodes[oldParent].child1 = newParent;
        else t->nodes[oldParent].child2 = newParent;
    } else {
        t->root = newParent;
    }
    dt_walk_back(t, newParent);
}

static void dt_remove_leaf(b3o_broadphase* t, uint32_t leaf) {                         /* dynamic_tree.cpp:251-292 */
    if (leaf == t->root) {
        t->root = B3O_NULL_NODE;
        return;
    }
    uint32_t parent = t->nodes[leaf].parent;
    uint32_t grand = t->nodes[parent].parent;
    uint32_t sibling = t->nodes[parent].child1 == leaf ? t->nodes[parent].child2 : t->nodes[parent].child1;
    if (grand != B3O_NULL_NODE) {
        if (t->nodes[grand].child1 == parent) t->nodes[grand].child1 = sibling;
        else t->nodes[grand].child2 = sibling;
        t->nodes[sibling].parent = grand;
        dt_free(t, parent);
        dt_walk_back(t, grand);
    } else {
        t->root = sibling;
        t->nodes[sibling].parent = B3O_NULL_NODE;
        dt_free(t, parent);
    }
}

static AABB aabb_of(const b3o_aabb* a) {
    AABB r;
    r.lo = v3(a->lower.x, a->lower.y, a->lower.z);
    r.hi = v3(a->upper.x, a->upper.y, a->upper.z);
    return r;
}

static void bp_buffer_move(b3o_broadphase* bp, uint32_t id) {                           /* broad_phase.cpp:41-59 */
    if (bp->moveCount == bp->moveCapacity) {
        bp->moveCapacity *= 2;
        bp->move = (uint32_t*)realloc(bp->move, 4 * (size_t)bp->moveCapacity);
    }
    bp->move[bp->moveCount++] = id;
}

b3o_broadphase* b3o_bp_create(void) {                                                   /* dynamic_tree.cpp:22-35, broad_phase.cpp:21-34 */
    b3o_broadphase* bp = (b3o_broadphase*)calloc(1, sizeof(b3o_broadphase));
    bp->root = B3O_NULL_NODE;
    bp->nodeCapacity = 32;
    bp->nodes = (DNode*)calloc(bp->nodeCapacity, sizeof(DNode));
    dt_add_to_free_list(bp, 0);
    bp->moveCapacity = 16;
    bp->move = (uint32_t*)calloc(bp->moveCapacity, 4);
    bp->pairCapacity = 16;
    bp->pairs = (uint32_t*)calloc(2 * (size_t)bp->pairCapacity, 4);
    return bp;
}

void b3o_bp_destroy(b3o_broadphase* bp) {
    if (!bp) return;
    free(bp->nodes);
    free(bp->move);
    free(bp->pairs);
    free(bp->unique);
    free(bp);
}

uint32_t b3o_bp_create_proxy(b3o_broadphase* bp, const b3o_aabb* aabb, uint32_t user) {   /* broad_phase.cpp:78-90 */
    AABB fat = aabb_of(aabb);
    aabb_extend(&fat, 0.2f); /* B3_AABB_EXTENSION (settings.h:54) */
    uint32_t id = dt_allocate(bp);                                                       /* InsertNode :102-114 */
    bp->nodes[id].aabb = fat;
    bp->nodes[id].user = user;
    bp->nodes[id].height = 0;
    dt_insert_leaf(bp, id);
    ++bp->proxyCount;
    bp_buffer_move(bp, id);
    return id;
}

void b3o_bp_destroy_proxy(b3o_broadphase* bp, uint32_t id) {                             /* broad_phase.cpp:92-97 */
    for (uint32_t i = 0; i < bp->moveCount; ++i)                                          /* UnbufferMove :61-70 */
        if (bp->move[i] == id) bp->move[i] = B3O_NULL_NODE;
    --bp->proxyCount;
    dt_remove_leaf(bp, id);                                                               /* RemoveNode :116-122 */
    dt_free(bp, id);
}

void b3o_bp_touch_proxy(b3o_broadphase* bp, uint32_t id) { bp_buffer_move(bp, id); }   /* broad_phase.cpp:151-154 */

int b3o_bp_move_proxy(b3o_broadphase* bp, uint32_t id, const b3o_aabb* aabb, const b3o_vec3* displacement) { /* :99-149 */
    AABB a = aabb_of(aabb);
    if (aabb_contains(&bp->nodes[id].aabb, &a)) return 0;
    AABB fat = a;
    aabb_extend(&fat, 0.2f);
    V3 d = scale(2.0f, v3(displacement->x, displacement->y, displacement->z)); /* B3_AABB_MULTIPLIER (settings.h:59) */
    if (d.x < 0.0f) fat.lo.x += d.x; else fat.hi.x += d.x;
    if (d.y < 0.0f) fat.lo.y += d.y; else fat.hi.y += d.y;
    if (d.z < 0.0f) fat.lo.z += d.z; else fat.hi.z += d.z;
    dt_remove_leaf(bp, id);                                                                /* UpdateNode :116-126 */
    bp->nodes[id].aabb = fat;
    dt_insert_leaf(bp, id);
    bp_buffer_move(bp, id);
    return 1;
}

static void bp_report(b3o_broadphase* bp, uint32_t id) {                                /* broad_phase.cpp:156-182 */
    if (id == bp->queryId) return;
    if (bp->pairCount == bp->pairCapacity) {
        bp->pairCapacity *= 2;
        bp->pairs = (uint32_t*)realloc(bp->pairs, 8 * (size_t)bp->pairCapacity);
    }
    bp->pairs[2 * bp->pairCount] = id < bp->queryId ? id : bp->queryId;
    bp->pairs[2 * bp->pairCount + 1] = id > bp->queryId ? id : bp->queryId;
    ++bp->pairCount;
}

static int pair_cmp(const void* a, const void* b) {                                     /* broad_phase.h:136-149 */
    const uint32_t* p = (const uint32_t*)a;
    const uint32_t* q = (const uint32_t*)b;
    if (p[0] != q[0]) return p[0] < q[0] ? -1 : 1;
    if (p[1] != q[1]) return p[1] < q[1] ? -1 : 1;
    return 0;
}

uint32_t b3o_bp_find_pairs(b3o_broadphase* bp, uint32_t cap, uint32_t* out) {         /* broad_phase.h:160-208 */
    bp->pairCount = 0;
    uint32_t stackCap = 256, *stack = (uint32_t*)malloc(4 * (size_t)stackCap);
    for (uint32_t i = 0; i < bp->moveCount; ++i) {
        bp->queryId = bp->move[i];
        if (bp->queryId == B3O_NULL_NODE) continue;
        AABB q = bp->nodes[bp->queryId].aabb;
        uint32_t top = 0;                                                                  /* dynamic_tree.h:150-182 */
        stack[top++] = bp->root;
        while (top > 0) {
            uint32_t node = stack[--top];
            if (node == B3O_NULL_NODE) continue;
            const DNode* n = bp->nodes + node;
            if (aabb_overlap(&n->aabb, &q)) {
                if (dleaf(n)) {
                    bp_report(bp, node);
                } else {
                    if (top + 2 > stackCap) {
                        stackCap *= 2;
                        stack = (uint32_t*)realloc(stack, 4 * (size_t)stackCap);
                    }
                    stack[top++] = n->child1;
                    stack[top++] = n->child2;
                }
            }
        }
    }
    free(stack);
    bp->moveCount = 0;
    qsort(bp->pairs, bp->pairCount, 8, pair_cmp);
    bp->uniqueCount = 0;
    uint32_t index = 0;
    while (index < bp->pairCount) {
        const uint32_t* primary = bp->pairs + 2 * index;
        if (bp->uniqueCount == bp->uniqueCapacity) {
            bp->uniqueCapacity = bp->uniqueCapacity ? 2 * bp->uniqueCapacity : 64;
            bp->unique = (uint32_t*)realloc(bp->unique, 16 * (size_t)bp->uniqueCapacity);
        }
        uint32_t* u = bp->unique + 4 * (size_t)bp->uniqueCount++;
        u[0] = primary[0];
        u[1] = primary[1];
        u[2] = bp->nodes[primary[0]].user;
        u[3] = bp->nodes[primary[1]].user;
        ++index;
        while (index < bp->pairCount && bp->pairs[2 * index] == primary[0] && bp->pairs[2 * index + 1] == primary[1])
            ++index;
    }
    if (out) memcpy(out, bp->unique, 16 * (size_t)(bp->uniqueCount < cap ? bp->uniqueCount : cap));
    return bp->uniqueCount;
}

int b3o_bp_test_overlap(const b3o_broadphase* bp, uint32_t a, uint32_t b) {              /* broad_phase.cpp:73-76 */
    return aabb_overlap(&bp->nodes[a].aabb, &bp->nodes[b].aabb);
}

void b3o_bp_get_aabb(const b3o_broadphase* bp, uint32_t id, b3o_aabb* out) {
    const AABB* a = &bp->nodes[id].aabb;
    out->lower.x = a->lo.x, out->lower.y = a->lo.y, out->lower.z = a->lo.z;
    out->upper.x = a->hi.x, out->upper.y = a->hi.y, out->upper.z = a->hi.z;
}

/* b3Sweep::GetTransform(0.0f) (transform.h:224-234). */
static XF sweep_transform0(const b3o_sweep* s) {
    const float t = 0.0f, u = 1.0f - t;
    V3 c = v3(u * s->worldCenter0.x + t * s->worldCenter.x, u * s->worldCenter0.y + t * s->worldCenter.y,
              u * s->worldCenter0.z + t * s->worldCenter.z);
    b3o_quat q = {u * s->orientation0.x + t * s->orientation.x, u * s->orientation0.y + t * s->orientation.y,
                  u * s->orientation0.z + t * s->orientation.z, u * s->orientation0.w + t * s->orientation.w};
    float len = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);                  /* quat.h:88-100 */
    if (len > FLT_EPSILON) {
        float k = 1.0f / len;
        q.x *= k, q.y *= k, q.z *= k, q.w *= k;
    }
    XF xf;
    M33 R = quat_mat33(q);
    xf.c[0] = R.c[0], xf.c[1] = R.c[1], xf.c[2] = R.c[2];
    V3 qv = v3(q.x, q.y, q.z), lc = v3(s->localCenter.x, s->localCenter.y, s->localCenter.z);
    V3 tt = scale(2.0f, cross(qv, lc));                                                   /* b3Mul(q, v), quat.h:225-232 */
    V3 r = add(add(lc, scale(q.w, tt)), cross(qv, tt));
    xf.p = sub(c, r);
    return xf;
}

void b3o_shape_aabbs(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t n, const uint32_t* shapeIdx,
                     const uint32_t* bodies, const b3o_xf* xfs, const b3o_sweep* sweeps, b3o_aabb* out,
                     b3o_vec3* disp) {
    for (uint32_t i = 0; i < n; ++i) {
        const b3o_shape* s = shapes + shapeIdx[i];
        const XF* xf2 = xfs + bodies[i];
        AABB a = shape_aabb(pool, s, xf2);
        V3 d = v3(0.0f, 0.0f, 0.0f);
        if (sweeps) {                                                                       /* body.cpp:181-202 */
            XF xf1 = sweep_transform0(sweeps + bodies[i]);
            AABB a1 = shape_aabb(pool, s, &xf1);
            a = aabb_combine(&a1, &a);
            d = sub(xf2->p, xf1.p);
        }
        out[i].lower.x = a.lo.x, out[i].lower.y = a.lo.y, out[i].lower.z = a.lo.z;
        out[i].upper.x = a.hi.x, out[i].upper.y = a.hi.y, out[i].upper.z = a.hi.z;
        if (disp) disp[i].x = d.x, disp[i].y = d.y, disp[i].z = d.z;
    }
}

/* ---------------------------------------------------------------- the contact manager */
/* b3ContactManager (contact_manager.cpp) over a world whose bodies carry one convex shape
 * each and are moved by a script instead of the solver: AddPair (:34-142) from the
 * broadphase's pairs, UpdateContacts (:167-231) with b3Contact::Update's narrowphase
 * (contact.cpp:41-140, convex contacts only), Destroy (:270-311); b3Body::SynchronizeShapes
 * (body.cpp:181-202) after each move.  Body b's shape is its proxy's user data. */
typedef struct KContact KContact;
typedef struct KEdge { KContact* contact; uint32_t other; struct KEdge *prev, *next; } KEdge;
struct KContact {
    uint32_t shapeA, shapeB;
    KEdge edgeA, edgeB;
    KContact *prev, *next;
    b3o_convex_cache cache;
    b3o_manifold manifold;
};
struct b3o_kworld {
    b3o_hull_pool pool;
    b3o_shape* shapes;
    b3o_broadphase* bp;
    uint32_t bodyCount, bodyCap;
    uint32_t *shapeOf, *type, *proxy, *awake;
    KEdge** edges;
    b3o_xf* xf;
    b3o_sweep* sweep;
    KContact* head;
    uint32_t contactCount;
};

b3o_kworld* b3o_kworld_create(const b3o_hull_pool* pool, const b3o_shape* shapes, uint32_t shapeCount) {
    b3o_kworld* w = (b3o_kworld*)calloc(1, sizeof(b3o_kworld));
    w->pool = *pool;
    w->shapes = (b3o_shape*)malloc(sizeof(b3o_shape) * (shapeCount ? shapeCount : 1));
    memcpy(w->shapes, shapes, sizeof(b3o_shape) * shapeCount);
    w->bp = b3o_bp_create();
    return w;
}

static void kw_destroy_contact(b3o_kworld* w, KContact* c) {                          /* contact_manager.cpp:270-311 */
    KEdge* lists[2] = {&c->edgeA, &c->edgeB};
    uint32_t owners[2] = {c->shapeA, c->shapeB};
    for (int k = 0; k < 2; ++k) {
        KEdge* e = lists[k];
        if (e->prev) e->prev->next = e->next;
        else w->edges[owners[k]] = e->next;
        if (e->next) e->next->prev = e->prev;
    }
    if (c->prev) c->prev->next = c->next;
    else w->head = c->next;
    if (c->next) c->next->prev = c->prev;
    --w->contactCount;
    free(c);
}

void b3o_kworld_destroy(b3o_kworld* w) {
    if (!w) return;
    while (w->head) kw_destroy_contact(w, w->head);
    b3o_bp_destroy(w->bp);
    free(w->shapes), free(w->shapeOf), free(w->type), free(w->proxy), free(w->awake), free(w->edges);
    free(w->xf), free(w->sweep);
    free(w);
}

void b3o_kworld_add_bodies(b3o_kworld* w, uint32_t n, const uint32_t* shapeIdx, const uint32_t* types,
                           const b3o_xf* xfs, const b3o_sweep* sweeps, uint32_t* proxyIds) {
    if (w->bodyCount + n > w->bodyCap) {
        uint32_t cap = w->bodyCap ? w->bodyCap : 64;
        while (cap < w->bodyCount + n) cap *= 2;
        w->shapeOf = (uint32_t*)realloc(w->shapeOf, 4 * (size_t)cap);
        w->type = (uint32_t*)realloc(w->type, 4 * (size_t)cap);
        w->proxy = (uint32_t*)realloc(w->proxy, 4 * (size_t)cap);
        w->awake = (uint32_t*)realloc(w->awake, 4 * (size_t)cap);
        w->edges = (KEdge**)realloc(w->edges, sizeof(KEdge*) * cap);
        w->xf = (b3o_xf*)realloc(w->xf, sizeof(b3o_xf) * cap);
        w->sweep = (b3o_sweep*)realloc(w->sweep, sizeof(b3o_sweep) * cap);
        w->bodyCap = cap;
    }
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = w->bodyCount++;
        w->shapeOf[b] = shapeIdx[i];
        w->type[b] = types[i];
        w->awake[b] = 1; /* b3BodyDef::awake (body.h:52) */
        w->edges[b] = NULL;
        w->xf[b] = xfs[i];
        w->sweep[b] = sweeps[i];
        /* b3Body::CreateShape: the AABB at m_xf, then CreateProxy (body.cpp:110-118) */
        AABB a = shape_aabb(&w->pool, w->shapes + shapeIdx[i], xfs + i);
        b3o_aabb box = {{a.lo.x, a.lo.y, a.lo.z}, {a.hi.x, a.hi.y, a.hi.z}};
        w->proxy[b] = b3o_bp_create_proxy(w->bp, &box, b);
        if (proxyIds) proxyIds[i] = w->proxy[b];
    }
}

static int kw_should_collide(const b3o_kworld* w, uint32_t a, uint32_t b) {            /* body.cpp:204-225 */
    return w->type[a] == 2u || w->type[b] == 2u; /* e_dynamicBody; no joints in this world */
}

static void kw_add_pair(b3o_kworld* w, uint32_t sA, uint32_t sB) {                      /* contact_manager.cpp:34-142 */
    if (sA == sB) return;
    for (KEdge* ce = w->edges[sB]; ce; ce = ce->next)
        if (ce->other == sA) return; /* a contact already exists (either order) */
    if (!kw_should_collide(w, sA, sB)) return;
    /* Create: swap so that typeA <= typeB (:233-268) */
    if (w->shapes[w->shapeOf[sA]].type > w->shapes[w->shapeOf[sB]].type) {
        uint32_t t = sA;
        sA = sB;
        sB = t;
    }
    KContact* c = (KContact*)calloc(1, sizeof(KContact));
    c->shapeA = sA;
    c->shapeB = sB;
    fresh_cache(&c->cache); /* convex_contact.cpp:31-36 */
    c->edgeA.contact = c, c->edgeA.other = sB;
    c->edgeA.next = w->edges[sA];
    if (w->edges[sA]) w->edges[sA]->prev = &c->edgeA;
    w->edges[sA] = &c->edgeA;
    c->edgeB.contact = c, c->edgeB.other = sA;
    c->edgeB.next = w->edges[sB];
    if (w->edges[sB]) w->edges[sB]->prev = &c->edgeB;
    w->edges[sB] = &c->edgeB;
    w->awake[sA] = w->awake[sB] = 1;
    c->next = w->head;
    if (w->head) w->head->prev = c;
    w->head = c;
    ++w->contactCount;
}

uint32_t b3o_kworld_find_new_contacts(b3o_kworld* w) {                                  /* contact_manager.cpp:156-165 */
    uint32_t n = b3o_bp_find_pairs(w->bp, 0, NULL);
    uint32_t before = w->contactCount;
    for (uint32_t i = 0; i < n; ++i) kw_add_pair(w, w->bp->unique[4 * i + 2], w->bp->unique[4 * i + 3]);
    return w->contactCount - before;
}

void b3o_kworld_update_contacts(b3o_kworld* w, uint32_t stats[3]) {                     /* contact_manager.cpp:167-231 */
    uint32_t destroyed = 0, collided = 0, skipped = 0;
    KContact* c = w->head;
    while (c) {
        uint32_t a = c->shapeA, b = c->shapeB;
        if (!kw_should_collide(w, a, b)) {
            KContact* q = c;
            c = c->next;
            kw_destroy_contact(w, q);
            ++destroyed;
            continue;
        }
        int activeA = w->awake[a] && w->type[a] != 0u, activeB = w->awake[b] && w->type[b] != 0u;
        if (!activeA && !activeB) {
            c = c->next;
            ++skipped;
            continue;
        }
        if (!b3o_bp_test_overlap(w->bp, w->proxy[a], w->proxy[b])) {
            KContact* q = c;
            c = c->next;
            kw_destroy_contact(w, q);
            ++destroyed;
            continue;
        }
        /* b3Contact::Update -> b3ConvexContact::Collide (contact.cpp:63-78, convex_contact.cpp:50-62) */
        uint32_t pair[2] = {w->shapeOf[a], w->shapeOf[b]};
        b3o_xf xfs[2] = {w->xf[a], w->xf[b]};
        b3o_quat sw[4] = {w->sweep[a].orientation0, w->sweep[a].orientation, w->sweep[b].orientation0,
                          w->sweep[b].orientation};
        b3o_collide_pairs(&w->pool, w->shapes, 1, pair, xfs, sw, &c->cache, &c->manifold);
        ++collided;
        c = c->next;
    }
    if (stats) stats[0] = destroyed, stats[1] = collided, stats[2] = skipped;
}

void b3o_kworld_move_bodies(b3o_kworld* w, uint32_t n, const uint32_t* bodies, const b3o_xf* xfs,
                            const b3o_sweep* sweeps) {
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b = bodies[i];
        w->xf[b] = xfs[i];
        w->sweep[b] = sweeps[i];
        uint32_t s = w->shapeOf[b];
        b3o_aabb box;
        b3o_vec3 d;
        uint32_t zero = 0;
        b3o_shape_aabbs(&w->pool, w->shapes, 1, &s, &zero, w->xf + b, w->sweep + b, &box, &d);
        b3o_bp_move_proxy(w->bp, w->proxy[b], &box, &d);
    }
}

uint32_t b3o_kworld_contacts(const b3o_kworld* w, uint32_t cap, uint32_t* bodyPairs, b3o_manifold* out,
                             b3o_convex_cache* caches) {
    uint32_t k = 0;
    for (const KContact* c = w->head; c; c = c->next, ++k) {
        if (k >= cap) continue;
        if (bodyPairs) bodyPairs[2 * k] = c->shapeA, bodyPairs[2 * k + 1] = c->shapeB;
        if (out) out[k] = c->manifold;
        if (caches) caches[k] = c->cache;
    }
    return k;
}

void b3o_kworld_proxies(const b3o_kworld* w, uint32_t n, const uint32_t* bodies, uint32_t* ids, b3o_aabb* fat) {
    for (uint32_t i = 0; i < n; ++i) {
        if (ids) ids[i] = w->proxy[bodies[i]];
        if (fat) b3o_bp_get_aabb(w->bp, w->proxy[bodies[i]], fat + i);
    }
}
