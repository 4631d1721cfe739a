"""Shared test helpers: manifold comparison, FP64 checks, workloads."""
import numpy as np

from bounce_amd.types import MANIFOLD_DTYPE


def manifold_bytes(m):
    return np.ascontiguousarray(m).view(np.uint8).reshape(len(m), -1)


def assert_manifolds_equal(got, want, what="manifolds"):
    """Bit-exact: point count, feature keys, manifold order and all 9 floats per point."""
    assert got.dtype == MANIFOLD_DTYPE and want.dtype == MANIFOLD_DTYPE
    assert len(got) == len(want)
    eq = (manifold_bytes(got) == manifold_bytes(want)).all(1)
    if not eq.all():
        bad = np.where(~eq)[0]
        i = int(bad[0])
        raise AssertionError("%s: %d of %d differ; first %d:\n got  %r\n want %r" % (
            what, len(bad), len(eq), i, got[i], want[i]))


def xf_matrix(x):
    """(R [3,3], p [3]) in float64 from 12 floats (columns, position)."""
    x = np.asarray(x, np.float64)
    return np.stack([x[0:3], x[3:6], x[6:9]], 1), x[9:12]


def world_separations(m, xfa, xfb, ra=0.0, rb=0.0):
    """FP64 b3WorldManifoldPoint separations (manifold.cpp:57-69) of one manifold."""
    Ra, pa = xf_matrix(xfa)
    Rb, pb = xf_matrix(xfb)
    out = []
    for i in range(int(m["pointCount"])):
        p = m["points"][i]
        nA = Ra @ p["localNormal1"].astype(np.float64)
        cA = Ra @ p["localPoint1"].astype(np.float64) + pa
        cB = Rb @ p["localPoint2"].astype(np.float64) + pb
        out.append(float((cB - cA) @ nA - ra - rb))
    return out


def sat_separation_fp64(pool, ha, hb, xfa, xfb):
    """Brute-force FP64 separating-axis value for two hulls: the maximum over all face
    normals of both hulls and all edge-pair cross products (oriented away from A's
    centroid) of the projection gap.  Equals the SAT result the narrowphase computes
    (max of the face and edge queries) up to FP32 rounding and the Gauss-map filter."""
    va, ea, fa, pa_, ca = pool.hull(ha)
    vb, eb, fb, pb_, cb = pool.hull(hb)
    Ra, ta = xf_matrix(xfa)
    Rb, tb = xf_matrix(xfb)
    A = va.astype(np.float64) @ Ra.T + ta
    B = vb.astype(np.float64) @ Rb.T + tb
    cA = Ra @ ca.astype(np.float64) + ta
    best = -np.inf
    for n in pa_[:, :3].astype(np.float64) @ Ra.T:
        best = max(best, (B @ n).min() - (A @ n).max())
    for n in pb_[:, :3].astype(np.float64) @ Rb.T:
        best = max(best, (A @ n).min() - (B @ n).max())
    EA = A[ea[1::2, 0]] - A[ea[0::2, 0]]
    EB = B[eb[1::2, 0]] - B[eb[0::2, 0]]
    nA = pa_[:, :3].astype(np.float64) @ Ra.T
    nB = pb_[:, :3].astype(np.float64) @ Rb.T
    UB, VB = nB[eb[0::2, 2]], nB[eb[1::2, 2]]
    for i in range(len(EA)):
        # Gauss-map arcs (U1, V1) and (-U2, -V2) must intersect (sat.cpp:60-71).
        U1, V1 = nA[ea[2 * i, 2]], nA[ea[2 * i + 1, 2]]
        ADC, BDC = -(EB @ U1), -(EB @ V1)
        CBA, DBA = UB @ EA[i], VB @ EA[i]
        mink = (CBA * DBA < 0) & (ADC * BDC < 0) & (CBA * BDC > 0)
        N = np.cross(EA[i], EB)
        L = np.linalg.norm(N, axis=1)
        ok = mink & (L > 0.005 * np.linalg.norm(EA[i]) * np.linalg.norm(EB, axis=1))
        N = N[ok] / L[ok, None]
        s = np.sign((A[ea[2 * i, 0]] - cA) @ N.T)
        s[s == 0] = 1
        N = N * s[:, None]
        gaps = (B @ N.T).min(0) - (A @ N.T).max(0)
        if len(gaps):
            best = max(best, gaps.max())
    return best


def hull_workload(lib, seed, n_hulls, n_pairs, gap_mean=-0.0127, gap_sigma=0.05, nmin=8, nmax=64):
    """Pool, shapes, pairs and transforms of the batched hull-hull workload at a given size."""
    from bounce_amd.workloads import gen_hull_pool, gen_hull_pairs
    pool = gen_hull_pool(lib, seed, n_hulls, nmin, nmax)
    pairs, xfs = gen_hull_pairs(lib, seed + 1, pool, n_pairs, gap_mean, gap_sigma)
    return pool, pool.shapes(), pairs, xfs


def load_golden(name):
    """A fixture written by tests/golden/make_golden.py from the compiled reference."""
    import os
    from bounce_amd.types import HullPool
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))
    pool = HullPool(d["desc"], d["centroids"], d["vertices"], d["edges"], d["faces"], d["planes"])
    g = {k: d[k] for k in d.files}
    g["pool"] = pool
    return g


# The b3ConvexCache fields a hull-hull step defines: b3ConvexContact's constructor sets only
# the simplex count and the feature state (convex_contact.cpp:35-36) and hull pairs never
# touch the simplex cache, so its other words keep whatever the contact's memory held.
HULL_CACHE_FIELDS = ("count", "state", "ftype", "index1", "index2")


def defined_hull_cache(c):
    """Copy of hull-pair cache records with the words the reference leaves undefined set
    to zero: the simplex cache apart from its count, and the feature type and indices of
    an empty feature cache (ReadState ignores them, sat.cpp:177-185)."""
    from bounce_amd.types import CACHE_EMPTY
    c = c.copy()
    for f in ("metric", "iterations", "sindex1", "sindex2"):
        c[f] = 0
    empty = c["state"] == CACHE_EMPTY
    for f in ("ftype", "index1", "index2"):
        c[f][empty] = 0
    return c


def defined_simplex_cache(c):
    """Copy of cache records with only the b3SimplexCache words the sensor path defines: the
    count, and the metric and the first `count` vertex indices when the count is not 0
    (b3Simplex::WriteCache, gjk.cpp:796-804).  The reference never writes `iterations`, and a
    triangle cache that b3MeshContact::TestOverlap did not reach keeps Report's count 0 over
    uninitialised words (mesh_contact.cpp:182-183)."""
    c = c.copy()
    c["iterations"] = 0
    c["metric"][c["count"] == 0] = 0
    for k in range(4):
        unused = c["count"] <= k
        c["sindex1"][unused, k] = 0
        c["sindex2"][unused, k] = 0
    return c


def assert_caches_equal(got, want, what="caches", fields=None):
    """Field-wise b3ConvexCache equality (the two padding bytes of b3SimplexCache are
    not part of the record's value)."""
    for f in fields or got.dtype.names:
        if not np.array_equal(got[f], want[f]):
            bad = np.where((got[f] != want[f]).reshape(len(got), -1).any(1))[0]
            raise AssertionError("%s: field %s differs in %d records; first %d: got %r want %r" % (
                what, f, len(bad), bad[0], got[bad[0]], want[bad[0]]))


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def mixed_workload(lib, seed, n_pairs, n_hulls=64, gap_mean=-0.02, gap_sigma=0.08, kinds=None):
    """Sphere/capsule/hull pairs (typeA <= typeB, collide.cpp:172-177) placed near contact
    along a random direction.  Returns pool, shapes, pairs, xfs."""
    from bounce_amd.types import SHAPE_DTYPE, SPHERE, CAPSULE, HULL
    from bounce_amd.workloads import gen_hull_pool
    from bounce_amd.shapes import identity_xf
    rng = np.random.default_rng(seed)
    pool = gen_hull_pool(lib, seed, n_hulls, 8, 32, 0.2, 1.0)
    kinds = kinds or [(SPHERE, SPHERE), (SPHERE, CAPSULE), (SPHERE, HULL), (CAPSULE, CAPSULE),
                      (CAPSULE, HULL), (HULL, HULL)]
    shapes = np.zeros(2 * n_pairs, SHAPE_DTYPE)
    pairs = np.arange(2 * n_pairs, dtype=np.uint32).reshape(n_pairs, 2)
    xfs = np.zeros((n_pairs, 2, 12), np.float32)

    def make(k, s):
        s["type"] = k
        if k == SPHERE:
            s["radius"] = rng.uniform(0.1, 0.5)
            s["v"][:3] = rng.uniform(-0.2, 0.2, 3) if rng.random() < 0.5 else 0.0
        elif k == CAPSULE:
            L = rng.uniform(0.0, 1.0) if rng.random() < 0.9 else 0.0
            s["radius"] = rng.uniform(0.1, 0.3)
            s["v"][:3] = (0.0, -L / 2, 0.0)
            s["v"][3:] = (0.0, L / 2, 0.0)
        else:
            s["hull"] = rng.integers(0, pool.count)
            s["radius"] = 0.0

    def support(s, R, u):
        if s["type"] == HULL:
            v = pool.hull(int(s["hull"]))[0].astype(np.float64)
            return (v @ R.T @ u).max()
        pts = [s["v"][:3]] if s["type"] == SPHERE else [s["v"][:3], s["v"][3:]]
        return max(float((R @ np.asarray(p, np.float64)) @ u) for p in pts) + float(s["radius"])

    for i in range(n_pairs):
        ka, kb = kinds[rng.integers(0, len(kinds))]
        make(ka, shapes[2 * i])
        make(kb, shapes[2 * i + 1])
        Ra, Rb = random_rotation(rng), random_rotation(rng)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pa = rng.uniform(-5, 5, 3)
        t = support(shapes[2 * i], Ra, u) + support(shapes[2 * i + 1], Rb, -u) + rng.normal(gap_mean, gap_sigma)
        if rng.random() < 0.05:
            t = rng.uniform(0.0, 0.05)  # deep: centres almost coincide
        xfs[i, 0] = identity_xf(pa, Ra)
        xfs[i, 1] = identity_xf(pa + t * u, Rb)
    return pool, shapes, pairs, xfs


def special_convex_cases():
    """Hand-placed pairs for the branches random placement rarely reaches: parallel
    capsules (2-point clip, collide_capsules.cpp:208-243), a capsule resting on a box face
    (2-point face contact, collide_capsule_hull.cpp:180-191), deep sphere/capsule inside a
    hull (SAT branch), coincident sphere centres (default normal), zero-length capsules."""
    from bounce_amd.types import SHAPE_DTYPE, SPHERE, CAPSULE, HULL, HullPool
    from bounce_amd.shapes import box_hull, identity_xf
    pool = HullPool.concat([box_hull(1, 1, 1), box_hull(2, 0.5, 1.5)])
    rows = []  # (shapeA, shapeB, xfA, xfB)

    def sph(r, c=(0, 0, 0)):
        s = np.zeros(1, SHAPE_DTYPE)[0]
        s["type"], s["radius"], s["v"][:3] = SPHERE, r, c
        return s

    def cap(r, L):
        s = np.zeros(1, SHAPE_DTYPE)[0]
        s["type"], s["radius"] = CAPSULE, r
        s["v"][:3], s["v"][3:] = (0, -L / 2, 0), (0, L / 2, 0)
        return s

    def hull(h):
        s = np.zeros(1, SHAPE_DTYPE)[0]
        s["type"], s["hull"] = HULL, h
        return s

    rz = np.array([[0, -1, 0], [1, 0, 0], [0, 0, 1]], np.float64)  # capsule axis along x
    for dy in (0.3, 0.45, 0.55, 0.7):
        for dx in (0.0, 0.4, 1.2):
            rows.append((cap(0.25, 1.0), cap(0.3, 1.5), identity_xf((0, 0, 0)), identity_xf((dx, dy, 0.0))))
            rows.append((cap(0.25, 1.0), hull(0), identity_xf((dx, 1.0 + dy - 0.3, 0.0), rz), identity_xf((0, 0, 0))))
            rows.append((cap(0.2, 2.0), hull(1), identity_xf((dx, 0.5 + dy - 0.3, 0.2), rz), identity_xf((0, 0, 0))))
    for c in ((0, 0, 0), (0.3, 0.1, -0.2), (0.9, 0.0, 0.0), (0.5, 0.5, 0.5)):
        rows.append((sph(0.2), hull(0), identity_xf(c), identity_xf((0, 0, 0))))
        rows.append((cap(0.2, 0.5), hull(1), identity_xf(c), identity_xf((0, 0, 0))))
        rows.append((cap(0.2, 0.0), hull(0), identity_xf(c), identity_xf((0, 0, 0))))
    rows.append((sph(0.3), sph(0.4), identity_xf((1, 1, 1)), identity_xf((1, 1, 1))))
    rows.append((sph(0.3), cap(0.2, 0.0), identity_xf((0, 0, 0)), identity_xf((0.1, 0.2, 0))))
    rows.append((cap(0.3, 0.0), cap(0.3, 1.0), identity_xf((0, 0.2, 0)), identity_xf((0.1, 0, 0))))
    rows.append((cap(0.3, 1.0), cap(0.3, 1.0), identity_xf((0, 0, 0)), identity_xf((0.0, 0, 0))))
    shapes = np.zeros(2 * len(rows), SHAPE_DTYPE)
    xfs = np.zeros((len(rows), 2, 12), np.float32)
    for i, (a, b, xa, xb) in enumerate(rows):
        shapes[2 * i], shapes[2 * i + 1] = a, b
        xfs[i, 0], xfs[i, 1] = xa, xb
    pairs = np.arange(2 * len(rows), dtype=np.uint32).reshape(-1, 2)
    return pool, shapes, pairs, xfs


def closest_point_triangle(p, a, b, c):
    """Closest point to p on triangle abc (Ericson, Real-Time Collision Detection 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    denom = 1.0 / (va + vb + vc)
    return a + ab * (vb * denom) + ac * (vc * denom)


def point_hull_distance(pool, h, R, t, p):
    """FP64 distance from point p to hull h (triangulated) placed at (R, t); 0 inside."""
    v, e, f, pl, _ = pool.hull(h)
    V = v.astype(np.float64) @ R.T + t
    n = pl[:, :3].astype(np.float64) @ R.T
    d = pl[:, 3].astype(np.float64) + n @ t
    if (n @ p - d).max() <= 0:
        return 0.0
    best = np.inf
    for fi in range(len(f)):
        h0 = f[fi]
        h1 = e[h0, 3]
        h2 = e[h1, 3]
        q = closest_point_triangle(p, V[e[h0, 0]], V[e[h1, 0]], V[e[h2, 0]])
        best = min(best, np.linalg.norm(q - p))
    return best


def grid_mesh(H, W, cell=0.5, amplitude=2.0, seed=0, lattice=8):
    """Heightfield on a (H+1) x (W+1) vertex grid triangulated as b3GridMesh
    (grid_mesh.h:57-77), heights from seeded value noise with smoothstep interpolation."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(-1.0, 1.0, size=(H // lattice + 2, W // lattice + 2))
    verts = np.zeros(((H + 1) * (W + 1), 3), np.float32)
    for i in range(H + 1):
        for j in range(W + 1):
            u, v = i / lattice, j / lattice
            i0, j0 = int(u), int(v)
            fu, fv = u - i0, v - j0
            su, sv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
            h = (g[i0, j0] * (1 - su) * (1 - sv) + g[i0 + 1, j0] * su * (1 - sv) +
                 g[i0, j0 + 1] * (1 - su) * sv + g[i0 + 1, j0 + 1] * su * sv)
            verts[i * (W + 1) + j] = ((j - 0.5 * W) * cell, amplitude * h, (i - 0.5 * H) * cell)
    tris = []
    for i in range(H):
        for j in range(W):
            v1 = i * (W + 1) + j
            v2 = (i + 1) * (W + 1) + j
            v3 = (i + 1) * (W + 1) + (j + 1)
            v4 = i * (W + 1) + (j + 1)
            tris.append((v3, v2, v1))
            tris.append((v1, v4, v3))
    return verts, np.array(tris, np.uint32)


def mesh_workload(lib, seed, H, W, n_shapes, gap_sigma=0.05):
    """Spheres, capsules and 8..32-point hulls dropped over a grid mesh, each resting near
    the surface: (pool, shapes, vertices, triangles, xfs [n][2][12] (shape, mesh))."""
    from bounce_amd.types import SHAPE_DTYPE, SPHERE, CAPSULE, HULL
    from bounce_amd.workloads import gen_hull_pool
    from bounce_amd.shapes import identity_xf
    rng = np.random.default_rng(seed)
    verts, tris = grid_mesh(H, W, seed=seed)
    pool = gen_hull_pool(lib, seed, 64, 8, 32, 0.1, 0.5)
    shapes = np.zeros(n_shapes, SHAPE_DTYPE)
    xfs = np.zeros((n_shapes, 2, 12), np.float32)
    half = 0.5 * 0.5 * np.array([W, H])
    for i in range(n_shapes):
        k = rng.random()
        s = shapes[i]
        R = random_rotation(rng)
        if k < 0.4:
            s["type"], s["radius"] = SPHERE, rng.uniform(0.1, 0.5)
            low = float(s["radius"])
        elif k < 0.7:
            L = rng.uniform(0.2, 1.0)
            s["type"], s["radius"] = CAPSULE, rng.uniform(0.1, 0.3)
            s["v"][:3], s["v"][3:] = (0, -L / 2, 0), (0, L / 2, 0)
            low = float(s["radius"]) + abs(R[1, 1]) * L / 2
        else:
            s["type"], s["hull"] = HULL, rng.integers(0, pool.count)
            v = pool.hull(int(s["hull"]))[0].astype(np.float64)
            low = -(v @ R.T)[:, 1].min()
        x, z = rng.uniform(-half[0] + 1, half[0] - 1), rng.uniform(-half[1] + 1, half[1] - 1)
        near = np.linalg.norm(verts[:, [0, 2]] - (x, z), axis=1) < 0.4
        ground = verts[near, 1].mean() if near.any() else 0.0
        y = ground + low + rng.normal(0.0, gap_sigma) - 0.03
        xfs[i, 0] = identity_xf((x, y, z), R)
        xfs[i, 1] = identity_xf((0, 0, 0))
    return pool, shapes, verts, tris, xfs


def plate_workload(seed, n_shapes, H=24, W=24, cell=0.3, extents=None):
    """Flat boxes (half-extents 0.18-0.25 m in x and z, 0.1 m high) lying on a flat grid of
    0.3 m cells, turned about the vertical and tilted by at most 0.01 rad, sunk 1-10 mm
    into the plane: every triangle under a box gives a face contact of up to four points,
    all with one normal, so a contact's single cluster holds more than 32 points while its
    triangle list (the fattened AABB's footprint) stays within 64.  `extents` replaces the
    four (x, z) half-extents.
    Returns (pool, shapes, vertices, triangles, xfs [n][2][12] (box, mesh))."""
    from bounce_amd.shapes import box_hull, identity_xf
    from bounce_amd.types import HULL, SHAPE_DTYPE, HullPool
    rng = np.random.default_rng(seed)
    verts, tris = grid_mesh(H, W, cell=cell, amplitude=0.0, seed=seed)
    ext = extents or [(0.25, 0.2), (0.2, 0.2), (0.25, 0.25), (0.22, 0.18)]
    pool = HullPool.concat([box_hull(ex, 0.1, ez) for ex, ez in ext])
    shapes = np.zeros(n_shapes, SHAPE_DTYPE)
    shapes["type"] = HULL
    shapes["hull"] = rng.integers(0, len(ext), n_shapes)
    xfs = np.zeros((n_shapes, 2, 12), np.float32)
    half = 0.5 * cell * np.array([W, H])
    for i in range(n_shapes):
        yaw = rng.uniform(0.0, 2.0 * np.pi)
        c, s = np.cos(yaw), np.sin(yaw)
        R = np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])
        tx, tz = rng.uniform(-0.01, 0.01, 2)
        Rx = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(tx), -np.sin(tx)], [0.0, np.sin(tx), np.cos(tx)]])
        Rz = np.array([[np.cos(tz), -np.sin(tz), 0.0], [np.sin(tz), np.cos(tz), 0.0], [0.0, 0.0, 1.0]])
        x, z = rng.uniform(-half[0] + 1.0, half[0] - 1.0, 2)
        y = 0.1 - rng.uniform(0.001, 0.01)
        xfs[i, 0] = identity_xf((x, y, z), Rx @ Rz @ R)
        xfs[i, 1] = identity_xf((0, 0, 0))
    return pool, shapes, verts, tris, xfs


# ---- the box pyramid (210 unit boxes on a static ground) stepped by the reference world

PYRAMID_STEPS = 300
PYRAMID_FULL_STEPS = (60, 120, 180, 240, 299)  # steps whose manifolds the fixture keeps whole


def pyramid_shapes(body_count):
    """Shape table of the pyramid world: shape i belongs to body i; body 0 is the ground
    (hull 0), every other body a unit box (hull 1)."""
    from bounce_amd.types import HULL, SHAPE_DTYPE
    s = np.zeros(body_count, SHAPE_DTYPE)
    s["type"] = HULL
    s["hull"] = 1
    s["hull"][0] = 0
    return s


def step_digest(out, caches):
    """SHA-256 over a step's manifolds (all bytes) and the defined fields of its hull-pair
    cache records."""
    import hashlib
    h = hashlib.sha256()
    h.update(manifold_bytes(out).tobytes())
    for f in HULL_CACHE_FIELDS:
        h.update(np.ascontiguousarray(caches[f]).tobytes())
    return np.frombuffer(h.digest(), np.uint8)


class CacheChain:
    """b3ConvexCache records carried from step to step by contact id, as each
    b3ConvexContact keeps its m_cache (convex_contact.h); new contacts start empty
    (convex_contact.cpp:35-36)."""

    def __init__(self, capacity=1 << 14):
        from bounce_amd.types import empty_caches
        self.caches = empty_caches(capacity)

    def gather(self, ids):
        return np.ascontiguousarray(self.caches[ids])

    def scatter(self, ids, caches):
        self.caches[ids] = caches


def pyramid_inputs(trace, step):
    """(ids, body pairs, xfs [n][2][12], sweeps [n][4][4]) of one step of a pyramid fixture."""
    a, b = int(trace["step_offsets"][step]), int(trace["step_offsets"][step + 1])
    ids = trace["ids"][a:b].astype(np.int64)
    pairs = trace["id_pairs"][ids]
    xfs = trace["body_xfs"][step][pairs]
    sweeps = trace["body_sweeps"][step][pairs].reshape(-1, 4, 4)
    return ids, pairs, xfs, sweeps


def pyramid_full_outputs(trace, step):
    """(manifolds, caches) the fixture keeps whole for one of PYRAMID_FULL_STEPS."""
    k = list(int(s) for s in trace["full_steps"]).index(step)
    a, b = int(trace["full_offsets"][k]), int(trace["full_offsets"][k + 1])
    return trace["full_out"][a:b], trace["full_caches"][a:b]


def quat_matrix(q):
    """Rotation matrices [n][3][3] of unit quaternions [n][4] (x, y, z, w) in float64."""
    x, y, z, w = (q[:, i].astype(np.float64) for i in range(4))
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def moving_mesh_steps(xfs0, steps, seed, fast_fraction=0.15):
    """Synthetic motion of the convex shapes of mesh contacts (body A; the mesh body stays
    at identity): per step a translation (mostly ~5 mm, a fast fraction ~8 cm, which leaves
    the 0.2 m fat AABB within a few steps) and a small rotation.  Returns, per step s,
    (xfs [n][2][12], sweeps [n][4][4] = (q_{s-1}, q_s, identity, identity), positions [n][3]);
    the orientations are quaternions and the transforms their float32 matrices."""
    rng = np.random.default_rng(seed)
    n = len(xfs0)
    R0 = np.stack([xfs0[:, 0, 0:3], xfs0[:, 0, 3:6], xfs0[:, 0, 6:9]], 2).astype(np.float64)
    # Quaternion of each initial rotation (Shepperd's method on the trace).
    q = np.zeros((n, 4))
    for i in range(n):
        m = R0[i]
        t = np.trace(m)
        if t > 0:
            s = np.sqrt(t + 1.0) * 2
            q[i] = [(m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s, 0.25 * s]
        else:
            k = int(np.argmax(np.diag(m)))
            a, b, c = k, (k + 1) % 3, (k + 2) % 3
            s = np.sqrt(1.0 + m[a, a] - m[b, b] - m[c, c]) * 2
            v = np.zeros(4)
            v[a] = 0.25 * s
            v[b] = (m[b, a] + m[a, b]) / s
            v[c] = (m[c, a] + m[a, c]) / s
            v[3] = (m[c, b] - m[b, c]) / s
            q[i] = v
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    pos = xfs0[:, 0, 9:12].astype(np.float64)
    speed = np.where(rng.random(n) < fast_fraction, 0.08, 0.005)
    vel = rng.normal(size=(n, 3))
    vel *= (speed / np.linalg.norm(vel, axis=1))[:, None]
    omega = rng.normal(0.0, 0.01, (n, 3))
    out = []
    for s in range(steps):
        q_prev = q.copy()
        if s > 0:
            pos = pos + vel
            dq = np.concatenate([0.5 * omega, np.ones((n, 1))], 1)
            x1, y1, z1, w1 = dq.T
            x2, y2, z2, w2 = q.T
            q = np.stack([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2, w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                          w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2, w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2], 1)
            q /= np.linalg.norm(q, axis=1, keepdims=True)
        R = quat_matrix(q)
        xfs = xfs0.copy()
        xfs[:, 0, 0:3], xfs[:, 0, 3:6], xfs[:, 0, 6:9] = R[:, :, 0], R[:, :, 1], R[:, :, 2]
        xfs[:, 0, 9:12] = pos
        sweeps = np.zeros((n, 4, 4), np.float32)
        sweeps[:, 0], sweeps[:, 1] = q_prev, q
        sweeps[:, 2, 3] = sweeps[:, 3, 3] = 1.0
        out.append((xfs.astype(np.float32), sweeps, pos.astype(np.float32)))
    return out


def check_persistent_mesh_fixture(g, contacts):
    """Replays tests/golden/mesh_contacts_persistent.npz through `contacts` (an object with
    collide / synchronize / read, created at the fixture's first transforms) and compares
    every step's manifolds and the triangle state after creation and every synchronize."""

    def state(k, what):
        c, l, ca, f = contacts.read()
        np.testing.assert_array_equal(c, g["tri_counts"][k], what)
        np.testing.assert_array_equal(f.view(np.uint32), g["fat"][k].view(np.uint32), what)
        for i in range(len(c)):
            m = min(int(c[i]), l.shape[1])
            np.testing.assert_array_equal(l[i, :m], g["tri_lists"][k][i, :m], what)
            assert_caches_equal(defined_hull_cache(ca[i, :m]), g["tri_caches"][k][i, :m], "%s contact %d" % (what, i))

    state(0, "after creation")
    for s in range(len(g["out"])):
        out, mc = contacts.collide(g["xfs"][s], g["sweeps"][s])
        np.testing.assert_array_equal(mc, g["manifold_counts"][s], "step %d" % s)
        assert_manifolds_equal(out.reshape(-1), g["out"][s].reshape(-1), "step %d manifolds" % s)
        contacts.synchronize(g["xfs"][s + 1], g["centers"][s])
        state(s + 1, "after synchronize %d" % s)


def warm_world_chain(checker, world, steps, idx=None, keep=None):
    """The listed contacts of a warm world (all by default) stepped through `checker`'s
    b3CollideShapeAndShape (oracle or compiled reference) for steps 0..steps-1, each
    contact's b3ConvexCache carried from step to step as b3ConvexContact keeps m_cache.
    Returns {step: (manifolds, caches)} for the steps in `keep` (default: the last)."""
    from bounce_amd.types import empty_caches
    from bounce_amd.workloads import warm_world_pair_inputs, warm_world_state
    contacts = world["contacts"] if idx is None else world["contacts"][idx]
    caches = empty_caches(len(contacts))
    keep = {steps - 1} if keep is None else set(keep)
    res = {}
    for s in range(steps):
        xfs, sweeps = warm_world_state(world, s)
        pairs, pxf, psw = warm_world_pair_inputs(world, xfs, sweeps, contacts)
        out = checker.collide(world["pool"], world["shapes"], pairs, pxf, sweeps=psw, caches=caches)
        if s in keep:
            res[s] = (out, caches.copy())
    return res


class WarmWorldDevice:
    """A warm world's body table, contact list, caches and manifolds in device buffers of
    a Context, stepped through b3amd_update_contacts."""

    def __init__(self, ctx, world):
        from bounce_amd.types import CACHE_DTYPE, MANIFOLD_DTYPE, empty_caches
        self.ctx, self.world = ctx, world
        ctx.upload_hulls(world["pool"])
        ctx.upload_shapes(world["shapes"])
        self.B = len(world["body_shapes"])
        self.n = len(world["contacts"])
        self.d_shapes = ctx.alloc(4 * self.B)
        self.d_shapes.upload(world["body_shapes"])
        self.d_contacts = ctx.alloc(8 * self.n)
        self.d_contacts.upload(world["contacts"])
        self.d_caches = ctx.alloc(CACHE_DTYPE.itemsize * self.n)
        self.d_caches.upload(empty_caches(self.n))
        self.d_out = ctx.alloc(MANIFOLD_DTYPE.itemsize * self.n)
        self.d_xfs = ctx.alloc(48 * self.B)
        self.d_sweeps = ctx.alloc(32 * self.B)

    def step(self, s, sweeps=True):
        from bounce_amd.workloads import warm_world_state
        xfs, sw = warm_world_state(self.world, s)
        self.d_xfs.upload(xfs)
        self.d_sweeps.upload(sw)
        self.ctx.update_contacts(self.B, self.d_shapes, self.d_xfs, self.n, self.d_contacts, self.d_out,
                                 d_body_sweeps=self.d_sweeps if sweeps else None, d_caches=self.d_caches)

    def read(self):
        from bounce_amd.types import CACHE_DTYPE, MANIFOLD_DTYPE
        self.ctx.synchronize()
        out = self.d_out.download(np.zeros(self.n, MANIFOLD_DTYPE))
        caches = self.d_caches.download(np.zeros(self.n, CACHE_DTYPE))
        return out, caches


def qhull_clouds(seed, n_ellipsoid=300, n_box=60, n_slab=30):
    """Point clouds for the hull builder: points on ellipsoids (4-64 points, semi-axes
    U[0.25, 2], the headline workload's), uniform in boxes (interior points), thin slabs
    (nearly coplanar), and fixed special sets: cube corners, a coplanar quad, points closer
    than LINEAR_SLOP (welded), a duplicated cloud."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_ellipsoid):
        n = int(rng.integers(4, 65))
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out.append(d * rng.uniform(0.25, 2.0, 3))
    for _ in range(n_box):
        out.append(rng.uniform(-1, 1, (int(rng.integers(4, 80)), 3)))
    for _ in range(n_slab):
        p = rng.uniform(-1, 1, (30, 3))
        p[:, 2] *= 1e-3
        out.append(p)
    out.append(np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], float))
    out.append(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0]], float))
    out.append(np.array([[0, 0, 0], [0, 0, 0.001], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float))
    out.append(np.concatenate([out[0], out[0]]))
    return [np.asarray(p, np.float32) for p in out]


def run_qhull(fn, points, simplify, returns_status=True):
    """(counts [3], vertices, edges, faces, planes, centroid) of b3QHull::Set on `points`
    through `fn` (the library's b3amd_quickhull or the reference's b3ref_qhull)."""
    import ctypes
    pts = np.ascontiguousarray(points, np.float32)
    m = 4 * max(len(pts), 4) + 16  # simplification can rebuild more vertices than it got
    counts = np.zeros(3, np.uint32)
    V, E = np.zeros((m, 3), np.float32), np.zeros((6 * m, 4), np.uint32)
    F, P, C = np.zeros(2 * m, np.uint32), np.zeros((2 * m, 4), np.float32), np.zeros(3, np.float32)
    fn.restype = ctypes.c_int if returns_status else None
    fn.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int] + [ctypes.c_void_p] * 6
    rc = fn(pts.ctypes.data, len(pts), int(simplify), counts.ctypes.data, V.ctypes.data, E.ctypes.data,
            F.ctypes.data, P.ctypes.data, C.ctypes.data)
    assert not returns_status or rc == 0
    v, he, f = (int(x) for x in counts)
    return counts, V[:v].copy(), E[:he].copy(), F[:f].copy(), P[:f].copy(), C


def qhull_defined(rec):
    """The records of a b3QHull::Set result that the reference defines: a vertex that no face
    uses keeps an unwritten slot (qhull.cpp:317-320), zeroed here, and the centroid, which
    averages it in (qhull.cpp:54-60), is dropped when one exists.  Returns (records, orphans)."""
    counts, V, E, F, P, C = rec
    used = np.zeros(len(V), bool)
    used[E[:, 0]] = True
    V = V.copy()
    V[~used] = 0.0
    orphans = int((~used).sum())
    return (counts, V, E, F, P, C if orphans == 0 else np.zeros(3, np.float32)), orphans


def qhull_digest(rec):
    import hashlib
    h = hashlib.sha256()
    for a in qhull_defined(rec)[0]:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def contact_init_workload(lib, oracle, seed, n, M=1):
    """Inputs of the step after the narrowphase for n contacts of M manifold slots: the new
    manifolds of a moving step and the old ones of the step before (from the oracle's
    narrowphase with caches carried, so most keys persist), random old impulses, bodies with
    random positions, orientations, velocities, mass data (some static) and materials, some
    contacts fast enough to trip the restitution threshold.
    Returns dict(bodies, state, material, counts, manifolds, old_counts, old_manifolds,
    old_impulses)."""
    from bounce_amd.types import (IMPULSES_DTYPE, MANIFOLD_DTYPE, MATERIAL_DTYPE, SOLVER_BODY_DTYPE,
                                  empty_caches)
    rng = np.random.default_rng(seed)
    npairs = n * M
    pool, shapes, pairs, xfs = mixed_workload(lib, seed, npairs, gap_mean=-0.02, gap_sigma=0.03)
    caches = empty_caches(npairs)
    old = oracle.collide(pool, shapes, pairs, xfs, caches=caches)
    x2 = xfs.copy()
    x2[:, 1, 9:12] += rng.normal(0.0, 0.002, (npairs, 3)).astype(np.float32)
    new = oracle.collide(pool, shapes, pairs, x2, caches=caches)
    nb = max(2, n)
    state = np.zeros(nb, SOLVER_BODY_DTYPE)
    state["x"] = rng.uniform(-5, 5, (nb, 3))
    q = rng.normal(size=(nb, 4))
    state["q"] = q / np.linalg.norm(q, axis=1, keepdims=True)
    state["v"] = rng.normal(0, 1.5, (nb, 3))
    state["w"] = rng.normal(0, 1.0, (nb, 3))
    static = rng.random(nb) < 0.1
    state["invMass"] = np.where(static, 0.0, rng.uniform(0.05, 2.0, nb))
    A = rng.normal(size=(nb, 3, 3))
    I = np.einsum("bij,bkj->bik", A, A) * 0.2 + np.eye(3) * 0.05
    I[static] = 0.0
    state["invI"] = I.transpose(0, 2, 1).reshape(nb, 9)
    state["localCenter"] = rng.normal(0, 0.05, (nb, 3))
    bodies = rng.integers(0, nb, (n, 2)).astype(np.uint32)
    material = np.zeros(n, MATERIAL_DTYPE)
    sa, sb = shapes[pairs[::M, 0]], shapes[pairs[::M, 1]]
    material["radiusA"], material["radiusB"] = sa["radius"], sb["radius"]
    material["frictionA"], material["frictionB"] = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    material["restitutionA"], material["restitutionB"] = rng.uniform(0, 0.6, n), rng.uniform(0, 0.6, n)
    manifolds = new.reshape(n, M)
    old_manifolds = old.reshape(n, M)
    counts = np.full(n, M, np.uint32) if M == 1 else rng.integers(1, M + 1, n).astype(np.uint32)
    old_counts = np.full(n, M, np.uint32) if M == 1 else rng.integers(0, M + 1, n).astype(np.uint32)
    old_impulses = np.zeros((n, M), IMPULSES_DTYPE)
    old_impulses["normalImpulse"] = rng.uniform(0, 3, (n, M, 4))
    old_impulses["tangentImpulse"] = rng.normal(0, 1, (n, M, 2))
    old_impulses["motorImpulse"] = rng.normal(0, 1, (n, M))
    return dict(bodies=bodies, state=state, material=material, counts=counts, manifolds=manifolds,
                old_counts=old_counts, old_manifolds=old_manifolds, old_impulses=old_impulses)


# ---------------------------------------------------------------- broadphase scripts
def broadphase_script(seed, n0=300, steps=20, extent=20.0, create_frac=0.03, destroy_frac=0.03,
                      touch_frac=0.01, move_frac=0.5, big=2):
    """A seeded sequence of b3BroadPhase calls on proxies named by creation order (handles):
    ("create", aabbs [k, 6]), ("destroy", handles), ("touch", handles),
    ("move", handles, aabbs [k, 6], displacements [k, 3]) and ("find",) once per step.
    AABBs are float32 boxes of half-extent 0.1-1 m in an `extent` cube, plus `big` ground-like
    slabs; a move translates a proxy's tight box by up to 0.3 m (many stay inside their fat
    box) and 5 % of moves jump up to 3 m."""
    rng = np.random.default_rng(seed)
    tight = []  # tight AABB per handle
    live = []
    ops = []

    def new_boxes(k):
        c = rng.uniform(0.0, extent, (k, 3))
        h = rng.uniform(0.1, 1.0, (k, 3))
        return np.concatenate([c - h, c + h], 1).astype(np.float32)

    first = new_boxes(n0)
    for g in range(min(big, n0)):
        first[g] = np.array([0.0, -0.5 + 3.0 * g, 0.0, extent, 0.5 + 3.0 * g, extent], np.float32)
    ops.append(("create", first))
    tight.extend(first)
    live.extend(range(n0))
    for _ in range(steps):
        ops.append(("find",))
        alive = np.array(live)
        k = max(1, int(move_frac * len(alive)))
        mv = rng.choice(alive, size=min(k, len(alive)), replace=False)
        d = rng.uniform(-0.3, 0.3, (len(mv), 3))
        jump = rng.random(len(mv)) < 0.05
        d[jump] = rng.uniform(-3.0, 3.0, (int(jump.sum()), 3))
        d = d.astype(np.float32)
        boxes = np.stack([tight[h] for h in mv]) + np.concatenate([d, d], 1)
        boxes = boxes.astype(np.float32)
        for j, h in enumerate(mv):
            tight[h] = boxes[j]
        ops.append(("move", mv.astype(np.int64), boxes, d))
        nd = int(destroy_frac * len(alive))
        if nd:
            dead = rng.choice(alive, size=nd, replace=False)
            ops.append(("destroy", dead.astype(np.int64)))
            live = [h for h in live if h not in set(dead.tolist())]
        nt = int(touch_frac * len(live))
        if nt:
            ops.append(("touch", rng.choice(np.array(live), size=nt, replace=False).astype(np.int64)))
        nc = int(create_frac * len(live)) + 1
        boxes = new_boxes(nc)
        ops.append(("create", boxes))
        base = len(tight)
        tight.extend(boxes)
        live.extend(range(base, base + nc))
    ops.append(("find",))
    return ops


def run_broadphase_script(bp, ops):
    """Runs a script on an object with create(aabbs, users) -> ids, destroy(ids), touch(ids),
    move(ids, aabbs, disps) -> moved flags and find_pairs() -> [k, 4]; returns the ids of
    every handle, the moved flags of every move and the pairs of every find (with users =
    handles)."""
    ids = []
    finds, moves = [], []
    for op in ops:
        if op[0] == "create":
            users = np.arange(len(ids), len(ids) + len(op[1]), dtype=np.uint32)
            ids.extend(int(x) for x in bp.create(op[1], users))
        elif op[0] == "destroy":
            bp.destroy(np.array([ids[h] for h in op[1]], np.uint32))
        elif op[0] == "touch":
            bp.touch(np.array([ids[h] for h in op[1]], np.uint32))
        elif op[0] == "move":
            moves.append(np.asarray(bp.move(np.array([ids[h] for h in op[1]], np.uint32), op[2], op[3])))
        else:
            finds.append(np.asarray(bp.find_pairs()))
    return np.array(ids, np.uint32), moves, finds


class ScalarBroadPhase:
    """Adapter of oracle_lib.BroadPhase (one proxy per call) to the batched script runner."""

    def __init__(self, bp):
        self.bp = bp

    def create(self, aabbs, users):
        return [self.bp.create(a, u) for a, u in zip(aabbs, users)]

    def destroy(self, ids):
        for i in ids:
            self.bp.destroy(i)

    def touch(self, ids):
        for i in ids:
            self.bp.touch(i)

    def move(self, ids, aabbs, disps):
        return [self.bp.move(i, a, d) for i, a, d in zip(ids, aabbs, disps)]

    def find_pairs(self):
        return self.bp.find_pairs(1 << 20)


# ---------------------------------------------------------------- scripted worlds
def kinematic_world(lib, seed, n_dyn=120, steps=40, spawn_every=10, spawn_count=12, extent=8.0, dt=1.0 / 60.0,
                    n_hulls=40, speed=2.5):
    """A world of scripted bodies for the contact-manager tests: a static 60 x 1 x 60 m ground
    box (top at y = 0) and dynamic bodies with 8-16-point hulls, moving with constant linear
    and angular velocities and bouncing off the ground (kinematically; no solver).
    Returns a dict: pool, shapes, spawns [(step, bodies, shape indices, types, xfs, sweeps)]
    (step -1: before the first step), states[s] = (xfs, sweeps) of every body for the move
    after step s (sweep from the step's state to the next)."""
    from bounce_amd.shapes import box_hull
    from bounce_amd.types import HullPool, SWEEP_DTYPE, SHAPE_DTYPE, HULL
    from bounce_amd.workloads import gen_hull_pool, quat_xf_rotation, _quat_mul
    rng = np.random.default_rng(seed)
    hulls = gen_hull_pool(lib, seed, n_hulls, 8, 16, 0.25, 0.6)
    pool = HullPool.join(HullPool.concat([box_hull(30.0, 0.5, 30.0)]), hulls)
    shapes = np.zeros(pool.count, SHAPE_DTYPE)
    shapes["type"] = HULL
    shapes["hull"] = np.arange(pool.count)
    n_total = 1 + n_dyn + spawn_count * (steps // spawn_every)
    shape_of = np.concatenate([[0], 1 + rng.integers(0, n_hulls, n_total - 1)]).astype(np.uint32)
    types = np.full(n_total, 2, np.uint32)
    types[0] = 0
    pos = np.zeros((n_total, 3), np.float32)
    pos[1:] = rng.uniform([-extent, 0.6, -extent], [extent, 2.0 * extent, extent], (n_total - 1, 3))
    vel = rng.uniform(-speed, speed, (n_total, 3)).astype(np.float32)
    vel[0] = 0.0
    q = rng.normal(size=(n_total, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[0] = [0.0, 0.0, 0.0, 1.0]
    q = q.astype(np.float32)
    w = rng.uniform(-2.0, 2.0, (n_total, 3)).astype(np.float32)
    w[0] = 0.0

    def xf_of(p, qq):
        x = np.zeros((len(p), 12), np.float32)
        x[:, :9] = quat_xf_rotation(qq)
        x[:, 9:] = p
        return x

    def sweep_of(p0, q0, p1, q1):
        s = np.zeros(len(p0), SWEEP_DTYPE)
        s["orientation0"], s["worldCenter0"], s["orientation"], s["worldCenter"] = q0, p0, q1, p1
        return s

    spawns = []
    first = 1 + n_dyn
    b = np.arange(first)
    spawns.append((-1, b, shape_of[b], types[b], xf_of(pos[b], q[b]), sweep_of(pos[b], q[b], pos[b], q[b])))
    live = first
    states = []
    for s in range(steps):
        if s > 0 and s % spawn_every == 0 and live + spawn_count <= n_total:
            b = np.arange(live, live + spawn_count)
            spawns.append((s, b, shape_of[b], types[b], xf_of(pos[b], q[b]), sweep_of(pos[b], q[b], pos[b], q[b])))
            live += spawn_count
        p1 = (pos + dt * vel).astype(np.float32)
        low = p1[:, 1] < 0.4
        low[0] = False
        vel[low, 1] = np.abs(vel[low, 1])
        p1[0] = pos[0]
        wq = np.concatenate([w, np.zeros((n_total, 1), np.float32)], 1)
        dq = 0.5 * dt * _quat_mul(wq.astype(np.float64), q.astype(np.float64))
        q1 = q.astype(np.float64) + dq
        q1 /= np.linalg.norm(q1, axis=1, keepdims=True)
        q1 = q1.astype(np.float32)
        q1[0] = q[0]
        b = np.arange(1, live)
        states.append((b, xf_of(p1[b], q1[b]), sweep_of(pos[b], q[b], p1[b], q1[b])))
        pos, q = p1, q1
    return {"pool": pool, "shapes": shapes, "spawns": spawns, "states": states, "steps": steps}


def run_kinematic_world(world, w, check=None):
    """Steps a world object (KWorld-like: add_bodies, find_new_contacts, update_contacts,
    move_bodies, contacts) through a kinematic_world script in b3World::Step's order: new
    shapes -> FindNewContacts, UpdateContacts, the move with SynchronizeShapes, then
    FindNewContacts (world.cpp:115-144, 308-336).  check(step, world) runs after each
    UpdateContacts."""
    spawns = {s[0]: s for s in w["spawns"]}
    _, b, sh, ty, xf, sw = spawns[-1]
    world.add_bodies(sh, ty, xf, sw)
    world.find_new_contacts()
    for s in range(w["steps"]):
        if s in spawns:
            _, b, sh, ty, xf, sw = spawns[s]
            world.add_bodies(sh, ty, xf, sw)
            world.find_new_contacts()
        stats = world.update_contacts()
        if check:
            check(s, world, stats)
        bodies, xfs, sweeps = w["states"][s]
        world.move_bodies(bodies, xfs, sweeps)
        world.find_new_contacts()


def rain_script(world, steps):
    """A gen_rain_world world as a run_kinematic_world script: bodies enter the script's world
    at their spawn steps (statics first), in spawn order (body indices of the script world are
    that order; `order` maps them back to rain bodies); after each step every body present
    moves to its next pose."""
    from bounce_amd.workloads import rain_state
    spawn = world["spawn"]
    order = []
    spawns = []
    st = np.nonzero(spawn < 0)[0]
    xf, sw, _ = rain_state(world, 0, st)
    spawns.append((-1, st, world["body_shapes"][st], np.zeros(len(st), np.uint32), xf, sw))
    order.extend(st.tolist())
    states = []
    for s in range(steps):
        b = np.nonzero(spawn == s)[0]
        if len(b):
            xf, sw, _ = rain_state(world, s, b)
            spawns.append((s, b, world["body_shapes"][b], np.full(len(b), 2, np.uint32), xf, sw))
            order.extend(b.tolist())
        present = np.array(order)
        moving = np.nonzero(spawn[present] >= 0)[0]
        xf, sw, _ = rain_state(world, s + 1, present[moving])
        states.append((moving, xf, sw))
    return {"pool": world["pool"], "shapes": world["shapes"], "spawns": spawns, "states": states, "steps": steps,
            "order": np.array(order)}


# ---- clip polygons beyond the reference's 32-vertex inline capacity

def prism_hull(n, r, h):
    """Right prism over a regular n-gon (radius r, half-height h): two n-gon caps."""
    from bounce_amd.shapes import polyhedron
    a = 2 * np.pi * np.arange(n) / n
    ring = np.stack([r * np.cos(a), np.zeros(n), r * np.sin(a)], 1)
    v = np.concatenate([ring + (0, -h, 0), ring + (0, h, 0)])
    faces = [list(range(n)), list(range(2 * n - 1, n - 1, -1))]  # outward: -y, +y
    faces += [[i, n + i, n + (i + 1) % n, (i + 1) % n] for i in range(n)]
    return polyhedron(v, faces)


def wide_prism_pairs(seed, n):
    """Prisms with 24-, 40- and 60-gon caps stacked cap to cap, the upper one turned about
    the vertical, tilted by a few milliradians and sunk up to 2 cm: the incident cap clipped by
    the reference cap's side planes gives polygons of up to twice the sides (48-120 vertices),
    beyond the reference's 32-vertex inline polygon (clip.cpp's b3StackArray grows on the
    heap).  Returns (pool, pairs, xfs)."""
    from bounce_amd.shapes import identity_xf
    from bounce_amd.types import HullPool
    rng = np.random.default_rng(seed)
    pool = HullPool.concat([prism_hull(k, r, 0.3) for k, r in ((24, 0.6), (40, 0.55), (60, 0.65))])
    pairs = rng.integers(0, pool.count, (n, 2)).astype(np.uint32)
    xfs = np.zeros((n, 2, 12), np.float32)
    for i in range(n):
        yaw = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(yaw), np.sin(yaw)
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        tilt = rng.normal(0, 0.002, 3)
        R = R @ np.array([[1, -tilt[2], tilt[1]], [tilt[2], 1, -tilt[0]], [-tilt[1], tilt[0], 1]])
        xfs[i, 0] = identity_xf((0, 0, 0))
        xfs[i, 1] = identity_xf((rng.normal(0, 0.05), 0.6 - rng.uniform(0.0, 0.02), rng.normal(0, 0.05)), R)
    return pool, pairs, xfs


def wide_prism_mesh(seed, n_shapes):
    """40- and 60-gon prisms lying on a flat grid of 4 m cells (each cap within one or two
    large triangles), turned about the vertical, sunk 1-10 mm: the cap and the triangle clip
    to polygons of 40-63 vertices.  Returns (pool, shapes, vertices, triangles, xfs)."""
    from bounce_amd.shapes import identity_xf
    from bounce_amd.types import HULL, SHAPE_DTYPE, HullPool
    rng = np.random.default_rng(seed)
    verts, tris = grid_mesh(8, 8, cell=4.0, amplitude=0.0, seed=seed)
    pool = HullPool.concat([prism_hull(40, 0.5, 0.1), prism_hull(60, 0.6, 0.1)])
    shapes = np.zeros(n_shapes, SHAPE_DTYPE)
    shapes["type"] = HULL
    shapes["hull"] = rng.integers(0, pool.count, n_shapes)
    xfs = np.zeros((n_shapes, 2, 12), np.float32)
    for i in range(n_shapes):
        yaw = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(yaw), np.sin(yaw)
        R = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
        x, z = rng.uniform(-14.0, 14.0, 2)
        xfs[i, 0] = identity_xf((x, 0.1 - rng.uniform(0.001, 0.01), z), R)
        xfs[i, 1] = identity_xf((0, 0, 0))
    return pool, shapes, verts, tris, xfs


def wide_plate_workload(seed, n_shapes):
    """Flat boxes of 1-1.8 m on a flat grid of 0.25 m cells: 70-200 triangles under each fat
    AABB (beyond the device lists' 64) and single clusters of 150-400 points."""
    return plate_workload(seed, n_shapes, H=32, W=32, cell=0.25,
                          extents=[(0.5, 0.5), (0.9, 0.6), (0.7, 0.7), (0.8, 0.9)])
