"""Seeded synthetic workloads (wrappers over the generators in libbounce_amd.so)."""
import ctypes
import os

import numpy as np

from .types import HullPool, ptr


def gen_hull_pool(lib, seed, count, nmin=8, nmax=64, amin=0.25, amax=2.0):
    """`count` hulls, each b3QHull::Set (no simplification, b3amd_quickhull) of N ~ U{nmin..nmax}
    points on an ellipsoid with semi-axes ~ U[amin, amax]; twins adjacent, the reference's
    numbering."""
    desc = np.zeros((count, 6), np.uint32)
    cen = np.zeros((count, 3), np.float32)
    verts = np.zeros((count * nmax, 3), np.float32)
    edges = np.zeros((count * (6 * nmax - 12), 4), np.uint32)
    faces = np.zeros(count * (2 * nmax - 4), np.uint32)
    planes = np.zeros((count * (2 * nmax - 4), 4), np.float32)
    sizes = np.zeros(3, np.uint64)
    f = lib.b3amd_gen_hull_pool
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float,
                  ctypes.c_float] + [ctypes.c_void_p] * 7
    rc = f(seed, count, nmin, nmax, amin, amax, ptr(desc), ptr(cen), ptr(verts), ptr(edges), ptr(faces),
           ptr(planes), ptr(sizes))
    if rc != 0:
        raise RuntimeError("b3amd_gen_hull_pool failed: %d" % rc)
    nv, ne, nf = (int(x) for x in sizes)
    return HullPool(desc, cen, verts[:nv].copy(), edges[:ne].copy(), faces[:nf].copy(), planes[:nf].copy())


def quickhull(lib, points, simplify=False):
    """b3QHull::Set of `points` [n, 3] (b3amd_quickhull): the hull record
    (vertices, edges, faces, planes, centroid) as HullPool.concat takes it, or None when the
    reference would leave the hull empty."""
    pts = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    m = 4 * max(len(pts), 4) + 16
    counts = np.zeros(3, np.uint32)
    V, E = np.zeros((m, 3), np.float32), np.zeros((6 * m, 4), np.uint32)
    F, P, C = np.zeros(2 * m, np.uint32), np.zeros((2 * m, 4), np.float32), np.zeros(3, np.float32)
    f = lib.b3amd_quickhull
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int] + [ctypes.c_void_p] * 6
    if f(ptr(pts), len(pts), int(bool(simplify)), ptr(counts), ptr(V), ptr(E), ptr(F), ptr(P), ptr(C)) != 0:
        raise RuntimeError("b3amd_quickhull failed")
    nv, ne, nf = (int(x) for x in counts)
    if nv == 0:
        return None
    return V[:nv].copy(), E[:ne].copy(), F[:nf].copy(), P[:nf].copy(), C.copy()


def save_hull_pool(lib, pool, path):
    """Writes `pool` as a binary hull file (b3amd_hull_file_write)."""
    f = lib.b3amd_hull_file_write
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
    ps = pool.struct()
    if f(os.fsencode(path), ctypes.byref(ps)) != 0:
        raise RuntimeError("b3amd_hull_file_write(%s) failed" % path)


def load_hull_pool(lib, path):
    """Reads a binary hull file (b3amd_hull_file_info / b3amd_hull_file_read) into a HullPool."""
    count = ctypes.c_uint32()
    sizes = np.zeros(3, np.uint64)
    f = lib.b3amd_hull_file_info
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
    if f(os.fsencode(path), ctypes.byref(count), ptr(sizes)) != 0:
        raise RuntimeError("%s is not a hull file" % path)
    n, (nv, ne, nf) = count.value, (int(x) for x in sizes)
    desc, cen = np.zeros((n, 6), np.uint32), np.zeros((n, 3), np.float32)
    verts, edges = np.zeros((nv, 3), np.float32), np.zeros((ne, 4), np.uint32)
    faces, planes = np.zeros(nf, np.uint32), np.zeros((nf, 4), np.float32)
    g = lib.b3amd_hull_file_read
    g.restype = ctypes.c_int
    g.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 6
    if g(os.fsencode(path), ptr(desc), ptr(cen), ptr(verts), ptr(edges), ptr(faces), ptr(planes)) != 0:
        raise RuntimeError("%s: corrupt hull file" % path)
    return HullPool(desc, cen, verts, edges, faces, planes)


def gen_hull_pairs(lib, seed, pool, n, gap_mean=0.0, gap_sigma=0.05):
    """n hull-hull pairs: (pairs [n,2] u32 hull indices, xfs [n,2,12] f32)."""
    pairs = np.zeros((n, 2), np.uint32)
    xfs = np.zeros((n, 2, 12), np.float32)
    f = lib.b3amd_gen_hull_pairs
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_float,
                  ctypes.c_void_p, ctypes.c_void_p]
    ps = pool.struct()
    rc = f(seed, ctypes.byref(ps), n, gap_mean, gap_sigma, ptr(pairs), ptr(xfs))
    if rc != 0:
        raise RuntimeError("b3amd_gen_hull_pairs failed: %d" % rc)
    return pairs, xfs


def place_along(lib, pool, ha, ra, hb, rb, dirs, gaps):
    """b3amd_gen_place_along: per i, the distance along dirs[i] at which hull hb[i]
    (rotation columns rb[i]) has SAT signed separation gaps[i] from hull ha[i] (rotation
    ra[i]) at the origin."""
    n = len(ha)
    ha, hb = np.ascontiguousarray(ha, np.uint32), np.ascontiguousarray(hb, np.uint32)
    ra, rb = np.ascontiguousarray(ra, np.float32).reshape(n, 9), np.ascontiguousarray(rb, np.float32).reshape(n, 9)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(n, 3)
    gaps = np.ascontiguousarray(gaps, np.float64).reshape(n)
    t = np.zeros(n, np.float64)
    f = lib.b3amd_gen_place_along
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32] + [ctypes.c_void_p] * 7
    ps = pool.struct()
    rc = f(ctypes.byref(ps), n, ptr(ha), ptr(ra), ptr(hb), ptr(rb), ptr(dirs), ptr(gaps), ptr(t))
    if rc != 0:
        raise RuntimeError("b3amd_gen_place_along failed: %d" % rc)
    return t


def gen_convex_pairs(lib, seed, n, hull_pool=2048, gap_mean=-0.02, gap_sigma=0.08):
    """n sphere/capsule pairs for the convex narrowphase, in equal shares of the five kinds
    it takes: sphere-sphere, sphere-capsule, sphere-hull, capsule-capsule, capsule-hull (type
    A <= type B, collide.cpp:172-177).  Spheres r ~ U[0.1, 0.5], offset from the body origin
    by U[-0.2, 0.2] per axis in half of them; capsules r ~ U[0.1, 0.3] with segment length
    U[0, 1] (a tenth of them zero length); hulls of 8-32 points, semi-axes U[0.2, 1] m.
    Body A at a random point of a 10 m cube, B along a random direction u at the distance
    where their supports along u touch, plus N(gap_mean, gap_sigma); 5 % nearly concentric.
    Returns (pool, shapes [2n], pairs [n][2] (shape 2i, 2i+1), xfs [n][2][12])."""
    from .types import CAPSULE, HULL, SHAPE_DTYPE, SPHERE
    rng = np.random.default_rng(seed)
    pool = gen_hull_pool(lib, seed, hull_pool, 8, 32, 0.2, 1.0)
    kinds = np.array([(SPHERE, SPHERE), (SPHERE, CAPSULE), (SPHERE, HULL), (CAPSULE, CAPSULE), (CAPSULE, HULL)])
    k = kinds[rng.integers(0, len(kinds), n)]
    shapes = np.zeros(2 * n, SHAPE_DTYPE)
    t = k.reshape(-1)
    shapes["type"] = t
    sph, cap, hul = t == SPHERE, t == CAPSULE, t == HULL
    shapes["radius"][sph] = rng.uniform(0.1, 0.5, sph.sum())
    off = rng.uniform(-0.2, 0.2, (2 * n, 3)) * (rng.random(2 * n) < 0.5)[:, None]
    shapes["v"][sph, :3] = off[sph]
    L = rng.uniform(0.0, 1.0, 2 * n) * (rng.random(2 * n) < 0.9)
    shapes["radius"][cap] = rng.uniform(0.1, 0.3, cap.sum())
    shapes["v"][cap, 1] = -L[cap] / 2
    shapes["v"][cap, 4] = L[cap] / 2
    shapes["hull"][hul] = rng.integers(0, pool.count, hul.sum())
    R = _uniform_rotations(rng, 2 * n)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    # Support of each shape along +u (A) or -u (B) in world orientation.
    dirs = np.repeat(u, 2, axis=0) * np.tile([1.0, -1.0], n)[:, None]
    loc = np.einsum("nji,nj->ni", R, dirs)  # direction in the shape's frame: R^T d
    sup = np.zeros(2 * n)
    sup[sph] = (off[sph] * loc[sph]).sum(1) + shapes["radius"][sph]
    seg = np.abs(L[cap] / 2 * loc[cap, 1])
    sup[cap] = seg + shapes["radius"][cap]
    Vmax = int(pool.desc[:, 1].max())
    vp = np.zeros((pool.count, Vmax, 3))
    for h in range(pool.count):
        v = pool.hull(h)[0].astype(np.float64)
        vp[h, :len(v)] = v
        vp[h, len(v):] = v[0]
    hi = np.where(hul)[0]
    for a in range(0, len(hi), 1 << 16):
        b = hi[a:a + (1 << 16)]
        sup[b] = np.einsum("nvk,nk->nv", vp[shapes["hull"][b]], loc[b]).max(1)
    gap = sup[0::2] + sup[1::2] + rng.normal(gap_mean, gap_sigma, n)
    deep = rng.random(n) < 0.05
    gap[deep] = rng.uniform(0.0, 0.05, deep.sum())
    pa = rng.uniform(-5.0, 5.0, (n, 3))
    xfs = np.zeros((n, 2, 12), np.float32)
    for j, p in ((0, pa), (1, pa + gap[:, None] * u)):
        Rj = R[j::2]
        xfs[:, j, 0:3], xfs[:, j, 3:6], xfs[:, j, 6:9] = Rj[:, :, 0], Rj[:, :, 1], Rj[:, :, 2]
        xfs[:, j, 9:12] = p
    pairs = np.arange(2 * n, dtype=np.uint32).reshape(n, 2)
    return pool, shapes, pairs, xfs


def _rotations(rng, n, tilt_sigma=0.05):
    """n rotation matrices: uniform yaw about y with small tilts about x and z."""
    yaw = rng.uniform(0.0, 2.0 * np.pi, n)
    ax = rng.normal(0.0, tilt_sigma, n)
    az = rng.normal(0.0, tilt_sigma, n)
    cy, sy, cx, sx, cz, sz = np.cos(yaw), np.sin(yaw), np.cos(ax), np.sin(ax), np.cos(az), np.sin(az)
    Ry = np.zeros((n, 3, 3))
    Ry[:, 0, 0], Ry[:, 0, 2], Ry[:, 1, 1], Ry[:, 2, 0], Ry[:, 2, 2] = cy, sy, 1.0, -sy, cy
    Rx = np.zeros((n, 3, 3))
    Rx[:, 0, 0], Rx[:, 1, 1], Rx[:, 1, 2], Rx[:, 2, 1], Rx[:, 2, 2] = 1.0, cx, -sx, sx, cx
    Rz = np.zeros((n, 3, 3))
    Rz[:, 0, 0], Rz[:, 0, 1], Rz[:, 1, 0], Rz[:, 1, 1], Rz[:, 2, 2] = cz, -sz, sz, cz, 1.0
    return Rx @ Rz @ Ry


AABB_EXTENSION = 0.2  # b3_aabbExtension (settings.h:54): fat AABBs of the broadphase


def gen_serving_worlds(lib, seed, n_worlds, bodies_min=10, bodies_max=2000, steps=4, hull_shapes=2048,
                       box_shapes=64, sphere_shapes=64, pile_fraction=0.5):
    """Worlds of boxes, spheres and hulls (8-32 points) in stacks and piles on a static ground.

    Body counts are log-uniform in [bodies_min, bodies_max]; a third of the bodies are boxes
    (half-extents U[0.2, 0.5] m), a third spheres (radius U[0.2, 0.5] m) and a third hulls
    (8-32 points on ellipsoids with semi-axes U[0.2, 0.5] m), with uniform yaw and tilts of
    N(0, 0.01 rad).  Bodies are grouped in clusters on a 4.5 m grid: a stack (one column of
    1-6 bodies) or, with probability pile_fraction, a pile (2 x 2 or 3 x 3 columns of 1-3
    bodies, columns spaced by the pile's widest body + 0.25 m).  In a column each body rests
    on the one below (or the ground) with a vertical gap ~ N(-5 mm, 5 mm) between their
    supports.  The contact list is what the broadphase would hand b3ContactManager: every
    pair of bodies whose fat AABBs (extended by 0.2 m) overlap, the static ground included
    (body 0: a 100 x 1 x 100 m box, top at y = 0), ordered by (lower body, higher body) with
    typeA <= typeB.  Each step jitters body positions by N(0, 1 mm).

    Returns (pool, shapes, worlds); a world is a dict with body_shapes [B] u32,
    pairs [P, 2] u32 body indices and xfs [steps, B, 12] f32.
    """
    from .shapes import box_hull
    from .types import SHAPE_DTYPE, SPHERE, HULL
    rng = np.random.default_rng(seed)
    hulls = gen_hull_pool(lib, seed ^ 0x5EED, hull_shapes, 8, 32, 0.2, 0.5)
    boxes = [box_hull(*rng.uniform(0.2, 0.5, 3)) for _ in range(box_shapes)]
    pool = HullPool.concat([box_hull(50.0, 0.5, 50.0)] + boxes + [hulls.hull(h) for h in range(hulls.count)])
    n_shapes = 1 + box_shapes + sphere_shapes + hull_shapes
    shapes = np.zeros(n_shapes, SHAPE_DTYPE)
    shapes["type"] = HULL
    shapes["hull"][: 1 + box_shapes] = np.arange(1 + box_shapes)
    sph = slice(1 + box_shapes, 1 + box_shapes + sphere_shapes)
    shapes["type"][sph] = SPHERE
    shapes["hull"][sph] = 0
    shapes["radius"][sph] = rng.uniform(0.2, 0.5, sphere_shapes)
    shapes["hull"][1 + box_shapes + sphere_shapes:] = 1 + box_shapes + np.arange(hull_shapes)
    # Padded vertex arrays per shape for support extents (spheres use their radius).
    vmax = 32
    verts = np.zeros((n_shapes, vmax, 3))
    for s in range(n_shapes):
        if shapes["type"][s] == HULL:
            v = pool.hull(int(shapes["hull"][s]))[0]
            verts[s, : len(v)] = v
            verts[s, len(v):] = v[0]
    worlds = []
    for _ in range(n_worlds):
        nb = int(np.exp(rng.uniform(np.log(bodies_min), np.log(bodies_max))))
        kind = rng.integers(0, 3, nb)
        body_shapes = np.where(kind == 0, 1 + rng.integers(0, box_shapes, nb),
                               np.where(kind == 1, sph.start + rng.integers(0, sphere_shapes, nb),
                                        sph.stop + rng.integers(0, hull_shapes, nb))).astype(np.uint32)
        R = _rotations(rng, nb, 0.01)
        # Support extents of each rotated body about its origin, per axis.
        is_sphere = shapes["type"][body_shapes] == SPHERE
        rv = np.matmul(verts[body_shapes], R.transpose(0, 2, 1))
        rad = shapes["radius"][body_shapes].astype(np.float64)[:, None]
        lo = np.where(is_sphere[:, None], -rad, rv.min(axis=1))
        hi = np.where(is_sphere[:, None], rad, rv.max(axis=1))
        # Clusters: (columns per side k, bodies per column), bodies taken in order.
        ks = np.where(rng.uniform(size=nb) < pile_fraction, rng.integers(2, 4, nb), 1)
        ncols = ks * ks
        col_c = np.repeat(np.arange(nb), ncols)
        col_q = np.arange(len(col_c)) - np.repeat(np.cumsum(ncols) - ncols, ncols)
        col_h = np.where(ks[col_c] > 1, rng.integers(1, 4, len(col_c)), rng.integers(1, 7, len(col_c)))
        cum = np.cumsum(col_h)
        last = int(np.searchsorted(cum, nb))                     # the column holding body nb - 1
        col_h, col_q, col_c = col_h[: last + 1].copy(), col_q[: last + 1], col_c[: last + 1]
        col_h[last] -= cum[last] - nb
        ks = ks[: col_c[-1] + 1]
        col = np.repeat(np.arange(len(col_h)), col_h)           # body -> column
        group = col_c[col]                                       # body -> cluster
        side = int(np.ceil(np.sqrt(len(ks))))
        wbody = np.max(hi[:, [0, 2]] - lo[:, [0, 2]], axis=1)
        cstart = np.searchsorted(group, np.arange(len(ks)))
        sp = np.maximum.reduceat(wbody, cstart) + 0.05          # the cluster's widest body + 5 cm
        c, q = col_c[col], col_q[col]
        k = ks[c]
        pos = np.zeros((nb, 3))
        pos[:, 0] = 4.5 * (c % side - side / 2.0) + sp[c] * (q % k - (k - 1) / 2.0)
        pos[:, 2] = 4.5 * (c // side - side / 2.0) + sp[c] * (q // k - (k - 1) / 2.0)
        # Each body rests on the one below with a gap g: y_b = top_below + g_b - lo_b,
        # top_b = y_b + hi_b, so y is a running sum within the column.
        g = rng.normal(-0.005, 0.005, nb)
        step_h = g - lo[:, 1] + hi[:, 1]
        run = np.cumsum(step_h)
        first = np.searchsorted(col, np.arange(len(col_h)))
        base = run[first] - step_h[first]                        # running sum before each column
        pos[:, 1] = run - base[col] - hi[:, 1]
        # Fat AABBs and their overlaps within each cluster (clusters 4.5 m apart never meet).
        flo, fhi = pos + lo - AABB_EXTENSION, pos + hi + AABB_EXTENSION
        order = np.argsort(group, kind="stable")
        starts = np.searchsorted(group[order], np.arange(len(ks) + 1))
        G = int(np.max(np.diff(starts)))
        idx = np.full((len(ks), G), -1, np.int64)
        for g in range(G):
            sel = starts[:-1] + g < starts[1:]
            idx[sel, g] = order[starts[:-1][sel] + g]
        ii, jj = np.triu_indices(G, 1)
        A, B = idx[:, ii].ravel(), idx[:, jj].ravel()
        ok = (A >= 0) & (B >= 0)
        A, B = A[ok], B[ok]
        hit = np.all((flo[A] <= fhi[B]) & (flo[B] <= fhi[A]), axis=1)
        A, B = np.minimum(A[hit], B[hit]), np.maximum(A[hit], B[hit])
        ground = np.nonzero(flo[:, 1] <= AABB_EXTENSION)[0]  # the ground's fat box reaches y = 0.2
        pairs = np.concatenate([np.stack([np.zeros(len(ground), np.int64), ground + 1], 1),
                                np.stack([A + 1, B + 1], 1)])
        pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]
        bodies = nb + 1  # body 0 = ground
        all_shapes = np.concatenate([[0], body_shapes]).astype(np.uint32)
        t = shapes["type"][all_shapes]
        swap = t[pairs[:, 0]] > t[pairs[:, 1]]
        pairs[swap] = pairs[swap][:, ::-1]
        xfs = np.zeros((steps, bodies, 12), np.float32)
        xfs[:, 0, 0], xfs[:, 0, 4], xfs[:, 0, 8], xfs[:, 0, 10] = 1.0, 1.0, 1.0, -0.5
        for s in range(steps):
            p = pos + (rng.normal(0.0, 0.001, (nb, 3)) if s else 0.0)
            xfs[s, 1:, 0:3], xfs[s, 1:, 3:6], xfs[s, 1:, 6:9] = R[:, :, 0], R[:, :, 1], R[:, :, 2]
            xfs[s, 1:, 9:12] = p
        worlds.append({"body_shapes": all_shapes, "pairs": pairs.astype(np.uint32), "xfs": xfs})
    return pool, shapes, worlds


def world_pair_inputs(world, step):
    """The batch-API inputs of one world step: (pairs [P, 2] shape indices, xfs [P, 2, 12])."""
    bp = world["pairs"].astype(np.int64)
    pairs = world["body_shapes"][bp].astype(np.uint32)
    xfs = world["xfs"][step][bp]
    return np.ascontiguousarray(pairs), np.ascontiguousarray(xfs)


def gen_heightfield(seed, H=256, W=256, cell=0.5, amplitude=2.0, lattice=8):
    """(H+1) x (W+1) vertex heightfield triangulated as b3GridMesh (grid_mesh.h:57-77):
    heights = amplitude x value noise (uniform lattice values every `lattice` cells,
    smoothstep interpolation).  Returns (vertices [V,3] f32, triangles [2HW,3] u32)."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(-1.0, 1.0, size=(H // lattice + 2, W // lattice + 2))
    i, j = np.meshgrid(np.arange(H + 1), np.arange(W + 1), indexing="ij")
    u, v = i / lattice, j / lattice
    i0, j0 = u.astype(np.int64), v.astype(np.int64)
    fu, fv = u - i0, v - j0
    su, sv = fu * fu * (3 - 2 * fu), fv * fv * (3 - 2 * fv)
    h = (g[i0, j0] * (1 - su) * (1 - sv) + g[i0 + 1, j0] * su * (1 - sv) +
         g[i0, j0 + 1] * (1 - su) * sv + g[i0 + 1, j0 + 1] * su * sv)
    verts = np.stack([(j - 0.5 * W) * cell, amplitude * h, (i - 0.5 * H) * cell], -1).reshape(-1, 3)
    ii, jj = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    v1 = (ii * (W + 1) + jj).ravel()
    v2 = ((ii + 1) * (W + 1) + jj).ravel()
    v3 = ((ii + 1) * (W + 1) + jj + 1).ravel()
    v4 = (ii * (W + 1) + jj + 1).ravel()
    tris = np.empty((2 * H * W, 3), np.uint32)
    tris[0::2] = np.stack([v3, v2, v1], -1)
    tris[1::2] = np.stack([v1, v4, v3], -1)
    return verts.astype(np.float32), tris


def _uniform_rotations(rng, n):
    """n uniformly distributed rotation matrices (normalised 4-D Gaussian quaternions)."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    x, y, z, w = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def gen_mesh_world(lib, seed, H=256, W=256, n_shapes=100000, cell=0.5, amplitude=2.0, gap_sigma=0.05,
                   hull_pool=2048):
    """Mixed shapes resting on a heightfield: 40% spheres (r ~ U[0.1, 0.5]), 30% capsules
    (length ~ U[0.2, 1], r ~ U[0.1, 0.3]), 30% hulls of 8-32 points (semi-axes
    U[0.1, 0.5] m), uniform over the grid interior, uniform orientations, lowest point at
    the terrain height under the shape's centre + N(0, gap_sigma).  Returns (pool,
    shapes [n], vertices, triangles, xfs [n][2][12] = (shape body, identity mesh body))."""
    from .types import CAPSULE, HULL, SHAPE_DTYPE, SPHERE
    rng = np.random.default_rng(seed)
    verts, tris = gen_heightfield(seed, H, W, cell, amplitude)
    pool = gen_hull_pool(lib, seed, hull_pool, 8, 32, 0.1, 0.5)
    n = n_shapes
    kind = rng.random(n)
    R = _uniform_rotations(rng, n)
    shapes = np.zeros(n, SHAPE_DTYPE)
    low = np.zeros(n)
    sph = kind < 0.4
    cap = (kind >= 0.4) & (kind < 0.7)
    hul = kind >= 0.7
    shapes["type"][sph] = SPHERE
    shapes["radius"][sph] = rng.uniform(0.1, 0.5, sph.sum())
    low[sph] = shapes["radius"][sph]
    L = rng.uniform(0.2, 1.0, n)
    shapes["type"][cap] = CAPSULE
    shapes["radius"][cap] = rng.uniform(0.1, 0.3, cap.sum())
    shapes["v"][cap, 1] = -L[cap] / 2
    shapes["v"][cap, 4] = L[cap] / 2
    low[cap] = shapes["radius"][cap] + np.abs(R[cap, 1, 1]) * L[cap] / 2
    shapes["type"][hul] = HULL
    hid = rng.integers(0, pool.count, n)
    shapes["hull"][hul] = hid[hul]
    for h in np.unique(hid[hul]):  # lowest vertex of each rotated hull
        sel = np.where(hul & (hid == h))[0]
        v = pool.hull(int(h))[0].astype(np.float64)
        low[sel] = -np.einsum("kj,nj->nk", v, R[sel, 1, :]).min(1)
    half = 0.5 * cell * np.array([W, H])
    x = rng.uniform(-half[0] + 1.0, half[0] - 1.0, n)
    z = rng.uniform(-half[1] + 1.0, half[1] - 1.0, n)
    # Terrain height under (x, z): bilinear over the grid cell (between its two triangles).
    gx, gz = (x + half[0]) / cell, (z + half[1]) / cell
    j0, i0 = np.floor(gx).astype(np.int64), np.floor(gz).astype(np.int64)
    fx, fz = gx - j0, gz - i0
    hv = verts[:, 1].astype(np.float64).reshape(H + 1, W + 1)
    ground = (hv[i0, j0] * (1 - fx) * (1 - fz) + hv[i0, j0 + 1] * fx * (1 - fz) +
              hv[i0 + 1, j0] * (1 - fx) * fz + hv[i0 + 1, j0 + 1] * fx * fz)
    y = ground + low + rng.normal(0.0, gap_sigma, n)
    xfs = np.zeros((n, 2, 12), np.float32)
    xfs[:, 0, 0:3], xfs[:, 0, 3:6], xfs[:, 0, 6:9] = R[:, :, 0], R[:, :, 1], R[:, :, 2]
    xfs[:, 0, 9:12] = np.stack([x, y, z], -1)
    xfs[:, 1, 0], xfs[:, 1, 4], xfs[:, 1, 8] = 1.0, 1.0, 1.0
    return pool, shapes, verts, tris, xfs


def quat_xf_rotation(q):
    """b3QuatMat33 (quat.h:306-318) in float32 and its operation order: the rotation columns
    [n][9] of unit quaternions [n][4] (x, y, z, w), as b3Body::SynchronizeTransform sets
    m_xf.rotation from the sweep orientation."""
    q = np.asarray(q, np.float32)
    x, y, z, w = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    one = np.float32(1.0)
    return np.stack([one - (yy + zz), xy + wz, xz - wy,
                     xy - wz, one - (xx + zz), yz + wx,
                     xz + wy, yz - wx, one - (xx + yy)], -1)


def _quat_mul(a, b):
    ax, ay, az, aw = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    bx, by, bz, bw = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([aw * bx + ax * bw + ay * bz - az * by,
                     aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw,
                     aw * bw - ax * bx - ay * by - az * bz], -1)


def gen_warm_world(lib, seed, n_hulls=200000, columns=62, half_width=25.0, amin=0.25, amax=0.5,
                   gap_mean=-0.005, gap_sigma=0.005, pitch_gap=0.05, extension=0.2, settling=0.1, hz=60.0,
                   period=120):
    """One large world for warm-started steps: n_hulls distinct hulls of 8-64 points (on
    ellipsoids with semi-axes U[amin, amax] m, uniform orientations) packed in layers of
    columns x columns into a container of 2 half_width x 2 half_width m (static boxes:
    floor, body 0, top at y = 0; walls at x = -/+half_width, bodies 1-2, and z =
    -/+half_width, bodies 3-4).  Lattice cell (i, j, k) holds body 5 + k columns^2 +
    j columns + i.  The bodies of a column (i, j) form a stack: each rests on the one
    below at SAT signed separation g ~ N(gap_mean, gap_sigma), placed along +y by
    b3amd_gen_place_along (on the floor: its lowest vertex at height g).  Columns stand on a
    square grid whose pitch is the mean lateral support width plus pitch_gap, from the -x
    and -z walls.

    The contact list is every pair whose AABBs, extended by `extension` on each side (the
    broadphase's fat AABBs, settings.h:54), overlap, dynamic-dynamic and dynamic-static,
    in (lower body, higher body) order sorted by that pair, the order b3BroadPhase's sorted
    pair buffer reports them in (broad_phase.h:160-208).

    Motion (warm_world_state): each body oscillates from its place, a fraction `settling`
    by up to ~2 cm and 0.05 rad, the rest by ~2 mm and 0.005 rad, at 1-4 times
    hz / period Hz (0.5-2 Hz at 60 Hz, period 120), so the motion repeats every `period`
    steps and warm_world_state needs only period + 1 distinct states.  The pile is
    synthetic: no host solver runs here.

    Returns a dict: pool, shapes (one per body), body_shapes [B] u32, contacts [P][2] u32,
    pos0 [B][3] f64, q0 [B][4] f64, stack_gaps [n_hulls] (each hull's drawn separation from
    what it rests on) and the motion parameters."""
    from .shapes import box_hull
    from .types import SHAPE_DTYPE, HULL
    rng = np.random.default_rng(seed)
    hulls = gen_hull_pool(lib, seed ^ 0xC5C5, n_hulls, 8, 64, amin, amax)
    layers = -(-n_hulls // (columns * columns))
    wall_half = 0.5 * (layers * 2.0 * amax + 2.0)
    H = half_width
    statics = [((H + 1.0, 0.5, H + 1.0), (0.0, -0.5, 0.0)),
               ((0.5, wall_half, H + 1.0), (-H - 0.5, wall_half, 0.0)),
               ((0.5, wall_half, H + 1.0), (H + 0.5, wall_half, 0.0)),
               ((H + 1.0, wall_half, 0.5), (0.0, wall_half, -H - 0.5)),
               ((H + 1.0, wall_half, 0.5), (0.0, wall_half, H + 0.5))]
    ns = len(statics)
    pool = HullPool.join(HullPool.concat([box_hull(*e) for e, _ in statics]), hulls)
    B = ns + n_hulls
    shapes = np.zeros(B, SHAPE_DTYPE)
    shapes["type"] = HULL
    shapes["hull"] = np.arange(B, dtype=np.uint32)
    body_shapes = np.arange(B, dtype=np.uint32)

    # Uniform orientations as quaternions (x, y, z, w); statics at identity.
    q0 = np.zeros((B, 4))
    q0[:, 3] = 1.0
    qd = rng.normal(size=(n_hulls, 4))
    q0[ns:] = qd / np.linalg.norm(qd, axis=1, keepdims=True)
    # Support extents of each rotated hull along +-x, +-y, +-z.
    R = quat_xf_rotation(q0[ns:].astype(np.float32)).astype(np.float64).reshape(-1, 3, 3)  # [n][col][row]
    d = hulls.desc.astype(np.int64)
    hi = np.zeros((n_hulls, 3))
    lo = np.zeros((n_hulls, 3))
    for c0 in range(0, n_hulls, 20000):
        c1 = min(n_hulls, c0 + 20000)
        idx = d[c0:c1, 0:1] + np.minimum(np.arange(64)[None, :], d[c0:c1, 1:2] - 1)
        v = hulls.vertices[idx].astype(np.float64)                     # [m][64][3]
        w = np.einsum("mvj,mja->mva", v, R[c0:c1])                    # world-axis components
        hi[c0:c1], lo[c0:c1] = w.max(1), -w.min(1)
    # Lattice: chains of support-to-support placements along x (i), z (j) and y (k).
    cells = layers * columns * columns
    ext_hi = np.zeros((cells, 3))
    ext_lo = np.zeros((cells, 3))
    ext_hi[:n_hulls], ext_lo[:n_hulls] = hi, lo
    ext_hi = ext_hi.reshape(layers, columns, columns, 3)
    ext_lo = ext_lo.reshape(layers, columns, columns, 3)
    gap = rng.normal(gap_mean, gap_sigma, (layers, columns, columns, 3))

    # Stacks: the body of layer k rests on the one of layer k - 1 in its column (the floor
    # for the first), support to support along y plus its gap; columns stand on a square
    # grid of the mean lateral width plus `pitch_gap`, against the -x and -z walls.
    pitch = float((hi[:, [0, 2]] + lo[:, [0, 2]]).mean()) + pitch_gap
    # Rise of each body over the one below it: the distance along +y at which their SAT
    # separation is the body's gap (exact, not a support-extent estimate).
    rise = np.zeros(cells)
    rise[:columns * columns] = ext_lo[0, ..., 1].ravel() + gap[0, ..., 1].ravel()
    up = np.arange(columns * columns, n_hulls)
    rot = quat_xf_rotation(q0[ns:].astype(np.float32))
    rise[up] = place_along(lib, hulls, up - columns * columns, rot[up - columns * columns], up, rot[up],
                           np.tile([0.0, 1.0, 0.0], (len(up), 1)), gap.reshape(-1, 3)[up, 1])
    y = np.cumsum(rise.reshape(layers, columns, columns), axis=0)
    gi = -H + (np.arange(columns) + 0.5) * pitch
    x = np.broadcast_to(gi[None, None, :], y.shape)
    z = np.broadcast_to(gi[None, :, None], y.shape)
    pos0 = np.zeros((B, 3))
    for a, (e, c) in enumerate(statics):
        pos0[a] = c
    pos0[ns:] = np.stack([x, y, z], -1).reshape(-1, 3)[:n_hulls]

    # Contacts: fat AABB overlaps.
    from scipy.spatial import cKDTree
    lo_b = np.zeros((B, 3))
    hi_b = np.zeros((B, 3))
    for a, (e, c) in enumerate(statics):
        lo_b[a], hi_b[a] = np.subtract(c, e), np.add(c, e)
    lo_b[ns:], hi_b[ns:] = pos0[ns:] - lo, pos0[ns:] + hi
    lo_b -= extension
    hi_b += extension
    tree = cKDTree(pos0[ns:])
    reach = 2.0 * (float(max(hi.max(), lo.max())) + extension)
    pr = tree.query_pairs(reach, p=np.inf, output_type="ndarray").astype(np.int64) + ns
    keep = np.all((lo_b[pr[:, 0]] <= hi_b[pr[:, 1]]) & (lo_b[pr[:, 1]] <= hi_b[pr[:, 0]]), axis=1)
    dyn = pr[keep]
    st = []
    for a in range(ns):
        m = np.all((lo_b[ns:] <= hi_b[a]) & (lo_b[a] <= hi_b[ns:]), axis=1)
        idx = np.nonzero(m)[0] + ns
        st.append(np.stack([np.full(len(idx), a), idx], -1))
    contacts = np.concatenate([dyn] + st)
    contacts = np.sort(contacts, axis=1)
    contacts = contacts[np.lexsort((contacts[:, 1], contacts[:, 0]))]

    # Motion parameters.
    n = n_hulls
    fast = rng.random(n) < settling
    amp = np.where(fast, 0.02, 0.002)[:, None] * rng.normal(size=(n, 3)) / np.sqrt(3.0)
    rot_amp = np.where(fast, 0.05, 0.005) * rng.uniform(0.5, 1.0, n)
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    return {"pool": pool, "shapes": shapes, "body_shapes": body_shapes, "contacts": contacts.astype(np.uint32),
            "pos0": pos0, "q0": q0, "statics": ns, "hz": hz, "amp": amp, "period": period,
            "freq": rng.integers(1, 5, n) * (hz / period), "phase": rng.uniform(0.0, 2.0 * np.pi, n),
            "rot_amp": rot_amp, "rot_axis": axis, "rot_phase": rng.uniform(0.0, 2.0 * np.pi, n),
            "layers": layers, "columns": columns, "stack_gaps": gap.reshape(-1, 3)[:n_hulls, 1].copy()}


def _warm_world_quat(world, s):
    ns = world["statics"]
    q = world["q0"].copy()
    t = 2.0 * np.pi * world["freq"] * ((s % world["period"]) / world["hz"])
    theta = world["rot_amp"] * (np.sin(t + world["rot_phase"]) - np.sin(world["rot_phase"]))
    half = 0.5 * theta[:, None]
    dq = np.concatenate([world["rot_axis"] * np.sin(half), np.cos(half)], 1)
    qs = _quat_mul(world["q0"][ns:], dq)
    q[ns:] = qs / np.linalg.norm(qs, axis=1, keepdims=True)
    return q.astype(np.float32)


def warm_world_state_index(world, s):
    """Index of step s's body state among the period + 1 distinct ones: 0 for the first
    step, then 1..period repeating (step s > 0 uses state (s - 1) % period + 1)."""
    return 0 if s == 0 else (s - 1) % world["period"] + 1


def warm_world_state(world, s):
    """Body state at step s: (xfs [B][12] f32: rotation columns from the orientation by
    b3QuatMat33, then the position; sweeps [B][2][4] f32: (orientation0, orientation) =
    the orientations at steps s - 1 and s, both the step-0 one at s = 0)."""
    ns = world["statics"]
    q = _warm_world_quat(world, s)
    q_prev = _warm_world_quat(world, s - 1) if s > 0 else q
    t = 2.0 * np.pi * world["freq"] * ((s % world["period"]) / world["hz"])
    pos = world["pos0"].copy()
    pos[ns:] += world["amp"] * (np.sin(t + world["phase"]) - np.sin(world["phase"]))[:, None]
    xfs = np.empty((len(q), 12), np.float32)
    xfs[:, 0:9] = quat_xf_rotation(q)
    xfs[:, 9:12] = pos
    sweeps = np.stack([q_prev, q], 1)
    return xfs, sweeps


def warm_world_pair_inputs(world, xfs, sweeps, contacts=None):
    """Batch-API inputs of one step for the listed contacts (all by default): (pairs [n][2]
    shape indices, xfs [n][2][12], sweeps [n][4][4])."""
    c = (world["contacts"] if contacts is None else contacts).astype(np.int64)
    pairs = world["body_shapes"][c].astype(np.uint32)
    return (np.ascontiguousarray(pairs), np.ascontiguousarray(xfs[c]),
            np.ascontiguousarray(sweeps[c].reshape(-1, 4, 4)))


def gen_rain_world(lib, seed, n_hulls=200000, columns=62, half_width=25.0, interval=8, fall_height=10.0,
                   spin=3.0, hz=60.0, gravity=9.8, base=None, **layout):
    """A world whose hulls rain into a container: the resting layout of gen_warm_world
    (stacks of 8-64-point hulls on a grid in a 50 x 50 m box of floor and four walls), but
    each hull enters the world (b3Body::CreateShape) at its spawn step `fall_height` m above
    its resting place and falls (height fall_height - g t^2 / 2 over its place, orientation
    turning towards its resting one from up to spin * t_fall rad away) until it lands and
    rests.  Layer k of column (i, j) spawns at step k * interval + a per-column offset in
    [0, interval), so each hull lands on the one below after that one has come to rest.
    The motion is kinematic (no solver runs), and contacts are created as hulls pass and
    reach their neighbours and destroyed as they move apart.

    Returns gen_warm_world's dict (or a copy of `base`, one made by gen_warm_world, whose
    layout it reuses) plus: spawn [B] step of each body's shape (-1: statics, present from the
    start), t_fall, spin axis and rate per body."""
    if base is None:
        w = gen_warm_world(lib, seed, n_hulls=n_hulls, columns=columns, half_width=half_width, **layout)
    else:
        w = dict(base)
        columns = w["columns"]
    rng = np.random.default_rng(seed ^ 0x5A1)
    ns, B = w["statics"], len(w["body_shapes"])
    n = B - ns
    level = np.arange(n) // (columns * columns)
    col = np.arange(n) % (columns * columns)
    offset = rng.integers(0, interval, columns * columns)
    spawn = np.full(B, -1, np.int64)
    spawn[ns:] = level * interval + offset[col]
    axis = rng.normal(size=(n, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    w.update({"spawn": spawn, "interval": interval, "fall_height": fall_height, "gravity": gravity, "dt": 1.0 / hz,
              "t_fall": float(np.sqrt(2.0 * fall_height / gravity)), "spin_axis": axis,
              "spin": rng.uniform(-spin, spin, n)})
    return w


def rain_pose(world, s, bodies=None):
    """Positions [m][3] and orientations [m][4] (float32) of the listed bodies (all by default)
    at step s; a body before its spawn step holds its spawn pose."""
    ns = world["statics"]
    b = np.arange(len(world["body_shapes"])) if bodies is None else np.asarray(bodies, np.int64)
    pos = world["pos0"][b].copy()
    q = world["q0"][b].copy()
    dyn = b >= ns
    d = b[dyn] - ns
    t = np.maximum(s - world["spawn"][b[dyn]], 0) * world["dt"]
    h = np.maximum(world["fall_height"] - 0.5 * world["gravity"] * t * t, 0.0)
    pos[dyn, 1] += h
    psi = world["spin"][d] * np.maximum(world["t_fall"] - t, 0.0)
    turning = psi != 0.0
    r = np.concatenate([world["spin_axis"][d] * (0.5 * psi)[:, None], np.ones((len(d), 1))], 1)
    qs = _quat_mul(world["q0"][b[dyn]], r)
    qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    qd = q[dyn]
    qd[turning] = qs[turning]  # exactly the resting orientation once landed
    q[dyn] = qd
    return pos.astype(np.float32), q.astype(np.float32)


def rain_state(world, s, bodies=None):
    """Body state at step s for the listed bodies: (xfs [m][12]: b3QuatMat33 of the
    orientation, then the position; sweeps [m] SWEEP_DTYPE from the pose at s - 1 to the one
    at s, both the spawn pose at a body's spawn step (b3Body's constructor); orientation
    pairs [m][2][4] (orientation0, orientation))."""
    from .types import SWEEP_DTYPE
    p1, q1 = rain_pose(world, s, bodies)
    p0, q0 = rain_pose(world, s - 1, bodies)
    b = np.arange(len(world["body_shapes"])) if bodies is None else np.asarray(bodies, np.int64)
    new = world["spawn"][b] >= s
    p0[new], q0[new] = p1[new], q1[new]
    xfs = np.empty((len(p1), 12), np.float32)
    xfs[:, :9] = quat_xf_rotation(q1)
    xfs[:, 9:] = p1
    sw = np.zeros(len(p1), SWEEP_DTYPE)
    sw["orientation0"], sw["worldCenter0"], sw["orientation"], sw["worldCenter"] = q0, p0, q1, p1
    return xfs, sw, np.ascontiguousarray(np.stack([q0, q1], 1))
