"""Seeded synthetic workloads (wrappers over the generators in libbounce_amd.so)."""
import ctypes

import numpy as np

from .types import HullPool, ptr


def gen_hull_pool(lib, seed, count, nmin=8, nmax=64, amin=0.25, amax=2.0):
    """`count` quickhull-style hulls of N ~ U{nmin..nmax} points on ellipsoids with
    semi-axes ~ U[amin, amax] (triangulated, twins adjacent)."""
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


def gen_serving_worlds(lib, seed, n_worlds, bodies_min=10, bodies_max=2000, steps=4, hull_shapes=2048,
                       box_shapes=64, sphere_shapes=64):
    """Worlds of boxes, spheres and hulls (8-32 points) stacked in piles on a static ground.

    Body counts are log-uniform in [bodies_min, bodies_max]; a third of the bodies are boxes
    (half-extents U[0.2, 0.5] m), a third spheres (radius U[0.2, 0.5] m) and a third hulls
    (8-32 points on ellipsoids with semi-axes U[0.2, 0.5] m).  Bodies stand in stacks of
    1-6 on a 1.4 m grid, each resting on the one below (or the ground) with a vertical gap
    ~ N(-5 mm, 5 mm) between their supports, uniform yaw and tilts of N(0, 0.01 rad).  Each body has one
    contact, with its support; body 0 is the ground (a 100 x 1 x 100 m box, top at y = 0).
    Each step jitters body positions by N(0, 1 mm).

    Returns (pool, shapes, worlds); a world is a dict with body_shapes [B] u32,
    pairs [P, 2] u32 body indices (typeA <= typeB) and xfs [steps, B, 12] f32.
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
    # Padded vertex arrays per shape for support heights (spheres use their radius).
    vmax = 32
    verts = np.zeros((n_shapes, vmax, 3))
    nverts = np.zeros(n_shapes, np.int64)
    for s in range(n_shapes):
        if shapes["type"][s] == HULL:
            v = pool.hull(int(shapes["hull"][s]))[0]
            verts[s, : len(v)] = v
            verts[s, len(v):] = v[0]
            nverts[s] = len(v)
    worlds = []
    for _ in range(n_worlds):
        nb = int(np.exp(rng.uniform(np.log(bodies_min), np.log(bodies_max))))
        kind = rng.integers(0, 3, nb)
        body_shapes = np.where(kind == 0, 1 + rng.integers(0, box_shapes, nb),
                               np.where(kind == 1, sph.start + rng.integers(0, sphere_shapes, nb),
                                        sph.stop + rng.integers(0, hull_shapes, nb))).astype(np.uint32)
        R = _rotations(rng, nb, 0.01)
        # Support extents along y of each rotated body about its origin.
        is_sphere = shapes["type"][body_shapes] == SPHERE
        vy = np.einsum("bj,bvj->bv", R[:, 1, :], verts[body_shapes])
        lo = np.where(is_sphere, -shapes["radius"][body_shapes], vy.min(axis=1))
        hi = np.where(is_sphere, shapes["radius"][body_shapes], vy.max(axis=1))
        # Stacks on a grid.
        heights = []
        left = nb
        while left > 0:
            h = min(left, int(rng.integers(1, 7)))
            heights.append(h)
            left -= h
        side = int(np.ceil(np.sqrt(len(heights))))
        pos = np.zeros((nb, 3))
        pairs = []
        b = 0
        for c, h in enumerate(heights):
            x, z = 1.4 * (c % side - side / 2.0), 1.4 * (c // side - side / 2.0)
            top, below = 0.0, -1  # ground top, ground body
            for _k in range(h):
                g = rng.normal(-0.005, 0.005)
                pos[b] = (x, top + g - lo[b], z)
                top = pos[b, 1] + hi[b]
                pairs.append((below, b))
                below = b
                b += 1
        bodies = nb + 1  # body 0 = ground
        pairs = np.asarray(pairs, np.int64) + 1
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
