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
