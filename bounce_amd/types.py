"""NumPy mirrors of the C records in include/bounce_amd.h.

Each dtype is byte-compatible with the reference type named beside it, so arrays of
these records can be handed to the C ABI (and to the reference) without conversion.
"""
import ctypes

import numpy as np

# b3Transform (transform.h:56-57): rotation columns x, y, z then position.
XF_FLOATS = 12

# b3amd_shape: b3SphereShape / b3CapsuleShape / b3HullShape fields read by the narrowphase.
SHAPE_DTYPE = np.dtype([
    ("type", "<u4"), ("hull", "<u4"), ("radius", "<f4"), ("v", "<f4", (6,)), ("pad", "<u4", (3,)),
])
assert SHAPE_DTYPE.itemsize == 48

# b3ManifoldPoint fields written by the narrowphase (manifold.h:66-76).
POINT_DTYPE = np.dtype([
    ("localNormal1", "<f4", (3,)), ("localPoint1", "<f4", (3,)), ("localPoint2", "<f4", (3,)),
    ("triangleKey", "<u4"), ("key", "<u4", (4,)),
])
assert POINT_DTYPE.itemsize == 56

# b3Manifold (manifold.h:79-93) without solver impulses.
MANIFOLD_DTYPE = np.dtype([
    ("pointCount", "<u4"), ("flags", "<u4"), ("reserved", "<u4", (2,)), ("points", POINT_DTYPE, (4,)),
])
assert MANIFOLD_DTYPE.itemsize == 240

# b3ConvexCache = b3SimplexCache (gjk.h:85-92) + b3SATFeaturePair (sat.h:74-80).
CACHE_DTYPE = np.dtype({
    "names": ["metric", "iterations", "count", "sindex1", "sindex2",
              "state", "ftype", "index1", "index2"],
    "formats": ["<f4", "<u4", "<u2", ("u1", (4,)), ("u1", (4,)), "<u4", "<u4", "<u4", "<u4"],
    "offsets": [0, 4, 8, 10, 14, 20, 24, 28, 32],
    "itemsize": 36,
})

# The step after the narrowphase (b3ContactSolver inputs and outputs, contact_solver.h).
SOLVER_BODY_DTYPE = np.dtype([("x", "<f4", (3,)), ("q", "<f4", (4,)), ("v", "<f4", (3,)), ("w", "<f4", (3,)),
                              ("invMass", "<f4"), ("invI", "<f4", (9,)), ("localCenter", "<f4", (3,))])
assert SOLVER_BODY_DTYPE.itemsize == 104
MATERIAL_DTYPE = np.dtype([("radiusA", "<f4"), ("radiusB", "<f4"), ("frictionA", "<f4"), ("frictionB", "<f4"),
                           ("restitutionA", "<f4"), ("restitutionB", "<f4"), ("reserved", "<u4", (2,))])
IMPULSES_DTYPE = np.dtype([("normalImpulse", "<f4", (4,)), ("persisting", "<u4", (4,)),
                           ("tangentImpulse", "<f4", (2,)), ("motorImpulse", "<f4"), ("reserved", "<u4")])
assert IMPULSES_DTYPE.itemsize == 48
VELOCITY_POINT_DTYPE = np.dtype([("rA", "<f4", (3,)), ("rB", "<f4", (3,)), ("normal", "<f4", (3,)),
                                 ("normalMass", "<f4"), ("normalImpulse", "<f4"), ("velocityBias", "<f4")])
VELOCITY_MANIFOLD_DTYPE = np.dtype([
    ("points", VELOCITY_POINT_DTYPE, (4,)), ("rA", "<f4", (3,)), ("rB", "<f4", (3,)), ("normal", "<f4", (3,)),
    ("tangent1", "<f4", (3,)), ("tangent2", "<f4", (3,)), ("tangentMass", "<f4", (4,)), ("tangentImpulse", "<f4", (2,)),
    ("motorImpulse", "<f4"), ("motorMass", "<f4"), ("pointCount", "<u4"), ("friction", "<f4"), ("restitution", "<f4"),
    ("reserved", "<u4", (2,))])
assert VELOCITY_MANIFOLD_DTYPE.itemsize == 304

# The step before the narrowphase: b3AABB3 (aabb3.h:26-28) and b3Sweep (transform.h:202-222).
AABB_DTYPE = np.dtype([("lower", "<f4", (3,)), ("upper", "<f4", (3,))])
assert AABB_DTYPE.itemsize == 24
SWEEP_DTYPE = np.dtype([("localCenter", "<f4", (3,)), ("orientation0", "<f4", (4,)), ("worldCenter0", "<f4", (3,)),
                        ("t0", "<f4"), ("orientation", "<f4", (4,)), ("worldCenter", "<f4", (3,))])
assert SWEEP_DTYPE.itemsize == 72

# b3BodyDef / b3ShapeDef / b3Body state of the host step (b3amd_body_def, b3amd_shape_def,
# b3amd_body_state in include/bounce_amd.h).
BODY_DEF_DTYPE = np.dtype([("type", "<u4"), ("awake", "<u4"), ("fixed_rotation", "<u4"), ("position", "<f4", (3,)),
                           ("orientation", "<f4", (4,)), ("linear_velocity", "<f4", (3,)),
                           ("angular_velocity", "<f4", (3,)), ("gravity_scale", "<f4"), ("linear_damping", "<f4"),
                           ("angular_damping", "<f4")])
SHAPE_DEF_DTYPE = np.dtype([("table_shape", "<u4"), ("density", "<f4"), ("friction", "<f4"),
                            ("restitution", "<f4"), ("sensor", "<u4")])
BODY_STATE_DTYPE = np.dtype([("xf", "<f4", (12,)), ("sweep", SWEEP_DTYPE), ("linear_velocity", "<f4", (3,)),
                             ("angular_velocity", "<f4", (3,)), ("mass", "<f4"), ("inv_mass", "<f4"),
                             ("inertia", "<f4", (9,)), ("inv_inertia", "<f4", (9,)), ("world_inv_inertia", "<f4", (9,)),
                             ("awake", "<u4"), ("sleep_time", "<f4")])

SPHERE, CAPSULE, HULL, MESH = 0, 1, 2, 3
# Manifold status bits (B3AMD_FLAG_* of bounce_amd.h).
FLAG_POLY_OVERFLOW, FLAG_INVALID_PAIR, FLAG_TRIANGLE_OVERFLOW = 1, 2, 4

CACHE_SEPARATION, CACHE_OVERLAP, CACHE_EMPTY = 0, 1, 2
FEATURE_EDGE1, FEATURE_FACE1, FEATURE_FACE2 = 0, 1, 2
NULL_KEY = 0xFFFFFFFF


def empty_caches(n):
    """n fresh b3ConvexCache records (count 0, feature state empty)."""
    c = np.zeros(n, CACHE_DTYPE)
    c["state"] = CACHE_EMPTY
    return c


class HullPoolStruct(ctypes.Structure):
    """b3amd_hull_pool / b3o_hull_pool."""
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("desc", ctypes.c_void_p),
        ("centroids", ctypes.c_void_p),
        ("vertices", ctypes.c_void_p),
        ("edges", ctypes.c_void_p),
        ("faces", ctypes.c_void_p),
        ("planes", ctypes.c_void_p),
    ]


class HullPool:
    """A flat pool of b3Hull records (hull.h:37-64) held in NumPy arrays."""

    def __init__(self, desc, centroids, vertices, edges, faces, planes):
        self.desc = np.ascontiguousarray(desc, dtype=np.uint32).reshape(-1, 6)
        self.centroids = np.ascontiguousarray(centroids, dtype=np.float32).reshape(-1, 3)
        self.vertices = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        self.edges = np.ascontiguousarray(edges, dtype=np.uint32).reshape(-1, 4)
        self.faces = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1)
        self.planes = np.ascontiguousarray(planes, dtype=np.float32).reshape(-1, 4)

    @property
    def count(self):
        return self.desc.shape[0]

    def struct(self):
        s = HullPoolStruct()
        s.count = self.count
        s.desc = self.desc.ctypes.data
        s.centroids = self.centroids.ctypes.data
        s.vertices = self.vertices.ctypes.data
        s.edges = self.edges.ctypes.data
        s.faces = self.faces.ctypes.data
        s.planes = self.planes.ctypes.data
        return s

    def hull(self, h):
        """(vertices, edges, faces, planes, centroid) of hull h as views."""
        vo, vn, eo, en, fo, fn = (int(x) for x in self.desc[h])
        return (self.vertices[vo:vo + vn], self.edges[eo:eo + en], self.faces[fo:fo + fn],
                self.planes[fo:fo + fn], self.centroids[h])

    @staticmethod
    def concat(hulls):
        """Pool from a list of (vertices, edges, faces, planes, centroid) tuples."""
        desc, cs, vs, es, fs, ps = [], [], [], [], [], []
        vo = eo = fo = 0
        for v, e, f, p, c in hulls:
            v = np.asarray(v, np.float32).reshape(-1, 3)
            e = np.asarray(e, np.uint32).reshape(-1, 4)
            f = np.asarray(f, np.uint32).reshape(-1)
            p = np.asarray(p, np.float32).reshape(-1, 4)
            desc.append([vo, len(v), eo, len(e), fo, len(f)])
            cs.append(np.asarray(c, np.float32).reshape(3))
            vs.append(v), es.append(e), fs.append(f), ps.append(p)
            vo += len(v)
            eo += len(e)
            fo += len(f)
        return HullPool(np.array(desc, np.uint32), np.array(cs, np.float32), np.concatenate(vs),
                        np.concatenate(es), np.concatenate(fs), np.concatenate(ps))

    @staticmethod
    def join(a, b):
        """Pool holding a's hulls, then b's (indices of b shift by a.count)."""
        shift = np.array([len(a.vertices), 0, len(a.edges), 0, len(a.faces), 0], np.uint32)
        return HullPool(np.concatenate([a.desc, b.desc + shift]), np.concatenate([a.centroids, b.centroids]),
                        np.concatenate([a.vertices, b.vertices]), np.concatenate([a.edges, b.edges]),
                        np.concatenate([a.faces, b.faces]), np.concatenate([a.planes, b.planes]))

    def shapes(self, radius=0.0):
        """One b3HullShape record per pool hull (shape index == hull index)."""
        s = np.zeros(self.count, SHAPE_DTYPE)
        s["type"] = HULL
        s["hull"] = np.arange(self.count, dtype=np.uint32)
        s["radius"] = radius
        return s


def ptr(a):
    """ctypes void pointer to a NumPy array's data (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)
