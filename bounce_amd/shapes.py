"""Host-side construction of b3Hull records (hull.h:37-64) from face lists."""
import numpy as np


def polyhedron(vertices, faces):
    """b3Hull record from vertices [V,3] and faces given as CCW (outward) vertex loops.

    Half-edges are numbered so that 2k and 2k+1 are twins (sat.cpp:123-128); each face's
    half-edges are linked through `next` in loop order; planes are unit outward normals
    through the first vertex; the centroid is the vertex mean.
    Returns (vertices, edges[HE,4], faces[F], planes[F,4], centroid)."""
    v = np.asarray(vertices, np.float64)
    uid = {}
    half = {}
    nxt = []
    for f, loop in enumerate(faces):
        ids = []
        for k in range(len(loop)):
            a, b = int(loop[k]), int(loop[(k + 1) % len(loop)])
            key = (min(a, b), max(a, b))
            if key not in uid:
                uid[key] = len(uid)
                h = 2 * uid[key]
            else:
                h = 2 * uid[key] + 1
            half[h] = (a, f)
            ids.append(h)
        nxt.append(ids)
    he = np.zeros((2 * len(uid), 4), np.uint32)
    for h, (a, f) in half.items():
        he[h, 0] = a
        he[h, 1] = h ^ 1
        he[h, 2] = f
    fedge = np.zeros(len(faces), np.uint32)
    for f, ids in enumerate(nxt):
        for k, h in enumerate(ids):
            he[h, 3] = ids[(k + 1) % len(ids)]
        fedge[f] = ids[0]
    planes = np.zeros((len(faces), 4), np.float64)
    for f, loop in enumerate(faces):
        p = v[list(loop)]
        n = np.zeros(3)
        for k in range(len(p)):  # Newell normal
            a, b = p[k], p[(k + 1) % len(p)]
            n += np.array([(a[1] - b[1]) * (a[2] + b[2]), (a[2] - b[2]) * (a[0] + b[0]), (a[0] - b[0]) * (a[1] + b[1])])
        n /= np.linalg.norm(n)
        planes[f, :3] = n
        planes[f, 3] = n @ p[0]
    return (v.astype(np.float32), he, fedge, planes.astype(np.float32), v.mean(0).astype(np.float32))


def box_hull(ex, ey, ez):
    """Box with half-extents (ex, ey, ez) centred at the origin, quad faces."""
    verts = np.array([[(1 if i & 1 else -1) * ex, (1 if i & 2 else -1) * ey, (1 if i & 4 else -1) * ez]
                      for i in range(8)], np.float32)
    faces = [
        [0, 4, 6, 2],  # -x
        [1, 3, 7, 5],  # +x
        [0, 1, 5, 4],  # -y
        [2, 6, 7, 3],  # +y
        [0, 2, 3, 1],  # -z
        [4, 5, 7, 6],  # +z
    ]
    return polyhedron(verts, faces)


def identity_xf(position=(0.0, 0.0, 0.0), rotation=None):
    """b3Transform as 12 floats: rotation columns then position."""
    x = np.zeros(12, np.float32)
    R = np.eye(3, dtype=np.float32) if rotation is None else np.asarray(rotation, np.float32)
    x[0:3], x[3:6], x[6:9] = R[:, 0], R[:, 1], R[:, 2]
    x[9:12] = position
    return x
