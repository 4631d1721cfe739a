"""CPU checks of the serving workload generator: each world's contact list is exactly the
set of body pairs whose fat AABBs (extended by b3_aabbExtension = 0.2 m, settings.h:54)
overlap, recomputed here by brute force over every body pair from the stored transforms
and the shapes' own vertices (independent of the generator's layout code); pairs are in
(lower body, higher body) order with typeA <= typeB (contact_manager.cpp:238-242); bodies
in piles have several contacts each."""
import numpy as np

from bounce_amd.types import HULL, SPHERE
from bounce_amd.workloads import gen_serving_worlds


def world_aabbs(pool, shapes, world, step=0):
    xf = world["xfs"][step].astype(np.float64)
    lo, hi = [], []
    for b, s in enumerate(world["body_shapes"]):
        R = np.stack([xf[b, 0:3], xf[b, 3:6], xf[b, 6:9]], 1)
        p = xf[b, 9:12]
        if shapes["type"][s] == SPHERE:
            c = R @ shapes["v"][s][:3].astype(np.float64) + p
            r = float(shapes["radius"][s])
            lo.append(c - r)
            hi.append(c + r)
        else:
            v = pool.hull(int(shapes["hull"][s]))[0].astype(np.float64) @ R.T + p
            lo.append(v.min(0))
            hi.append(v.max(0))
    return np.array(lo) - 0.2, np.array(hi) + 0.2


def test_serving_contacts_are_fat_aabb_overlaps(lib):
    pool, shapes, worlds = gen_serving_worlds(lib, 7, 12, bodies_min=10, bodies_max=400, steps=1)
    per_body = []
    for w in worlds:
        lo, hi = world_aabbs(pool, shapes, w)
        n = len(lo)
        i, j = np.triu_indices(n, 1)
        gap = np.maximum(lo[i] - hi[j], lo[j] - hi[i]).max(1)  # > 0: apart along some axis
        sure = set(zip(i[gap < -1e-4], j[gap < -1e-4]))
        maybe = set(zip(i[gap <= 1e-4], j[gap <= 1e-4]))
        got = set(map(tuple, np.sort(w["pairs"].astype(np.int64), 1)))
        assert sure <= got <= maybe
        assert len(got) == len(w["pairs"])
        t = shapes["type"][w["body_shapes"]]
        assert (t[w["pairs"][:, 0]] <= t[w["pairs"][:, 1]]).all()
        key = np.sort(w["pairs"].astype(np.int64), 1)
        assert (np.diff(key[:, 0] * n + key[:, 1]) > 0).all()  # (lower, higher) order
        per_body.append(len(got) / (n - 1))
        assert shapes["type"][w["body_shapes"][0]] == HULL  # body 0 is the ground box
    assert 1.5 < np.mean(per_body) < 4.0
