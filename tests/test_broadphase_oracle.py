"""The step before the narrowphase in the oracle: b3BroadPhase over b3DynamicTree restated
(oracle/b3oracle.c: node pool, surface-area insertion, removal, AABB query, FindPairs) and
the contact manager's AddPair / UpdateContacts on scripted worlds.

Pinned against the reference's own b3BroadPhase and b3World compiled from its sources
(oracle/_ref): proxy ids, fat AABBs, every FindPairs list in order, contact lists,
manifolds and caches; against brute force: each FindPairs list is exactly the sorted set
of (moved, other) proxy pairs whose fat AABBs overlap; and the rule the device broadphase
assigns proxy ids by (creates pop the group the last destroy at that count pushed) is
checked against the reference's ids."""
import numpy as np
import pytest

import oracle_lib
from util import (broadphase_script, run_broadphase_script, ScalarBroadPhase, kinematic_world,
                  run_kinematic_world, assert_manifolds_equal, assert_caches_equal, defined_hull_cache)

ref = oracle_lib.reference()
needs_ref = pytest.mark.skipif(ref is None, reason="reference not built (oracle/_ref)")


def oracle_bp():
    return oracle_lib.BroadPhase(oracle_lib.oracle().lib, "b3o_bp_")


def ref_bp():
    return oracle_lib.BroadPhase(ref.lib, "b3ref_bp_")


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_broadphase_oracle_equals_reference(seed):
    ops = broadphase_script(seed, n0=250, steps=15)
    a, b = oracle_bp(), ref_bp()
    ida, mva, fa = run_broadphase_script(ScalarBroadPhase(a), ops)
    idb, mvb, fb = run_broadphase_script(ScalarBroadPhase(b), ops)
    assert np.array_equal(ida, idb)
    assert len(mva) == len(mvb) and all(np.array_equal(x, y) for x, y in zip(mva, mvb))
    assert len(fa) == len(fb)
    for k, (x, y) in enumerate(zip(fa, fb)):
        assert np.array_equal(x, y), "find %d: %d vs %d pairs" % (k, len(x), len(y))
    assert sum(len(x) for x in fa) > 1000
    live = set(range(len(ida)))
    for op in ops:
        if op[0] == "destroy":
            live -= set(op[1].tolist())
    for h in live:
        assert np.array_equal(a.aabb(ida[h]).view(np.uint32), b.aabb(idb[h]).view(np.uint32))
    rng = np.random.default_rng(seed)
    hs = np.array(sorted(live))
    for _ in range(300):
        x, y = rng.choice(hs, 2, replace=False)
        assert a.test_overlap(ida[x], ida[y]) == b.test_overlap(idb[x], idb[y])


def test_find_pairs_is_the_sorted_moved_overlap_set():
    ops = broadphase_script(7, n0=200, steps=12)
    bp = oracle_bp()
    ids, moved, live = [], set(), set()
    checked = 0
    for op in ops:
        if op[0] == "create":
            for a in op[1]:
                h = len(ids)
                ids.append(bp.create(a, h))
                moved.add(h)
                live.add(h)
        elif op[0] == "destroy":
            for h in op[1]:
                bp.destroy(ids[h])
                live.discard(int(h))
                moved.discard(int(h))
        elif op[0] == "touch":
            for h in op[1]:
                bp.touch(ids[h])
                moved.add(int(h))
        elif op[0] == "move":
            for h, a, d in zip(op[1], op[2], op[3]):
                if bp.move(ids[h], a, d):
                    moved.add(int(h))
        else:
            hs = sorted(live)
            fat = {h: bp.aabb(ids[h]).astype(np.float64) for h in hs}
            want = set()
            for h in moved:
                for g in hs:
                    if g != h and np.all(fat[h][:3] <= fat[g][3:]) and np.all(fat[h][3:] >= fat[g][:3]):
                        want.add((min(ids[h], ids[g]), max(ids[h], ids[g])))
            got = bp.find_pairs()
            assert [tuple(p) for p in got[:, :2].tolist()] == sorted(want)
            assert all(ids[u] == p for u, p in zip(got[:, 2], got[:, 0]))
            checked += len(got)
            moved = set()
    assert checked > 1000


@needs_ref
def test_proxy_ids_follow_the_node_pool_groups():
    """A create takes the leaf id of the group the last destroy at the current proxy count
    pushed, or the next never-used node (two when the tree is not empty)."""
    rng = np.random.default_rng(11)
    bp = ref_bp()
    groups, fresh, count, live = [], 0, 0, []
    box = np.array([0, 0, 0, 1, 1, 1], np.float32)
    for step in range(3000):
        if live and (rng.random() < 0.45 or count > 300):
            i = int(rng.integers(0, len(live)))
            pid = live.pop(i)
            bp.destroy(pid)
            groups.append((pid, count >= 2))
            count -= 1
        else:
            if groups:
                pid, parent = groups.pop()
                assert parent == (count >= 1)
            else:
                pid = fresh
                fresh += 2 if count >= 1 else 1
            got = bp.create(box + rng.uniform(0, 10, 1).astype(np.float32), step)
            assert got == pid
            live.append(pid)
            count += 1
        if rng.random() < 0.3 and live:
            pid = live[int(rng.integers(0, len(live)))]
            bp.move(pid, box + rng.uniform(-20, 20, 1).astype(np.float32), np.ones(3, np.float32))


def _kworld_compare(w, a, b):
    def check(s, _w, stats):
        pa, ma, ca = a.contacts()
        pb, mb, cb = b.contacts()
        assert np.array_equal(pa, pb), "step %d: contact lists differ (%d vs %d)" % (s, len(pa), len(pb))
        assert_manifolds_equal(ma, mb, "step %d manifolds" % s)
        assert_caches_equal(defined_hull_cache(ca), defined_hull_cache(cb), "step %d caches" % s)
        ida, fa = a.proxies(np.arange(a.bodies))
        idb, fb = b.proxies(np.arange(b.bodies))
        assert np.array_equal(ida, idb) and np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        counts.append((len(pa), int((ma["pointCount"] > 0).sum())))
    counts = []
    return check, counts


@needs_ref
def test_kinematic_world_oracle_equals_reference():
    from bounce_amd import load_library
    w = kinematic_world(load_library(), 5, n_dyn=200, steps=40, extent=4.0, speed=3.0)
    a = oracle_lib.KWorld(oracle_lib.oracle().lib, "b3o_kworld_", w["pool"], w["shapes"])
    b = oracle_lib.KWorld(ref.lib, "b3ref_kworld_", w["pool"], w["shapes"])

    class Both:
        bodies = 0

        def add_bodies(self, *args):
            x, y = a.add_bodies(*args), b.add_bodies(*args)
            assert np.array_equal(x, y)
            self.bodies = a.bodies

        def find_new_contacts(self):
            x, y = a.find_new_contacts(), b.find_new_contacts()
            assert x == y
            created.append(x)

        def update_contacts(self):
            x, y = a.update_contacts(), b.update_contacts()
            assert np.array_equal(x, y), (x, y)
            destroyed.append(int(x[0]))
            return x

        def move_bodies(self, *args):
            a.move_bodies(*args)
            b.move_bodies(*args)

    created, destroyed = [], []
    check, counts = _kworld_compare(w, a, b)
    run_kinematic_world(Both(), w, check)
    # contacts are created and destroyed throughout, and many touch
    assert sum(created) > 300 and sum(destroyed) > 100
    assert sum(c[1] for c in counts) > 200


@needs_ref
def test_rain_world_oracle_equals_reference():
    """Hulls raining into a container (gen_rain_world, 6 x 6 columns of 8): the oracle's
    contact manager equals the reference's b3World at every step, and contacts come and go as
    the hulls fall past their neighbours and land."""
    from bounce_amd import load_library
    from bounce_amd.workloads import gen_rain_world
    from util import rain_script
    w = rain_script(gen_rain_world(load_library(), 3, n_hulls=288, columns=6, half_width=2.9, fall_height=3.0),
                    130)
    a = oracle_lib.KWorld(oracle_lib.oracle().lib, "b3o_kworld_", w["pool"], w["shapes"])
    b = oracle_lib.KWorld(ref.lib, "b3ref_kworld_", w["pool"], w["shapes"])
    created, destroyed = [], []

    class Both:
        bodies = 0

        def add_bodies(self, *args):
            assert np.array_equal(a.add_bodies(*args), b.add_bodies(*args))
            self.bodies = a.bodies

        def find_new_contacts(self):
            x, y = a.find_new_contacts(), b.find_new_contacts()
            assert x == y
            created.append(x)

        def update_contacts(self):
            x, y = a.update_contacts(), b.update_contacts()
            assert np.array_equal(x, y), (x, y)
            destroyed.append(int(x[0]))
            return x

        def move_bodies(self, *args):
            a.move_bodies(*args)
            b.move_bodies(*args)

    check, counts = _kworld_compare(w, a, b)
    run_kinematic_world(Both(), w, check)
    assert sum(created) > 500 and sum(destroyed) > 50
    assert counts[-1][1] > 200  # resting contacts with points at the end
