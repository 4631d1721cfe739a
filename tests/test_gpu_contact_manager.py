"""GPU parity of the step before the narrowphase as a whole: the device b3ContactManager
(b3amd_contact_manager_*) over the device broadphase steps scripted worlds in
b3World::Step's order (new shapes -> FindNewContacts, UpdateContacts, the move with
SynchronizeShapes, FindNewContacts; world.cpp:115-144, 308-336) and, after every
UpdateContacts, equals the oracle's restated b3ContactManager (pinned to the reference's own
b3World in tests/test_broadphase_oracle.py) and, where it is built here, the reference
itself: the contact list in m_contactList order, every manifold and cache, proxy ids and fat
AABBs.  Contacts are created and destroyed throughout (bodies fly through each other and
spawn in batches)."""
import numpy as np
import pytest

import bounce_amd
import oracle_lib
from bounce_amd.types import SWEEP_DTYPE
from util import (kinematic_world, run_kinematic_world, assert_manifolds_equal, assert_caches_equal,
                  defined_hull_cache)

pytestmark = pytest.mark.gpu


class DeviceWorld:
    """The kinematic-world interface over a Context, Broadphase and ContactManager; one shape
    per body, added in body order, so shape index == body index."""

    def __init__(self, ctx, w):
        self.ctx = ctx
        ctx.upload_hulls(w["pool"])
        ctx.upload_shapes(w["shapes"])
        n = sum(len(s[1]) for s in w["spawns"])
        self.n = n
        self.xf = np.zeros((n, 12), np.float32)
        self.sw = np.zeros(n, SWEEP_DTYPE)
        self.flags = np.zeros(n, np.uint32)
        self.d_xf = bounce_amd.DeviceBuffer(ctx, 48 * n)
        self.d_q = bounce_amd.DeviceBuffer(ctx, 32 * n)
        self.d_sw = bounce_amd.DeviceBuffer(ctx, 72 * n)
        self.d_ids = bounce_amd.DeviceBuffer(ctx, 4 * n)
        self.bp = ctx.broadphase()
        self.m = ctx.contact_manager(self.bp)
        self.bodies = 0
        self.proxy = np.zeros(n, np.uint32)

    def close(self):
        self.m.close()
        self.bp.close()
        for b in (self.d_xf, self.d_q, self.d_sw, self.d_ids):
            b.free()

    def _upload(self):
        self.d_xf.upload(self.xf)
        q = np.ascontiguousarray(np.stack([self.sw["orientation0"], self.sw["orientation"]], 1))
        self.d_q.upload(q)
        self.d_sw.upload(self.sw)

    def add_bodies(self, shape_idx, types, xfs, sweeps):
        b = np.arange(self.bodies, self.bodies + len(shape_idx))
        self.xf[b] = xfs
        self.sw[b] = sweeps
        self.flags[b] = np.asarray(types, np.uint32) | 4
        self._upload()
        self.m.set_bodies(self.flags[: b[-1] + 1])
        ids = self.m.add_shapes(shape_idx, b, self.d_xf)
        self.proxy[b] = ids
        self.bodies += len(b)
        return ids

    def find_new_contacts(self):
        return self.m.find_new_contacts()

    def update_contacts(self):
        self._upload()
        return self.m.update_contacts(self.d_xf, self.d_q)

    def move_bodies(self, bodies, xfs, sweeps):
        self.xf[bodies] = xfs
        self.sw[bodies] = sweeps
        self._upload()
        self.d_ids.upload(np.ascontiguousarray(bodies, np.uint32))
        self.m.synchronize(len(bodies), self.d_ids, self.d_xf, self.d_sw)

    def contacts(self):
        pairs, born, overlap, out, caches = self.m.read()
        assert np.array_equal(overlap, (out["pointCount"] > 0).astype(np.uint32))
        return pairs, out, caches

    def proxies(self, bodies):
        fat, _, _ = self.bp.read(0, self.bp.info()[1])
        return self.proxy[bodies], fat[self.proxy[bodies]]


def _run(ctx, w, others):
    dev = DeviceWorld(ctx, w)
    stats = {"created": 0, "destroyed": 0, "touching": 0, "steps": 0, "max_contacts": 0}

    class All:
        def add_bodies(self, *args):
            ids = dev.add_bodies(*args)
            for o in others:
                assert np.array_equal(ids, o.add_bodies(*args))

        def find_new_contacts(self):
            k = dev.find_new_contacts()
            for o in others:
                assert k == o.find_new_contacts()
            stats["created"] += k

        def update_contacts(self):
            st = dev.update_contacts()
            for o in others:
                assert np.array_equal(st, o.update_contacts()), st
            stats["destroyed"] += int(st[0])
            return st

        def move_bodies(self, *args):
            dev.move_bodies(*args)
            for o in others:
                o.move_bodies(*args)

    def check(s, _w, st):
        pa, ma, ca = dev.contacts()
        for o in others:
            pb, mb, cb = o.contacts()
            assert np.array_equal(pa, pb), "step %d: contact lists differ (%d vs %d)" % (s, len(pa), len(pb))
            assert_manifolds_equal(ma, mb, "step %d manifolds" % s)
            assert_caches_equal(defined_hull_cache(ca), defined_hull_cache(cb), "step %d caches" % s)
            ida, fa = dev.proxies(np.arange(dev.bodies))
            idb, fb = o.proxies(np.arange(dev.bodies))
            assert np.array_equal(ida, idb) and np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
        stats["touching"] += int((ma["pointCount"] > 0).sum())
        stats["steps"] += 1
        stats["max_contacts"] = max(stats["max_contacts"], len(pa))

    try:
        run_kinematic_world(All(), w, check)
    finally:
        dev.close()
    return stats


@pytest.fixture(scope="module")
def ctx():
    with bounce_amd.Context(0) as c:
        yield c


def test_gpu_contact_manager_equals_reference_small(ctx):
    print("kinematic world seed 5, 200 bodies, 40 steps", flush=True)
    w = kinematic_world(bounce_amd.load_library(), 5, n_dyn=200, steps=40, extent=4.0, speed=3.0)
    others = [oracle_lib.KWorld(oracle_lib.oracle().lib, "b3o_kworld_", w["pool"], w["shapes"])]
    ref = oracle_lib.reference()
    if ref is not None:
        others.append(oracle_lib.KWorld(ref.lib, "b3ref_kworld_", w["pool"], w["shapes"]))
    st = _run(ctx, w, others)
    assert st["created"] > 1000 and st["destroyed"] > 500 and st["touching"] > 2000


def test_gpu_contact_manager_equals_oracle_large(ctx):
    """2,000 bodies, batches of 100 spawned every 10 steps, 30 steps: the slot arrays and
    the sorted keys grow past their first capacity while contacts come and go."""
    print("kinematic world seed 9, 2000 bodies, 30 steps", flush=True)
    w = kinematic_world(bounce_amd.load_library(), 9, n_dyn=2000, steps=30, extent=8.0, speed=3.0,
                        spawn_count=100)
    others = [oracle_lib.KWorld(oracle_lib.oracle().lib, "b3o_kworld_", w["pool"], w["shapes"])]
    st = _run(ctx, w, others)
    assert st["created"] > 10000 and st["destroyed"] > 5000 and st["max_contacts"] > 4096


def test_gpu_contact_manager_body_types_and_errors(ctx):
    """Static and kinematic bodies never pair with each other, a contact whose bodies become
    non-dynamic is destroyed, one whose bodies both sleep is skipped, and bad inputs are
    refused."""
    from bounce_amd.shapes import box_hull
    from bounce_amd.types import HullPool, SHAPE_DTYPE, HULL
    pool = HullPool.concat([box_hull(0.5, 0.5, 0.5)])
    shapes = np.zeros(1, SHAPE_DTYPE)
    shapes["type"] = HULL
    ctx.upload_hulls(pool)
    ctx.upload_shapes(shapes)
    n = 4
    xf = np.zeros((n, 12), np.float32)
    xf[:, [0, 4, 8]] = 1.0
    xf[:, 9] = [0.0, 0.9, 1.8, 2.7]  # a row of touching boxes
    d_xf = bounce_amd.DeviceBuffer(ctx, 48 * n)
    d_xf.upload(xf)
    with ctx.broadphase() as bp, ctx.contact_manager(bp) as m:
        m.set_bodies([0 | 4, 1 | 4, 2 | 4, 2])  # static, kinematic, dynamic, dynamic asleep
        m.add_shapes([0, 0, 0, 0], [0, 1, 2, 3], d_xf)
        assert m.find_new_contacts() == 2  # (1,2) and (2,3); (0,1) is static-kinematic
        pairs, _, _, _, _ = m.read()
        assert sorted(map(tuple, pairs.tolist())) == [(1, 2), (2, 3)]
        # AddPair woke body 3 (contact_manager.cpp:116-121)
        assert m.bodies(4)[3] & 4
        st = m.update_contacts(d_xf)
        assert list(st) == [0, 2, 0]
        m.set_bodies([0 | 4, 1 | 4, 2, 2])  # both dynamic bodies asleep: (2,3) skipped; (1,2) has kinematic 1 awake
        st = m.update_contacts(d_xf)
        assert list(st) == [0, 1, 1]
        m.set_bodies([0 | 4, 1 | 4, 1 | 4, 2 | 4])  # body 2 becomes kinematic: (1,2) destroyed
        st = m.update_contacts(d_xf)
        assert list(st) == [1, 1, 0] and m.count() == 1
        with pytest.raises(bounce_amd.BounceError):
            m.set_bodies([0, 0])  # shapes on bodies 2 and 3
        with pytest.raises(bounce_amd.BounceError):
            m.add_shapes([5], [0], d_xf)  # shape beyond the table
        d_bad = bounce_amd.DeviceBuffer(ctx, 256)
        d_bad.upload(np.array([77], np.uint32))
        d_sw = bounce_amd.DeviceBuffer(ctx, 72 * n)
        d_sw.upload(np.zeros(n, SWEEP_DTYPE))
        m.synchronize(1, d_bad, d_xf, d_sw)
        with pytest.raises(bounce_amd.BounceError):
            m.find_new_contacts()
        d_bad.free(), d_sw.free()
    d_xf.free()


def test_gpu_rain_world_equals_reference(ctx):
    """Hulls raining into a container (gen_rain_world, 6 x 6 columns of 8 hulls of 8-64
    points, 130 steps): broadphase, contact manager and narrowphase on the device equal the
    reference's own b3World (and the oracle) after every step: the list order as contacts are
    created when hulls fall past and reach their neighbours and destroyed as they move apart,
    every manifold and cache carried from each contact's creation."""
    from bounce_amd.workloads import gen_rain_world
    from util import rain_script
    print("rain world seed 3, 288 hulls, 130 steps", flush=True)
    w = rain_script(gen_rain_world(bounce_amd.load_library(), 3, n_hulls=288, columns=6, half_width=2.9,
                                   fall_height=3.0), 130)
    others = [oracle_lib.KWorld(oracle_lib.oracle().lib, "b3o_kworld_", w["pool"], w["shapes"])]
    ref = oracle_lib.reference()
    if ref is not None:
        others.append(oracle_lib.KWorld(ref.lib, "b3ref_kworld_", w["pool"], w["shapes"]))
    st = _run(ctx, w, others)
    assert st["created"] > 8000 and st["destroyed"] > 4000 and st["touching"] > 15000
